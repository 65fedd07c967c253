// Level-set analysis on the GPU that reproduces the reference's BFS
// *discovery order* exactly (R/findlevel.h:7-83, R/utils_cuda.h:29-97,
// where it runs on a single GPU thread).
//
// Order rule (the sequential algorithm restated): level 0 holds the rows
// with no off-diagonal entry, ascending.  A row v joins level k+1 when the
// last of its dependencies is visited; in the sequential walk that is its
// dependency u* with the largest level-order position, and v lands at the
// position sorted by (position of u*, offset of v inside column u*).
//
// Parallel form, one level per round inside one cooperative launch:
//   A  every frontier node u (position p) decrements rem[v] and raises
//      par[v] to p for every off-diagonal v in its column (atomics);
//   B  node u counts its children: v with rem[v] == 0 and par[v] == p;
//   S  exclusive scan of the counts in frontier order (grid-wide);
//   C  node u writes its children, in column order, at e + offset.
// A grid barrier separates the phases.
#include <hip/hip_cooperative_groups.h>

#include "rb_internal.h"

namespace cg = cooperative_groups;

namespace rb {

namespace {

constexpr int kBfsThreads = 256;

struct BfsCtl {
    int s, e;     // current frontier [s, e) of levelItem
    int nlevel;   // levels emitted so far
    int stalled;  // 1 when a round adds nothing while rows remain
};

// Exclusive scan of one value per thread across a 256-thread block.
// Returns the exclusive prefix; *total receives the block sum.
__device__ inline int block_excl_scan(int v, int *lds, int *total) {
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int w = 0; w < kBfsThreads / 64; w++) {
            int t = lds[w];
            lds[w] = run;
            run += t;
        }
        lds[kBfsThreads / 64] = run;
    }
    __syncthreads();
    int res = x - v + lds[wid];
    *total = lds[kBfsThreads / 64];
    __syncthreads();
    return res;
}

__device__ inline int ld_relaxed(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// workgroup scope: while one workgroup walks narrow levels it is the only
// reader and writer of rem / par / levelItem (every other workgroup waits
// at the grid barrier, whose device-scope release / acquire publishes the
// walk's results and makes earlier writes visible to it)
__device__ inline int ld_wg(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Grid-wide exclusive scan of cnt[0..len) in place (contiguous chunk per
// block).  Returns the grand total to every thread.  Uses two grid syncs.
__device__ int grid_excl_scan(cg::grid_group &g, int *cnt, int len,
                              int *blocksum, int *lds) {
    const int G = gridDim.x;
    const int chunk = (len + G - 1) / G;
    const int b0 = blockIdx.x * chunk;
    const int b1 = min(len, b0 + chunk);
    int local = 0;
    for (int i = b0 + threadIdx.x; i < b1; i += blockDim.x) local += cnt[i];
    int tot;
    (void)block_excl_scan(local, lds, &tot);
    if (threadIdx.x == 0) blocksum[blockIdx.x] = tot;
    g.sync();
    // every block scans the block sums redundantly (G is small)
    int before = 0, grand = 0;
    {
        int run = 0;
        for (int base = 0; base < G; base += blockDim.x) {
            int i = base + threadIdx.x;
            int v = i < G ? ld_relaxed(&blocksum[i]) : 0;
            int t;
            int ex = block_excl_scan(v, lds, &t);
            if (i == blockIdx.x) lds[8] = run + ex;
            run += t;
            __syncthreads();
        }
        before = lds[8];
        grand = run;
        __syncthreads();
    }
    int run = before;
    for (int base = b0; base < b1; base += blockDim.x) {
        int i = base + threadIdx.x;
        int v = i < b1 ? cnt[i] : 0;
        int t;
        int ex = block_excl_scan(v, lds, &t);
        if (i < b1) cnt[i] = run + ex;
        run += t;
    }
    g.sync();
    return grand;
}

constexpr int kNarrow = 4096;  // frontier size handled by one workgroup
constexpr int kIlp = 6;        // frontier nodes per thread, cached levels
constexpr int kE = 4;          // column entries per node kept in registers

// Exclusive scan of cl[0..len) in place (one workgroup); returns the sum.
__device__ int narrow_scan(int *cl, int len, int *lds) {
    const int tid = threadIdx.x;
    const int per = (len + blockDim.x - 1) / blockDim.x;
    const int lo = min(len, tid * per), hi = min(len, lo + per);
    int local = 0;
    for (int i = lo; i < hi; i++) local += cl[i];
    int total;
    int run = block_excl_scan(local, lds, &total);
    for (int i = lo; i < hi; i++) {
        const int t = cl[i];
        cl[i] = run;
        run += t;
    }
    __syncthreads();
    return total;
}

// The frontier of the next level, kept in LDS by a cached level when that
// level is cached too: its nodes and their column ranges [a, b).
struct Front {
    int u[kIlp * kBfsThreads];
    int a[kIlp * kBfsThreads];
    int b[kIlp * kBfsThreads];
};

// One narrow level [s, e) of at most kIlp * kBfsThreads nodes.  Each thread
// keeps its nodes, their entry ranges and the first kE entries of their
// columns in registers from phase A to C, and phase B remembers which of
// those entries are children, so B reads only rem / par and C only writes:
// a level costs about four dependent memory round trips instead of a dozen.
// Entries past the first kE of a column are re-read (the generic rule).
__device__ int level_cached(const int *__restrict__ colptr,
                            const int *__restrict__ rowidx, int *rem, int *par,
                            int *levelItem, int s, int e, int *cl, int *lds,
                            Front &fr, bool &held) {
    const int tid = threadIdx.x, len = e - s;
    const bool in = held;  // this level's nodes and column ranges are in fr
    int u[kIlp], q0[kIlp], q1[kIlp], vc[kIlp][kE];
    unsigned kid[kIlp];
    if (in) {
#pragma unroll
        for (int k = 0; k < kIlp; k++) {
            const int i = tid + k * kBfsThreads;
            u[k] = i < len ? fr.u[i] : -1;
            q0[k] = i < len ? fr.a[i] : 0;
            q1[k] = i < len ? fr.b[i] : 0;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kIlp; k++) {
            const int i = tid + k * kBfsThreads;
            u[k] = i < len ? ld_wg(&levelItem[s + i]) : -1;
        }
#pragma unroll
        for (int k = 0; k < kIlp; k++) {
            q0[k] = u[k] >= 0 ? colptr[u[k]] : 0;
            q1[k] = u[k] >= 0 ? colptr[u[k] + 1] : 0;
        }
    }
#pragma unroll
    for (int k = 0; k < kIlp; k++)
#pragma unroll
        for (int j = 0; j < kE; j++)
            vc[k][j] = q0[k] + j < q1[k] ? rowidx[q0[k] + j] : -1;
    // A: decrement dependants, raise their parent position
#pragma unroll
    for (int k = 0; k < kIlp; k++) {
        const int p = s + tid + k * kBfsThreads;
#pragma unroll
        for (int j = 0; j < kE; j++) {
            const int v = vc[k][j];
            if (v < 0 || v == u[k]) continue;
            __hip_atomic_fetch_add(&rem[v], -1, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_fetch_max(&par[v], p, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        for (int q = q0[k] + kE; q < q1[k]; q++) {
            const int v = rowidx[q];
            if (v == u[k]) continue;
            __hip_atomic_fetch_add(&rem[v], -1, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_fetch_max(&par[v], p, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    // B: children (rem 0, parent p) per node.  The cached entries' rem and
    // par are read together (a non-candidate reads row 0, in range, and is
    // ignored), and with them the entries' column ranges, which a child
    // takes into the next level's frontier in LDS.  A narrow level cost
    // about five dependent trips to memory (frontier, column pointers,
    // entries, atomics, rem / par) at ~2 us each on the 1500^2 Laplacian;
    // with the frontier in LDS three are left.  Reading rem / par entry by
    // entry behind branches cost another ~0.5 us.
    int rv[kIlp][kE], pv[kIlp][kE], ca[kIlp][kE], cb[kIlp][kE];
#pragma unroll
    for (int k = 0; k < kIlp; k++)
#pragma unroll
        for (int j = 0; j < kE; j++) {
            const int v = vc[k][j];
            const int va = v >= 0 && v != u[k] ? v : 0;
            rv[k][j] = ld_wg(&rem[va]);
            pv[k][j] = ld_wg(&par[va]);
            ca[k][j] = colptr[va];
            cb[k][j] = colptr[va + 1];
        }
#pragma unroll
    for (int k = 0; k < kIlp; k++) {
        const int i = tid + k * kBfsThreads;
        const int p = s + i;
        int c = 0;
        unsigned bits = 0;
#pragma unroll
        for (int j = 0; j < kE; j++) {
            const int v = vc[k][j];
            const bool cand = v >= 0 && v != u[k];
            if (cand && rv[k][j] == 0 && pv[k][j] == p) {
                c++;
                bits |= 1u << j;
            }
        }
        for (int q = q0[k] + kE; q < q1[k]; q++) {
            const int v = rowidx[q];
            if (v != u[k] && ld_wg(&rem[v]) == 0 && ld_wg(&par[v]) == p) c++;
        }
        kid[k] = bits;
        if (i < len) cl[i] = c;
    }
    __syncthreads();
    const int total = narrow_scan(cl, len, lds);
    // C: children in column order at the node's offset; when the next
    // level is cached too, its frontier also goes to fr (every read of fr
    // was in A, before this level's barriers)
    const bool out = total <= kIlp * kBfsThreads;
#pragma unroll
    for (int k = 0; k < kIlp; k++) {
        const int i = tid + k * kBfsThreads;
        if (i >= len) continue;
        const int p = s + i;
        int o = e + cl[i];
#pragma unroll
        for (int j = 0; j < kE; j++)
            if ((kid[k] >> j) & 1u) {
                if (out) {
                    fr.u[o - e] = vc[k][j];
                    fr.a[o - e] = ca[k][j];
                    fr.b[o - e] = cb[k][j];
                }
                __hip_atomic_store(&levelItem[o++], vc[k][j], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        for (int q = q0[k] + kE; q < q1[k]; q++) {
            const int v = rowidx[q];
            if (v != u[k] && ld_wg(&rem[v]) == 0 && ld_wg(&par[v]) == p) {
                if (out) {
                    fr.u[o - e] = v;
                    fr.a[o - e] = colptr[v];
                    fr.b[o - e] = colptr[v + 1];
                }
                __hip_atomic_store(&levelItem[o++], v, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
    }
    held = out;
    return total;
}

// One narrow level of any width up to kNarrow: the same phases, each
// re-reading the frontier's columns.
__device__ int level_generic(const int *__restrict__ colptr,
                             const int *__restrict__ rowidx, int *rem, int *par,
                             int *levelItem, int s, int e, int *cl, int *lds) {
    const int tid = threadIdx.x, len = e - s;
    for (int p = s + tid; p < e; p += blockDim.x) {
        const int u = ld_wg(&levelItem[p]);
        for (int q = colptr[u]; q < colptr[u + 1]; q++) {
            const int v = rowidx[q];
            if (v == u) continue;
            __hip_atomic_fetch_add(&rem[v], -1, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_fetch_max(&par[v], p, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    for (int i = tid; i < len; i += blockDim.x) {
        const int p = s + i;
        const int u = ld_wg(&levelItem[p]);
        int c = 0;
        for (int q = colptr[u]; q < colptr[u + 1]; q++) {
            const int v = rowidx[q];
            if (v != u && ld_wg(&rem[v]) == 0 && ld_wg(&par[v]) == p) c++;
        }
        cl[i] = c;
    }
    __syncthreads();
    const int total = narrow_scan(cl, len, lds);
    for (int i = tid; i < len; i += blockDim.x) {
        const int p = s + i;
        const int u = ld_wg(&levelItem[p]);
        int o = e + cl[i];
        for (int q = colptr[u]; q < colptr[u + 1]; q++) {
            const int v = rowidx[q];
            if (v != u && ld_wg(&rem[v]) == 0 && ld_wg(&par[v]) == p)
                __hip_atomic_store(&levelItem[o++], v, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
    return total;
}

// Consecutive narrow levels inside workgroup 0: the same A / B / scan / C
// phases, separated by workgroup barriers.  Frontier counts live in LDS.
// rem / par / levelItem are read, updated and written at workgroup scope
// (see ld_wg).
__device__ void bfs_narrow(const int *__restrict__ colptr,
                           const int *__restrict__ rowidx, int m, int *rem,
                           int *par, int *levelItem, int *levelPtr,
                           BfsCtl *ctl, int *lds) {
    __shared__ int cl[kNarrow];
    __shared__ Front fr;
    const int tid = threadIdx.x;
    int s = ld_relaxed(&ctl->s), e = ld_relaxed(&ctl->e);
    int nl = ld_relaxed(&ctl->nlevel);
    int stalled = 0;
    bool held = false;  // the level's frontier is in fr
    while (e < m && e - s <= kNarrow) {
        int total;
        if (e - s <= kIlp * kBfsThreads) {
            total = level_cached(colptr, rowidx, rem, par, levelItem, s, e, cl, lds,
                                 fr, held);
        } else {
            total = level_generic(colptr, rowidx, rem, par, levelItem, s, e, cl, lds);
            held = false;
        }
        if (tid == 0) levelPtr[nl + 1] = e + total;
        nl++;
        s = e;
        e += total;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __syncthreads();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        if (total == 0) {
            stalled = 1;
            break;
        }
    }
    if (tid == 0) {
        ctl->s = s;
        ctl->e = e;
        ctl->nlevel = nl;
        if (stalled) ctl->stalled = 1;
    }
}

__global__ void __launch_bounds__(kBfsThreads)
k_bfs(const int *__restrict__ colptr, const int *__restrict__ rowidx, int m,
      int *rem, int *par, int *cnt, int *levelItem, int *levelPtr,
      int *blocksum, BfsCtl *ctl) {
    cg::grid_group g = cg::this_grid();
    __shared__ int lds[16];
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int gsz = gridDim.x * blockDim.x;

    // ---- level 0: rows without off-diagonal entries, ascending ----
    {
        const int G = gridDim.x;
        const int chunk = (m + G - 1) / G;
        const int b0 = blockIdx.x * chunk;
        const int b1 = min(m, b0 + chunk);
        for (int i = b0 + threadIdx.x; i < b1; i += blockDim.x)
            cnt[i] = rem[i] == 0 ? 1 : 0;
        g.sync();
        int total = grid_excl_scan(g, cnt, m, blocksum, lds);
        for (int i = b0 + threadIdx.x; i < b1; i += blockDim.x)
            if (rem[i] == 0) levelItem[cnt[i]] = i;
        if (gtid == 0) {
            levelPtr[0] = 0;
            levelPtr[1] = total;
            ctl->s = 0;
            ctl->e = total;
            ctl->nlevel = 1;
            ctl->stalled = (total == 0 && m > 0) ? 1 : 0;
        }
        g.sync();
    }

    while (true) {
        const int s = ld_relaxed(&ctl->s);
        const int e = ld_relaxed(&ctl->e);
        if (e >= m || ld_relaxed(&ctl->stalled)) break;
        const int len = e - s;
        if (len <= kNarrow) {
            // narrow frontier: workgroup 0 walks consecutive narrow levels
            // with workgroup barriers only; the rest wait at the grid sync
            if (blockIdx.x == 0)
                bfs_narrow(colptr, rowidx, m, rem, par, levelItem, levelPtr,
                           ctl, lds);
            g.sync();
            continue;
        }
        // A: decrement dependants, raise their parent position
        for (int p = s + gtid; p < e; p += gsz) {
            const int u = levelItem[p];
            for (int q = colptr[u]; q < colptr[u + 1]; q++) {
                const int v = rowidx[q];
                if (v == u) continue;
                atomicSub(&rem[v], 1);
                atomicMax(&par[v], p);
            }
        }
        g.sync();
        // B: count the children each frontier node discovers
        {
            const int G = gridDim.x;
            const int chunk = (len + G - 1) / G;
            const int b0 = blockIdx.x * chunk;
            const int b1 = min(len, b0 + chunk);
            for (int i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
                const int p = s + i;
                const int u = levelItem[p];
                int c = 0;
                for (int q = colptr[u]; q < colptr[u + 1]; q++) {
                    const int v = rowidx[q];
                    if (v != u && ld_relaxed(&rem[v]) == 0 &&
                        ld_relaxed(&par[v]) == p)
                        c++;
                }
                cnt[i] = c;
            }
        }
        g.sync();
        const int total = grid_excl_scan(g, cnt, len, blocksum, lds);
        // C: emit children in column order
        for (int i = gtid; i < len; i += gsz) {
            const int p = s + i;
            const int u = levelItem[p];
            int o = e + cnt[i];
            for (int q = colptr[u]; q < colptr[u + 1]; q++) {
                const int v = rowidx[q];
                if (v != u && ld_relaxed(&rem[v]) == 0 &&
                    ld_relaxed(&par[v]) == p)
                    levelItem[o++] = v;
            }
        }
        if (gtid == 0) {
            const int nl = ctl->nlevel;
            levelPtr[nl + 1] = e + total;
            ctl->nlevel = nl + 1;
            ctl->s = e;
            ctl->e = e + total;
            if (total == 0) ctl->stalled = 1;
        }
        g.sync();
    }
}

// One level [s, e) of any width inside one workgroup: the A / B / scan /
// C phases of the narrow levels, with the counts in device memory (cnt,
// indexed by frontier position) and a scan in chunks of the workgroup's
// width.  A thread handles the same positions in B and C, so it reads back
// only its own counts.
__device__ int level_wide(const int *__restrict__ colptr,
                          const int *__restrict__ rowidx, int *rem, int *par,
                          int *levelItem, int s, int e, int *cnt, int *lds) {
    const int tid = threadIdx.x;
    for (int p = s + tid; p < e; p += blockDim.x) {
        const int u = ld_wg(&levelItem[p]);
        for (int q = colptr[u]; q < colptr[u + 1]; q++) {
            const int v = rowidx[q];
            if (v == u) continue;
            __hip_atomic_fetch_add(&rem[v], -1, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_fetch_max(&par[v], p, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    int carry = 0;
    for (int base = s; base < e; base += blockDim.x) {  // uniform trip count
        const int p = base + tid;
        int c = 0;
        if (p < e) {
            const int u = ld_wg(&levelItem[p]);
            for (int q = colptr[u]; q < colptr[u + 1]; q++) {
                const int v = rowidx[q];
                if (v != u && ld_wg(&rem[v]) == 0 && ld_wg(&par[v]) == p) c++;
            }
        }
        int tot = 0;
        const int off = block_excl_scan(c, lds, &tot);
        if (p < e) cnt[p] = carry + off;
        carry += tot;
    }
    for (int p = s + tid; p < e; p += blockDim.x) {
        const int u = ld_wg(&levelItem[p]);
        int o = e + cnt[p];
        for (int q = colptr[u]; q < colptr[u + 1]; q++) {
            const int v = rowidx[q];
            if (v != u && ld_wg(&rem[v]) == 0 && ld_wg(&par[v]) == p)
                __hip_atomic_store(&levelItem[o++], v, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
    return carry;
}

// The level analysis of every triangular block at once, one workgroup per
// block (blockIdx.x = t): the same discovery order as the whole-matrix
// walk (R/findlevel.h:7-83), restricted to the block's own entries.  The
// CSC holds the block-diagonal matrix (every row's in-block entries; row
// ids are permuted ids, block t owns [ups[t], ups[t] + bms[t])), so the
// blocks share no row and each workgroup reads and writes only its own
// rows: no grid barrier, only workgroup barriers.  Writes the block's rows
// in discovery order to item[ups[t] ..] (permuted ids), its level starts
// (block-local, level 0 at 0) to lvptr[lvbase[t] ..] and its level count
// to nlv[t]; a block that cannot finish (a row never released) sets
// *stalled.
__global__ void __launch_bounds__(kBfsThreads)
k_bfs_blocks(const int *__restrict__ colptr, const int *__restrict__ rowidx,
             const int *__restrict__ rowcnt, const int *__restrict__ ups,
             const int *__restrict__ bms, const int *__restrict__ lvbase,
             int *item, int *lvptr, int *nlv, int *rem, int *par, int *cnt,
             int *stalled) {
    __shared__ int lds[16];
    __shared__ int cl[kNarrow];
    __shared__ Front fr;
    const int t = blockIdx.x, tid = threadIdx.x;
    const int up = ups[t], bm = bms[t];
    int *lp = lvptr + lvbase[t];
    if (bm == 0) {
        if (tid == 0) {
            lp[0] = 0;
            nlv[t] = 0;
        }
        return;
    }
    int *it = item + up;
    for (int i = tid; i < bm; i += blockDim.x) {
        rem[up + i] = rowcnt[up + i] - 1;
        par[up + i] = -1;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    // level 0: the rows without an in-block dependency, ascending
    int e = 0;
    for (int base = 0; base < bm; base += blockDim.x) {  // uniform trip count
        const int i = base + tid;
        const int f = i < bm && ld_wg(&rem[up + i]) == 0 ? 1 : 0;
        int tot = 0;
        const int off = block_excl_scan(f, lds, &tot);
        if (f)
            __hip_atomic_store(&it[e + off], up + i, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
        e += tot;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    if (tid == 0) {
        lp[0] = 0;
        lp[1] = e;
    }
    int s = 0, nl = 1;
    bool bad = e == 0;
    bool held = false;  // the level's frontier is in fr
    while (e < bm && !bad) {
        const int len = e - s;
        int total;
        if (len <= kIlp * kBfsThreads) {
            total = level_cached(colptr, rowidx, rem, par, it, s, e, cl, lds, fr, held);
        } else {
            total = len <= kNarrow
                        ? level_generic(colptr, rowidx, rem, par, it, s, e, cl, lds)
                        : level_wide(colptr, rowidx, rem, par, it, s, e, cnt + up, lds);
            held = false;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __syncthreads();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        if (total == 0) {
            bad = true;
            break;
        }
        if (tid == 0) lp[nl + 1] = e + total;
        nl++;
        s = e;
        e += total;
    }
    if (tid == 0) {
        nlv[t] = nl;
        if (bad) atomicOr(stalled, 1);
    }
}

__global__ void k_bfs_init(const int *__restrict__ rowcnt, int m, int *rem,
                           int *par) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) {
        rem[i] = rowcnt[i] - 1;
        par[i] = -1;
    }
}

}  // namespace

rb_status bfs_blocks(const int *d_colptr, const int *d_rowidx,
                     const int *d_rowcnt, int m, int nblocks, const int *d_ups,
                     const int *d_bms, const int *d_lvbase, int *d_item,
                     int *d_lvptr, int *d_nlv, hipStream_t s) {
    if (m == 0 || nblocks == 0) return RB_OK;
    Arena tmp(12 * (size_t)m + (1 << 20));  // rem, par, cnt and counters
    int *rem, *par, *cnt, *stalled;
    RB_TRY(tmp.alloc(&rem, m));
    RB_TRY(tmp.alloc(&par, m));
    RB_TRY(tmp.alloc(&cnt, m));
    RB_TRY(tmp.alloc(&stalled, 1));
    RB_HIP(hipMemsetAsync(stalled, 0, sizeof(int), s));
    k_bfs_blocks<<<nblocks, kBfsThreads, 0, s>>>(d_colptr, d_rowidx, d_rowcnt,
                                                 d_ups, d_bms, d_lvbase, d_item,
                                                 d_lvptr, d_nlv, rem, par, cnt,
                                                 stalled);
    RB_HIP(hipGetLastError());
    int h = 0;
    hipError_t e = hipMemcpyAsync(&h, stalled, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    tmp.release();
    if (e != hipSuccess) return status_of_hip(e, "k_bfs_blocks");
    if (h) {
        set_error(RB_ERR_MISSING_DIAG,
                  "block level analysis stalled: a row is never released "
                  "(missing diagonal or cyclic dependency)");
        return RB_ERR_MISSING_DIAG;
    }
    return RB_OK;
}

rb_status bfs_levels(const int *d_colptr, const int *d_rowidx,
                     const int *d_rowcnt, int m, int *d_levelItem,
                     int *d_levelPtr, int *nlevel, hipStream_t s) {
    *nlevel = 0;
    if (m == 0) {
        int z = 0;
        RB_HIP(hipMemcpyAsync(d_levelPtr, &z, sizeof(int),
                              hipMemcpyHostToDevice, s));
        RB_HIP(hipStreamSynchronize(s));
        return RB_OK;
    }
    int dev = 0;
    RB_HIP(hipGetDevice(&dev));
    int ncu = 0;
    RB_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                                 dev));
    int per_cu = 0;
    RB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_bfs,
                                                        kBfsThreads, 0));
    if (per_cu < 1) per_cu = 1;
    int grid = ncu * (per_cu < 2 ? per_cu : 2);
    int need = (m + kBfsThreads - 1) / kBfsThreads;
    if (grid > need) grid = need;
    if (grid < 1) grid = 1;

    Arena tmp(12 * (size_t)m + (1 << 20));  // rem, par, cnt and counters
    int *rem, *par, *cnt, *blocksum;
    BfsCtl *ctl;
    RB_TRY(tmp.alloc(&rem, m));
    RB_TRY(tmp.alloc(&par, m));
    RB_TRY(tmp.alloc(&cnt, m));
    RB_TRY(tmp.alloc(&blocksum, grid));
    RB_TRY(tmp.alloc(&ctl, 1));
    k_bfs_init<<<(m + 255) / 256, 256, 0, s>>>(d_rowcnt, m, rem, par);
    RB_HIP(hipGetLastError());
    void *args[] = {(void *)&d_colptr, (void *)&d_rowidx, (void *)&m,
                    (void *)&rem,      (void *)&par,      (void *)&cnt,
                    (void *)&d_levelItem, (void *)&d_levelPtr,
                    (void *)&blocksum, (void *)&ctl};
    hipError_t e = hipLaunchCooperativeKernel((const void *)k_bfs, dim3(grid),
                                              dim3(kBfsThreads), args, 0, s);
    if (e != hipSuccess) {
        tmp.release();
        return status_of_hip(e, "hipLaunchCooperativeKernel(k_bfs)");
    }
    BfsCtl h;
    e = hipMemcpyAsync(&h, ctl, sizeof(h), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    tmp.release();
    if (e != hipSuccess) return status_of_hip(e, "k_bfs");
    if (h.stalled) {
        set_error(RB_ERR_MISSING_DIAG,
                  "level analysis stalled: a row is never released "
                  "(missing diagonal or cyclic dependency)");
        return RB_ERR_MISSING_DIAG;
    }
    *nlevel = h.nlevel;
    return RB_OK;
}

}  // namespace rb
