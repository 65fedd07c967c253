// Plan construction on the device: validation, level analysis, level-order
// permutation, recursive partition, block extraction into the slot storage,
// selector statistics and the local level analysis of triangular blocks.
//
// Reference counterparts: levelset_reordering_colrow_csc_cuda
// (R/utils_cuda.h:706-779), mat_preprocessing_cuda (R/utils_cuda.h:241-482),
// the per-block loop of recblocking_solver_cuda
// (R/recblocking_solver_cuda.h:133-581 forward, :671-1121 backward), and
// the statistics kernels cal_longrow / find_level_nnz (R/utils_cuda.h:
// 781-830).
//
// Storage ("slot space"): every block owns blk_m consecutive slots, one per
// block-local row, in the reference's block order (tri0, sq1, tri1, ...).
// A slot holds that row's off-diagonal entries inside the block, as a CSR
// segment: rowptr over all slots, block-local column ids, values.  The
// diagonal of every row is kept apart as 1/diag (invd, permuted order).
#include <chrono>
#include <cstring>

#include "rb_internal.h"

namespace rb {

namespace {

struct Geom {
    int m, lv, step, subst;
};

// Block index of entry (r, c) of the permuted matrix (closed form of the
// bisection in R/utils_cuda.h:283-478): in coordinates mirrored for the
// backward case, leaf tr = row / step and tc = col / step (the last leaf
// absorbs the remainder); equal leaves give triangular block 2*tr,
// otherwise the square block of split point t = tr with the bits below
// the highest differing bit cleared, block index 2*t - 1.
__host__ __device__ inline int block_of(const Geom &g, int r, int c) {
    const int T = 1 << g.lv;
    const int rr = g.subst ? g.m - 1 - r : r;
    const int cc = g.subst ? g.m - 1 - c : c;
    int tr, tc;
    if (g.step == 0) {
        tr = tc = T - 1;
    } else {
        tr = rr / g.step;
        tc = cc / g.step;
        if (tr > T - 1) tr = T - 1;
        if (tc > T - 1) tc = T - 1;
    }
    if (tr == tc) return 2 * tr;
    const unsigned x = (unsigned)(tr ^ tc);
#ifdef __HIP_DEVICE_COMPILE__
    const int h = 31 - __clz((int)x);
#else
    const int h = 31 - __builtin_clz(x);
#endif
    const int t = (tr >> h) << h;
    return 2 * t - 1;
}

struct DevBlock {
    int up, left, slot_base, kind;
};

constexpr int kT = 256;

inline int nblk(long n, int t = kT) { return (int)((n + t - 1) / t); }

// ---------------------------------------------------------------- checks
// Bit flags: 1 bad structure, 2 missing diagonal, 4 zero diagonal.
__global__ void k_validate(const int *__restrict__ colptr,
                           const int *__restrict__ rowidx,
                           const double *__restrict__ val, int n, int nnz,
                           int subst, int *err) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int s = colptr[j], e = colptr[j + 1];
    int f = 0;
    if (s < 0 || e < s || e > nnz) {
        atomicOr(err, 1);
        return;
    }
    int prev = -1;
    for (int q = s; q < e; q++) {
        const int r = rowidx[q];
        if (r < 0 || r >= n || r <= prev) f |= 1;
        if (subst == 0 ? r < j : r > j) f |= 1;
        prev = r;
    }
    if (e == s) {
        f |= 2;
    } else {
        const int dq = subst == 0 ? s : e - 1;
        if (rowidx[dq] != j) f |= 2;
        else if (val[dq] == 0.0) f |= 4;
    }
    if (f) atomicOr(err, f);
}

__global__ void k_histogram(const int *__restrict__ idx, int nnz, int *cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nnz) atomicAdd(&cnt[idx[i]], 1);
}

// final permutation lf (both directions: new column i = old column lf[i])
// and its inverse newpos
__global__ void k_perm(const int *__restrict__ bfs_item, int m, int subst,
                       int *lf, int *newpos) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int old = subst == 0 ? bfs_item[i] : bfs_item[m - 1 - i];
    lf[i] = old;
    newpos[old] = i;
}

__global__ void k_pcolcnt(const int *__restrict__ colptr,
                          const int *__restrict__ lf, int m, int *cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) cnt[i] = colptr[lf[i] + 1] - colptr[lf[i]];
}

// wave per column: copy entries in order, mapping row ids
__global__ void k_pcolcopy(const int *__restrict__ colptr,
                           const int *__restrict__ rowidx,
                           const double *__restrict__ val,
                           const int *__restrict__ lf,
                           const int *__restrict__ newpos,
                           const int *__restrict__ pcolptr, int m,
                           int *prowidx, double *pval, int *ecol) {
    const int w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (w >= m) return;
    const int src = colptr[lf[w]];
    const int len = colptr[lf[w] + 1] - src;
    const int dst = pcolptr[w];
    for (int k = lane; k < len; k += 64) {
        prowidx[dst + k] = newpos[rowidx[src + k]];
        pval[dst + k] = val[src + k];
        ecol[dst + k] = w;
    }
}

// column ids of the level-ordered CSR mapped back to original ids
// The CSR of the original matrix, from the permuted one (row r of which is
// original row lf[r]): row lengths into original order, then each row's
// entries with original column ids, in the permuted row's entry order.
__global__ void k_orig_len(const int *__restrict__ prptr, const int *__restrict__ lf,
                           int m, int *olen) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < m) olen[lf[r]] = prptr[r + 1] - prptr[r];
}
__global__ void k_orig_rows(const int *__restrict__ prptr, const int *__restrict__ ccol,
                            const double *__restrict__ cval, const int *__restrict__ lf,
                            int m, const int *__restrict__ optr, int *ocol,
                            double *oval) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    int d = optr[lf[r]];
    for (int k = prptr[r]; k < prptr[r + 1]; k++, d++) {
        ocol[d] = lf[ccol[k]];
        oval[d] = cval[k];
    }
}

__global__ void k_f64_to_f32(const double *__restrict__ a, long n, float *o) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o[i] = (float)a[i];
}

__global__ void k_iota(int *a, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

__global__ void k_gather_csr(const int *__restrict__ order,
                             const int *__restrict__ ecol,
                             const double *__restrict__ pval, int nnz,
                             int *ccol, double *cval) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nnz) {
        const int q = order[k];
        ccol[k] = ecol[q];
        cval[k] = pval[q];
    }
}

// Count pass over permuted rows: runs of equal block index along a sorted
// row become one slot count; the diagonal becomes invd.
__global__ void k_slot_count(Geom g, const DevBlock *__restrict__ blk,
                             const int *__restrict__ bm,
                             const int *__restrict__ rptr,
                             const int *__restrict__ ccol,
                             const double *__restrict__ cval, int *cnt,
                             double *invd) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= g.m) return;
    int cur = -1, run = 0;
    for (int q = rptr[r]; q < rptr[r + 1]; q++) {
        const int c = ccol[q];
        if (c == r) {
            invd[r] = 1.0 / cval[q];
            continue;
        }
        const int b = block_of(g, r, c);
        if (r - blk[b].up < 0 || r - blk[b].up >= bm[b]) continue;  // guard
        if (b != cur) {
            if (cur >= 0) cnt[blk[cur].slot_base + (r - blk[cur].up)] = run;
            cur = b;
            run = 0;
        }
        run++;
    }
    if (cur >= 0) cnt[blk[cur].slot_base + (r - blk[cur].up)] = run;
}

__global__ void k_slot_fill(Geom g, const DevBlock *__restrict__ blk,
                            const int *__restrict__ bm,
                            const int *__restrict__ rptr,
                            const int *__restrict__ ccol,
                            const double *__restrict__ cval,
                            const int *__restrict__ rowptr, int *col,
                            double *val) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= g.m) return;
    int cur = -1, o = 0;
    for (int q = rptr[r]; q < rptr[r + 1]; q++) {
        const int c = ccol[q];
        if (c == r) continue;
        const int b = block_of(g, r, c);
        if (r - blk[b].up < 0 || r - blk[b].up >= bm[b]) continue;  // guard
        if (b != cur) {
            cur = b;
            o = rowptr[blk[b].slot_base + (r - blk[b].up)];
        }
        col[o] = c - blk[b].left;
        val[o] = cval[q];
        o++;
    }
}

// One workgroup per block: nnz and cal_longrow statistics over its slots.
// out[b*6 + ...] = {nnz, m_new, lenmax, longrow, longlen, 0}
__global__ void k_block_stats(const DevBlock *__restrict__ blk,
                              const int *__restrict__ bm,
                              const int *__restrict__ cnt, int *out) {
    const int b = blockIdx.x;
    const int base = blk[b].slot_base, n = bm[b];
    long long s_nnz = 0;
    int s_new = 0, s_max = 0, s_lr = 0, s_ll = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int c = cnt[base + i];
        s_nnz += c;
        s_new += c > 0;
        s_max = c > s_max ? c : s_max;
        if (c > kLongRow) {
            s_lr++;
            s_ll += c;
        }
    }
    __shared__ long long sh_nnz[kT];
    __shared__ int sh[4][kT];
    sh_nnz[threadIdx.x] = s_nnz;
    sh[0][threadIdx.x] = s_new;
    sh[1][threadIdx.x] = s_max;
    sh[2][threadIdx.x] = s_lr;
    sh[3][threadIdx.x] = s_ll;
    __syncthreads();
    for (int o = kT / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sh_nnz[threadIdx.x] += sh_nnz[threadIdx.x + o];
            sh[0][threadIdx.x] += sh[0][threadIdx.x + o];
            sh[1][threadIdx.x] = max(sh[1][threadIdx.x], sh[1][threadIdx.x + o]);
            sh[2][threadIdx.x] += sh[2][threadIdx.x + o];
            sh[3][threadIdx.x] += sh[3][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[b * 6 + 0] = (int)sh_nnz[0];
        out[b * 6 + 1] = sh[0][0];
        out[b * 6 + 2] = sh[1][0];
        out[b * 6 + 3] = sh[2][0];
        out[b * 6 + 4] = sh[3][0];
        out[b * 6 + 5] = 0;
    }
}

// flags over the slot space for square blocks: nonempty (DCSR) and long
__global__ void k_sq_flags(const int *__restrict__ slot_kind,
                           const int *__restrict__ cnt, int nslots,
                           int *f_dcsr, int *f_long) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nslots) return;
    const int sq = slot_kind[i];
    f_dcsr[i] = sq && cnt[i] > 0;
    f_long[i] = sq && cnt[i] > kLongRow;
}

__global__ void k_slot_kind(const DevBlock *__restrict__ blk,
                            const int *__restrict__ bm, int nb, int *kind,
                            int *local) {
    const int b = blockIdx.y;
    if (b >= nb) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < bm[b];
         i += gridDim.x * blockDim.x) {
        kind[blk[b].slot_base + i] = blk[b].kind;
        local[blk[b].slot_base + i] = i;
    }
}

__global__ void k_compact(const int *__restrict__ flag,
                          const int *__restrict__ pos,
                          const int *__restrict__ local, int nslots,
                          int *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nslots && flag[i]) out[pos[i]] = local[i];
}

// Block-diagonal matrix D of the triangular blocks (CSC, permuted ids).
__global__ void k_dcol_count(Geom g, const int *__restrict__ pcolptr,
                             const int *__restrict__ prowidx, int *cnt) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.m) return;
    int k = 0;
    for (int q = pcolptr[c]; q < pcolptr[c + 1]; q++)
        k += block_of(g, prowidx[q], c) == block_of(g, c, c);
    cnt[c] = k;
}

__global__ void k_dcol_fill(Geom g, const int *__restrict__ pcolptr,
                            const int *__restrict__ prowidx,
                            const int *__restrict__ dptr, int *didx) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.m) return;
    const int tb = block_of(g, c, c);
    int o = dptr[c];
    for (int q = pcolptr[c]; q < pcolptr[c + 1]; q++) {
        const int r = prowidx[q];
        if (block_of(g, r, c) == tb) didx[o++] = r;
    }
}

__global__ void k_drowcnt(Geom g, const DevBlock *__restrict__ blk,
                          const int *__restrict__ cnt, int *rc) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= g.m) return;
    const int b = block_of(g, r, r);
    rc[r] = cnt[blk[b].slot_base + (r - blk[b].up)] + 1;
}

// level of every node of D and its triangular block id (sort key)
// Block-local row ids of the blocks' discovery orders: position i of the
// item array lies in the triangular block that owns row i (the blocks tile
// the rows in order, each block's order stored at its first row).
__global__ void k_local_ids(Geom g, const int *__restrict__ item,
                            const int *__restrict__ tri_up, int *lvitem) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.m) return;
    lvitem[i] = item[i] - tri_up[block_of(g, i, i) >> 1];
}

// nnz_lv of find_level_nnz (R/utils_cuda.h:781-796): sum of row counts over
// the *position* range [lvptr[L], lvptr[L+1]) of the block.
__global__ void k_nnzlv(const int *__restrict__ lvptr,
                        const int *__restrict__ lvbase,
                        const int *__restrict__ nlv,
                        const int *__restrict__ slot_base,
                        const int *__restrict__ rowptr, int *nnzlv,
                        const int *__restrict__ lvitem_base) {
    const int t = blockIdx.x;
    const int nl = nlv[t];
    const int *lp = lvptr + lvbase[t];
    const int sb = slot_base[t];
    for (int L = threadIdx.x; L < nl; L += blockDim.x) {
        const int a = lp[L], b = lp[L + 1];
        nnzlv[lvitem_base[t] + L] =
            (rowptr[sb + b] - rowptr[sb + a]) + (b - a);
    }
}

template <class T>
rb_status d2h(std::vector<T> &h, const T *d, size_t n, hipStream_t s) {
    h.resize(n);
    if (n == 0) return RB_OK;
    RB_HIP(hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost,
                          s));
    RB_HIP(hipStreamSynchronize(s));
    return RB_OK;
}

template <class T>
rb_status h2d(T *d, const std::vector<T> &h, hipStream_t s) {
    if (h.empty()) return RB_OK;
    RB_HIP(hipMemcpyAsync(d, h.data(), h.size() * sizeof(T),
                          hipMemcpyHostToDevice, s));
    RB_HIP(hipStreamSynchronize(s));
    return RB_OK;
}

int bits_for(int n) {
    int b = 1;
    while ((1LL << b) < (long long)n + 1) b++;
    return b;
}

}  // namespace

// Closed-form partition (R/recblocking_preprocess.h:15-296).
static void partition_blocks(rb_plan *p) {
    const int T = 1 << p->lv, m = p->m, step = p->step;
    p->blocks.assign(2 * T - 1, Block());
    for (int t = 0; t < T; t++) {
        Block &b = p->blocks[2 * t];
        b.kind = 0;
        if (p->subst == 0) {
            b.up = t * step;
            b.down = (t == T - 1) ? m : (t + 1) * step;
        } else {
            b.down = m - t * step;
            b.up = (t == T - 1) ? 0 : m - (t + 1) * step;
        }
        b.left = b.up;
        b.right = b.down;
        b.blk_m = b.blk_n = b.down - b.up;
    }
    for (int t = 1; t < T; t++) {
        Block &b = p->blocks[2 * t - 1];
        const int lb = t & -t;
        b.kind = 1;
        if (p->subst == 0) {
            b.up = t * step;
            b.down = (t + lb == T) ? m : (t + lb) * step;
            b.left = (t - lb) * step;
            b.right = t * step;
        } else {
            b.down = m - t * step;
            b.up = (t + lb == T) ? 0 : m - (t + lb) * step;
            b.left = m - t * step;
            b.right = m - (t - lb) * step;
        }
        b.blk_m = b.down - b.up;
        b.blk_n = b.right - b.left;
    }
}

rb_status build_plan(rb_plan *p, const int *d_colptr, const int *d_rowidx,
                     const double *d_val, hipStream_t s) {
    auto t0 = std::chrono::steady_clock::now();
    const int m = p->m, nnz = p->nnz;
    // RB_PLAN_TRACE=1: host wall-clock ms at the end of each step (kernels
    // are asynchronous: their time shows at the step that next waits)
    const bool trace = getenv("RB_PLAN_TRACE") != nullptr;
    std::vector<std::pair<const char *, double>> marks;
    auto mark = [&](const char *what) {
        if (trace)
            marks.emplace_back(what, std::chrono::duration<double, std::milli>(
                                         std::chrono::steady_clock::now() - t0)
                                         .count());
    };
    // released at the end; one chunk holds the temporaries below (about 17
    // ints per row and 8 ints + 1 double per entry)
    Arena tmp(68 * (size_t)m + 32 * (size_t)nnz + (1 << 20));
    struct Guard {
        Arena &a;
        ~Guard() { a.release(); }
    } guard{tmp};

    // ---- 0. validation
    {
        int hp[2] = {0, 0};
        RB_HIP(hipMemcpyAsync(&hp[0], d_colptr, sizeof(int),
                              hipMemcpyDeviceToHost, s));
        RB_HIP(hipMemcpyAsync(&hp[1], d_colptr + m, sizeof(int),
                              hipMemcpyDeviceToHost, s));
        RB_HIP(hipStreamSynchronize(s));
        if (hp[0] != 0 || hp[1] != nnz) {
            set_error(RB_ERR_BAD_STRUCTURE,
                      "colptr[0] must be 0 and colptr[n] must equal nnz");
            return RB_ERR_BAD_STRUCTURE;
        }
        int *d_err;
        RB_TRY(tmp.alloc(&d_err, 1));
        RB_HIP(hipMemsetAsync(d_err, 0, sizeof(int), s));
        if (m > 0)
            k_validate<<<nblk(m), kT, 0, s>>>(d_colptr, d_rowidx, d_val, m,
                                               nnz, p->subst, d_err);
        RB_HIP(hipGetLastError());
        int herr = 0;
        RB_HIP(hipMemcpyAsync(&herr, d_err, sizeof(int),
                              hipMemcpyDeviceToHost, s));
        RB_HIP(hipStreamSynchronize(s));
        if (herr & 1) {
            set_error(RB_ERR_BAD_STRUCTURE,
                      "CSC structure invalid: row id out of range or outside "
                      "the triangle, or row ids not strictly ascending in a "
                      "column");
            return RB_ERR_BAD_STRUCTURE;
        }
        if (herr & 2) {
            set_error(RB_ERR_MISSING_DIAG,
                      "a column lacks its diagonal entry in the required "
                      "position (first for forward, last for backward)");
            return RB_ERR_MISSING_DIAG;
        }
        if (herr & 4) {
            set_error(RB_ERR_ZERO_DIAG, "zero diagonal value");
            return RB_ERR_ZERO_DIAG;
        }
    }

    Arena &A = p->arena;
    mark("validation");
    // ---- 1. row counts and global level analysis
    int *rowcnt, *bfs_item;
    RB_TRY(tmp.alloc(&rowcnt, m));
    RB_TRY(tmp.alloc(&bfs_item, m));
    RB_TRY(A.alloc(&p->d_levelPtr, m + 2));
    RB_HIP(hipMemsetAsync(rowcnt, 0, sizeof(int) * (size_t)(m > 0 ? m : 1), s));
    if (nnz > 0) k_histogram<<<nblk(nnz), kT, 0, s>>>(d_rowidx, nnz, rowcnt);
    RB_HIP(hipGetLastError());
    RB_TRY(bfs_levels(d_colptr, d_rowidx, rowcnt, m, bfs_item, p->d_levelPtr,
                      &p->nlevel, s));
    auto t_lv = std::chrono::steady_clock::now();

    mark("levels");
    // ---- 2. level-order permutation of the CSC
    int *newpos, *pcolcnt, *ecol;
    RB_TRY(A.alloc(&p->d_levelItem, m));
    RB_TRY(tmp.alloc(&newpos, m));
    RB_TRY(tmp.alloc(&pcolcnt, m + 1));
    RB_TRY(A.alloc(&p->d_pcolptr, m + 1));
    RB_TRY(A.alloc(&p->d_prowidx, nnz));
    RB_TRY(A.alloc(&p->d_pval, nnz));
    RB_TRY(tmp.alloc(&ecol, nnz));
    if (m > 0) {
        k_perm<<<nblk(m), kT, 0, s>>>(bfs_item, m, p->subst, p->d_levelItem,
                                      newpos);
        k_pcolcnt<<<nblk(m), kT, 0, s>>>(d_colptr, p->d_levelItem, m,
                                         pcolcnt);
        RB_HIP(hipGetLastError());
    }
    RB_HIP(hipMemsetAsync(pcolcnt + m, 0, sizeof(int), s));
    RB_TRY(exclusive_scan_i32(pcolcnt, p->d_pcolptr, m + 1, s));
    if (m > 0) {
        k_pcolcopy<<<nblk((long)m * 64), kT, 0, s>>>(
            d_colptr, d_rowidx, d_val, p->d_levelItem, newpos, p->d_pcolptr,
            m, p->d_prowidx, p->d_pval, ecol);
        RB_HIP(hipGetLastError());
    }

    mark("permutation");
    // ---- 3. permuted CSR with ascending columns (stable sort by row)
    int *iota, *sorted_rows, *order, *prowcnt, *prptr, *ccol;
    double *cval;
    RB_TRY(tmp.alloc(&iota, nnz));
    RB_TRY(tmp.alloc(&sorted_rows, nnz));
    RB_TRY(tmp.alloc(&order, nnz));
    RB_TRY(tmp.alloc(&prowcnt, m + 1));
    RB_TRY(tmp.alloc(&prptr, m + 1));
    RB_TRY(tmp.alloc(&ccol, nnz));
    RB_TRY(tmp.alloc(&cval, nnz));
    if (nnz > 0) {
        k_iota<<<nblk(nnz), kT, 0, s>>>(iota, nnz);
        RB_HIP(hipGetLastError());
        RB_TRY(sort_pairs_i32(p->d_prowidx, sorted_rows, iota, order, nnz,
                              bits_for(m), s));
        k_gather_csr<<<nblk(nnz), kT, 0, s>>>(order, ecol, p->d_pval, nnz,
                                              ccol, cval);
    }
    RB_HIP(hipMemsetAsync(prowcnt, 0, sizeof(int) * (size_t)(m + 1), s));
    if (nnz > 0)
        k_histogram<<<nblk(nnz), kT, 0, s>>>(p->d_prowidx, nnz, prowcnt);
    RB_HIP(hipGetLastError());
    RB_TRY(exclusive_scan_i32(prowcnt, prptr, m + 1, s));

    // ---- 9, begun here: the fused schedule's host passes need only the
    // permuted CSR structure and the levels, so they run on a thread while
    // steps 4-8 run; the upload follows step 8
    FusedJob *fjob = nullptr;
    struct JobGuard {  // an early return abandons the job (joins the thread)
        FusedJob *&j;
        ~JobGuard() {
            if (j) (void)build_fused_finish(j, false);
        }
    } jguard{fjob};
    auto t_fs0 = std::chrono::steady_clock::now();
    if (p->want_schedule != RB_SCHED_BLOCKED && m > 0)
        RB_TRY(build_fused_start(p, prptr, ccol, cval, p->fused_waves, s, &fjob));
    auto t_fs1 = std::chrono::steady_clock::now();
    mark("csr");
    // ---- 4. partition and slot space
    p->ntri = 1 << p->lv;
    p->nblocks = 2 * p->ntri - 1;
    p->step = m / p->ntri;
    partition_blocks(p);
    const int nb = p->nblocks;
    std::vector<DevBlock> hb(nb);
    std::vector<int> hbm(nb);
    long long ns = 0;
    for (int i = 0; i < nb; i++) {
        Block &b = p->blocks[i];
        b.slot_base = (int)ns;
        ns += b.blk_m;
        hb[i] = DevBlock{b.up, b.left, b.slot_base, b.kind};
        hbm[i] = b.blk_m;
    }
    if (ns > 0x7fffffffLL - 1) {
        set_error(RB_ERR_INVALID_ARG, "lv too large for this matrix");
        return RB_ERR_INVALID_ARG;
    }
    p->nslots = (int)ns;
    const int NS = p->nslots;
    DevBlock *d_blk;
    int *d_bm, *cnt;
    RB_TRY(tmp.alloc(&d_blk, nb));
    RB_TRY(tmp.alloc(&d_bm, nb));
    RB_TRY(h2d(d_blk, hb, s));
    RB_TRY(h2d(d_bm, hbm, s));
    RB_TRY(tmp.alloc(&cnt, NS + 1));
    RB_TRY(A.alloc(&p->d_invd, m));
    RB_TRY(A.alloc(&p->d_rowptr, NS + 1));
    RB_HIP(hipMemsetAsync(cnt, 0, sizeof(int) * (size_t)(NS + 1), s));
    Geom g{m, p->lv, p->step, p->subst};
    if (m > 0)
        k_slot_count<<<nblk(m), kT, 0, s>>>(g, d_blk, d_bm, prptr, ccol, cval, cnt,
                                            p->d_invd);
    RB_HIP(hipGetLastError());
    RB_TRY(exclusive_scan_i32(cnt, p->d_rowptr, NS + 1, s));
    const int offnnz = nnz - m;  // off-diagonal entries
    RB_TRY(A.alloc(&p->d_col, offnnz));
    RB_TRY(A.alloc(&p->d_val, offnnz));
    if (m > 0)
        k_slot_fill<<<nblk(m), kT, 0, s>>>(g, d_blk, d_bm, prptr, ccol, cval,
                                           p->d_rowptr, p->d_col, p->d_val);
    RB_HIP(hipGetLastError());

    mark("slots");
    // ---- 5. block statistics (blk_nnz, cal_longrow)
    int *d_stats;
    RB_TRY(tmp.alloc(&d_stats, nb * 6));
    k_block_stats<<<nb, kT, 0, s>>>(d_blk, d_bm, cnt, d_stats);
    RB_HIP(hipGetLastError());
    std::vector<int> st;
    RB_TRY(d2h(st, d_stats, (size_t)nb * 6, s));
    for (int i = 0; i < nb; i++) {
        Block &b = p->blocks[i];
        b.blk_nnz = st[i * 6 + 0] + (b.kind == 0 ? b.blk_m : 0);
        if (b.kind == 1) {
            b.m_new = st[i * 6 + 1];
            b.lenmax = st[i * 6 + 2];
            b.longrow = st[i * 6 + 3];
            b.longlen = st[i * 6 + 4];
        }
    }

    mark("block stats");
    // ---- 6. DCSR row lists and long-row lists of square blocks
    {
        int *kind, *local, *fd, *fl, *pd, *pl;
        RB_TRY(tmp.alloc(&kind, NS + 1));
        RB_TRY(tmp.alloc(&local, NS + 1));
        RB_TRY(tmp.alloc(&fd, NS + 1));
        RB_TRY(tmp.alloc(&fl, NS + 1));
        RB_TRY(tmp.alloc(&pd, NS + 1));
        RB_TRY(tmp.alloc(&pl, NS + 1));
        if (NS > 0) {
            dim3 gk(nblk(NS < 65536 ? NS : 65536), nb);
            k_slot_kind<<<gk, kT, 0, s>>>(d_blk, d_bm, nb, kind, local);
            k_sq_flags<<<nblk(NS), kT, 0, s>>>(kind, cnt, NS, fd, fl);
            RB_HIP(hipGetLastError());
        }
        RB_HIP(hipMemsetAsync(fd + NS, 0, sizeof(int), s));
        RB_HIP(hipMemsetAsync(fl + NS, 0, sizeof(int), s));
        RB_TRY(exclusive_scan_i32(fd, pd, NS + 1, s));
        RB_TRY(exclusive_scan_i32(fl, pl, NS + 1, s));
        std::vector<int> hpd, hpl;
        RB_TRY(d2h(hpd, pd, (size_t)NS + 1, s));
        RB_TRY(d2h(hpl, pl, (size_t)NS + 1, s));
        p->ndcsr = hpd[NS];
        p->nlong = hpl[NS];
        for (int i = 0; i < nb; i++) {
            p->blocks[i].dcsr_base = hpd[p->blocks[i].slot_base];
            p->blocks[i].long_base = hpl[p->blocks[i].slot_base];
        }
        RB_TRY(A.alloc(&p->d_dcsr_rows, p->ndcsr));
        RB_TRY(A.alloc(&p->d_long_rows, p->nlong));
        if (NS > 0) {
            k_compact<<<nblk(NS), kT, 0, s>>>(fd, pd, local, NS,
                                              p->d_dcsr_rows);
            k_compact<<<nblk(NS), kT, 0, s>>>(fl, pl, local, NS,
                                              p->d_long_rows);
            RB_HIP(hipGetLastError());
        }
    }

    mark("dcsr");
    // ---- 7. local level analysis of the triangular blocks: each block's
    // own BFS over its in-block entries (the block-diagonal matrix D), all
    // blocks at once, one workgroup each
    const int ntri = p->ntri;
    {
        int *dcnt, *dptr, *didx, *drc, *ditem;
        const int dnnz_max = nnz;
        RB_TRY(tmp.alloc(&dcnt, m + 1));
        RB_TRY(tmp.alloc(&dptr, m + 1));
        RB_TRY(tmp.alloc(&didx, dnnz_max));
        RB_TRY(tmp.alloc(&drc, m));
        RB_TRY(tmp.alloc(&ditem, m));
        if (m > 0) {
            k_dcol_count<<<nblk(m), kT, 0, s>>>(g, p->d_pcolptr, p->d_prowidx,
                                                dcnt);
            RB_HIP(hipGetLastError());
        }
        RB_HIP(hipMemsetAsync(dcnt + m, 0, sizeof(int), s));
        RB_TRY(exclusive_scan_i32(dcnt, dptr, m + 1, s));
        // per-triangular-block bases
        std::vector<int> tri_up(ntri), tri_m(ntri), tri_lvbase(ntri),
            tri_sb(ntri);
        int lb = 0;
        for (int t = 0; t < ntri; t++) {
            Block &b = p->blocks[2 * t];
            tri_up[t] = b.up;
            tri_m[t] = b.blk_m;
            tri_lvbase[t] = lb;
            tri_sb[t] = b.slot_base;
            b.lvptr_base = lb;
            b.lv_base = b.up;  // the blocks tile the rows in order
            lb += b.blk_m + 1;
        }
        RB_TRY(A.alloc(&p->d_lvptr, lb));
        RB_TRY(A.alloc(&p->d_lvitem, m));
        RB_TRY(A.alloc(&p->d_nnzlv, lb));
        int *d_triup, *d_trim, *d_trilvb, *d_nlv, *d_trisb;
        RB_TRY(tmp.alloc(&d_triup, ntri));
        RB_TRY(tmp.alloc(&d_trim, ntri));
        RB_TRY(tmp.alloc(&d_trilvb, ntri));
        RB_TRY(tmp.alloc(&d_nlv, ntri));
        RB_TRY(tmp.alloc(&d_trisb, ntri));
        RB_TRY(h2d(d_triup, tri_up, s));
        RB_TRY(h2d(d_trim, tri_m, s));
        RB_TRY(h2d(d_trilvb, tri_lvbase, s));
        RB_TRY(h2d(d_trisb, tri_sb, s));
        RB_HIP(hipMemsetAsync(d_nlv, 0, sizeof(int) * ntri, s));
        if (m > 0) {
            k_dcol_fill<<<nblk(m), kT, 0, s>>>(g, p->d_pcolptr, p->d_prowidx,
                                               dptr, didx);
            k_drowcnt<<<nblk(m), kT, 0, s>>>(g, d_blk, cnt, drc);
            RB_HIP(hipGetLastError());
            RB_TRY(bfs_blocks(dptr, didx, drc, m, ntri, d_triup, d_trim,
                              d_trilvb, ditem, p->d_lvptr, d_nlv, s));
            k_local_ids<<<nblk(m), kT, 0, s>>>(g, ditem, d_triup, p->d_lvitem);
            RB_HIP(hipGetLastError());
        }
        std::vector<int> hnlv;
        RB_TRY(d2h(hnlv, d_nlv, ntri, s));
        // close each level table with blk_m
        std::vector<int> hlvptr;
        RB_TRY(d2h(hlvptr, p->d_lvptr, (size_t)lb, s));
        for (int t = 0; t < ntri; t++) {
            Block &b = p->blocks[2 * t];
            int nl = hnlv[t];
            if (b.blk_m == 0) nl = 1;
            int *lp = hlvptr.data() + tri_lvbase[t];
            lp[0] = 0;
            lp[nl] = b.blk_m;
            b.nlv = nl;
        }
        RB_TRY(h2d(p->d_lvptr, hlvptr, s));
        std::vector<int> hnl(ntri);
        for (int t = 0; t < ntri; t++) hnl[t] = p->blocks[2 * t].nlv;
        RB_TRY(h2d(d_nlv, hnl, s));
        k_nnzlv<<<ntri, kT, 0, s>>>(p->d_lvptr, d_trilvb, d_nlv, d_trisb,
                                     p->d_rowptr, p->d_nnzlv, d_trilvb);
        RB_HIP(hipGetLastError());
        RB_HIP(hipStreamSynchronize(s));
    }

    mark("local levels");
    // ---- 8. selector (R/recblocking_solver_cuda.h:176-367, :441-546)
    for (int i = 0; i < nb; i++) {
        Block &b = p->blocks[i];
        if (b.kind == 0) {
            if (b.blk_m == b.blk_nnz) {
                b.method = 0;
                b.nlv = 1;
                b.nnzr = b.blk_m > 0 ? b.blk_nnz / b.blk_m : 0;
            } else {
                b.nnzr = b.blk_nnz / b.blk_m;
                if (b.nlv > 20000)
                    b.method = 1;
                else if ((b.nnzr <= 15 && b.nlv <= 20) ||
                         (b.nnzr == 1 && b.nlv <= 100))
                    b.method = 2;
                else
                    b.method = 3;
            }
            if (b.method == 2) {
                RB_TRY(d2h(b.lvptr_h, p->d_lvptr + b.lvptr_base,
                           (size_t)b.nlv + 1, s));
                RB_TRY(d2h(b.nnzlv_h, p->d_nnzlv + b.lvptr_base,
                           (size_t)b.nlv, s));
            }
        } else {
            b.empty_ratio =
                b.blk_m > 0 ? 100.0 * (double)(b.blk_m - b.m_new) / b.blk_m
                            : 0.0;
            if (b.blk_nnz == 0 || b.m_new == 0) {
                b.nnzr = -1;
                b.method = -1;
            } else {
                b.nnzr = (b.blk_nnz - b.longlen) / b.m_new;
                if (b.nnzr <= 12 && b.empty_ratio <= 50)
                    b.method = 0;
                else if (b.nnzr <= 12 && b.empty_ratio > 50)
                    b.method = 1;
                else if (b.nnzr > 12 && b.empty_ratio <= 15)
                    b.method = 2;
                else
                    b.method = 3;
            }
        }
    }

    mark("selector");
    // ---- 9. fused schedule storage (chain-tiled partition of the rows)
    auto t_f0 = std::chrono::steady_clock::now();
    if (fjob) {
        FusedJob *j = fjob;
        fjob = nullptr;
        RB_TRY(build_fused_finish(j, true));
    }
    auto t_f1 = std::chrono::steady_clock::now();

    mark("fused");
    // ---- 10. FP32 + FP64-refinement storage: FP32 copies of the slot
    // values, and the original matrix as CSR for the FP64 residuals (rows,
    // b, r and the x gathers all in the caller's order: the level-ordered
    // rows put neighbouring rows of a grid ~n apart, every access on a
    // cache line of its own)
    if (p->mixed && m > 0) {
        RB_TRY(A.alloc(&p->d_val32, offnnz));
        RB_TRY(A.alloc(&p->d_invd32, m));
        if (offnnz > 0)
            k_f64_to_f32<<<nblk(offnnz), kT, 0, s>>>(p->d_val, offnnz, p->d_val32);
        k_f64_to_f32<<<nblk(m), kT, 0, s>>>(p->d_invd, m, p->d_invd32);
        RB_TRY(A.alloc(&p->d_rptr, m + 1));
        RB_TRY(A.alloc(&p->d_rcol, nnz));
        RB_TRY(A.alloc(&p->d_rval, nnz));
        int *olen;
        RB_TRY(tmp.alloc(&olen, m + 1));
        RB_HIP(hipMemsetAsync(olen + m, 0, sizeof(int), s));
        k_orig_len<<<nblk(m), kT, 0, s>>>(prptr, p->d_levelItem, m, olen);
        RB_HIP(hipGetLastError());
        RB_TRY(exclusive_scan_i32(olen, p->d_rptr, m + 1, s));
        k_orig_rows<<<nblk(m), kT, 0, s>>>(prptr, ccol, cval, p->d_levelItem, m,
                                           p->d_rptr, p->d_rcol, p->d_rval);
        RB_HIP(hipGetLastError());
        RB_TRY(A.alloc(&p->d_bperm32, m));
        RB_TRY(A.alloc(&p->d_xperm32, m));
    }

    mark("fp32 copies");
    // ---- 11. solve scratch
    RB_TRY(A.alloc(&p->d_bperm, m));
    RB_TRY(A.alloc(&p->d_xperm, m));
    RB_TRY(A.alloc(&p->d_tickets, ntri));
    RB_TRY(A.alloc(&p->d_status, 1));
    RB_HIP(hipMemsetAsync(p->d_status, 0, sizeof(int), s));
    RB_TRY(A.alloc(&p->d_ncol, 2));
    RB_HIP(hipMemsetD32Async((hipDeviceptr_t)p->d_ncol, 1, 2, s));
    RB_HIP(hipStreamSynchronize(s));
    auto t1 = std::chrono::steady_clock::now();
    p->preprocess_ms =
        std::chrono::duration<double, std::milli>(t1 - t0).count();
    auto ms = [](auto a, auto b) {
        return std::chrono::duration<double, std::milli>(b - a).count();
    };
    p->phase_ms[0] = ms(t0, t_lv);
    // the fused build's own share: structure copies, then the wait for its
    // host passes and the upload (the passes overlap steps 4-8)
    p->phase_ms[2] = ms(t_fs0, t_fs1) + ms(t_f0, t_f1);
    p->phase_ms[1] = p->preprocess_ms - p->phase_ms[0] - p->phase_ms[2];
    mark("scratch");
    if (trace) {
        fprintf(stderr, "plan build (cumulative ms):");
        for (auto &mk : marks) fprintf(stderr, " %s %.1f", mk.first, mk.second);
        fprintf(stderr, " (%zu temporary buffers)\n", tmp.ptrs.size());
    }
    return RB_OK;
}

}  // namespace rb
