// Solve kernels and the blocked executor.
//
// The blocked executor walks the blocks in the reference's order
// (L_calculate / U_calculate, R/recblocking_calculate.h:15-325) on one
// stream with no host synchronisation: a triangular block runs one of
//   method 0  diagonal scale            (R/utils_sptrsv_cuda.h:202-216)
//   method 2  level-set, one launch per local level, thread or wave per row
//             (R/utils_sptrsv_cuda.h:116-200; rows taken from the block's
//             local level order instead of assuming it is the identity)
//   method 1/3 sync-free CSR pull: a wave takes 64 consecutive rows by
//             ticket (lane per row) and polls the x slots it needs; a slot
//             holds a sentinel NaN until solved, so the value itself is the
//             ready flag (replaces the CSC push with in-degree counters and
//             FP atomics of R/utils_sptrsv_cuda.h:53-114, and cuSPARSE
//             csrsv2 for method 1, which ignores the stored diagonal)
// and a square block one SpMV with the subtraction b -= y fused in
//   method 0/2 over all rows (CSR), 1/3 over the non-empty rows (DCSR)
//   (R/utils_spmv_cuda.h:60-203), with G lanes per row, G the power of two
//   at or above the block's mean row length (the reference uses one thread
//   or one warp), plus a workgroup per row longer than 2048 entries (the
//   long-row pass :30-58).
// Besides the reference selector's kernels a triangular block may run the
// sync-free kernel or a one-workgroup kernel that keeps recent x in an LDS
// ring; at plan build, tune_blocks times the candidates of every large
// block (and the lanes per row of every large square block) and keeps the
// fastest.  Every sum runs in a fixed order, so a solve is bitwise
// reproducible, and the kernels of one block agree bit for bit whenever
// they read a row's entries in the same order.
//
// One walk solves up to 32 right-hand sides: every kernel takes its column
// from the grid and the walk's column count from device memory, so a graph
// captured for a grid width (a power of two) serves every count up to it.
// Batches of factors give each row the column count of its own factor; the
// kernels leave the row alone in later columns.
//
// The kernels are templates on the value type: the FP64 walk is the plain
// solve; the FP32 walk (FP32 copies of the slot values) is the inner solver
// of the FP32 + FP64 iterative refinement at the end of this file, whose
// correction loop runs on the device.
#include <algorithm>
#include <utility>
#include <vector>

#include "rb_internal.h"

namespace rb {

namespace {

constexpr int kT = 256;

// value type <-> same-width unsigned bits, and the type's sentinel
template <class T> struct Bits;
template <> struct Bits<double> {
    using U = unsigned long long;
    static constexpr U kPend = kPending;
    static constexpr U kQNaN = 0x7FF8000000000000ull;
};
template <> struct Bits<float> {
    using U = unsigned int;
    static constexpr U kPend = kPending32;
    static constexpr U kQNaN = 0x7FC00000u;
};

template <class T>
__device__ inline T ld_x(const T *p) {
    using U = typename Bits<T>::U;
    const U u = __hip_atomic_load((const U *)p, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
    return __builtin_bit_cast(T, u);
}

template <class T>
__device__ inline void st_x(T *p, T v) {
    using U = typename Bits<T>::U;
    __hip_atomic_store((U *)p, __builtin_bit_cast(U, v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

template <class T>
__device__ inline bool is_pending(T v) {
    return __builtin_bit_cast(typename Bits<T>::U, v) == Bits<T>::kPend;
}

// Right-hand side blockIdx.y of a walk over several: b and x in the
// caller's layout (leading dimensions ldb, ldx), the permuted vectors and
// tickets of column y at y * ld and y * ntri.
template <class T>
__global__ void k_prologue(const T *__restrict__ b, long ldb,
                           const int *__restrict__ lf, int m, long ld,
                           T *bperm, T *xperm, int *tickets, int ntri,
                           int *ncolw, const int *__restrict__ left) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    // the walk's column count, read by its kernels (graphs are captured
    // for a few grid widths, each solving the first *ncolw columns); zero
    // when a refinement has no column left to correct (*left == 0)
    const bool skip = left && *left == 0;
    if (ncolw && i == 0 && c == 0) *ncolw = skip ? 0 : gridDim.y;
    if (skip) return;
    if (i < m) {
        bperm[c * ld + i] = b[c * ldb + lf[i]];
        ((typename Bits<T>::U *)xperm)[c * ld + i] = Bits<T>::kPend;
    }
    if (i < ntri) tickets[c * ntri + i] = 0;
}

// x back to the caller's order; after a watchdog trip (status set) every
// entry is a quiet NaN, so no solve of a tripped plan hands out a finite x
template <class T>
__global__ void k_epilogue(const T *__restrict__ xperm, long ld,
                           const int *__restrict__ lf, int m, T *x, long ldx,
                           const int *status, const int *__restrict__ left,
                           const int *__restrict__ rowcap) {
    if (left && *left == 0) return;  // the walk was skipped
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i < m)
        x[c * ldx + lf[i]] = (*status & 7) ? __builtin_bit_cast(T, Bits<T>::kQNaN)
                             : rowcap && c >= rowcap[i] ? T(0)  // not solved
                                                        : xperm[c * ld + i];
}

// rowcap (may be null): per permuted row, the columns its factor has in a
// batch; the walk kernels leave the row alone in later columns (nothing
// outside the factor depends on it)
__device__ inline bool skip_row(const int *__restrict__ rowcap, int i) {
    return rowcap && (int)blockIdx.y >= rowcap[i];
}

template <class T>
__global__ void k_tri_diag(const T *__restrict__ invd,
                           const T *__restrict__ bperm, int up, int n,
                           T *xperm, long ld, const int *__restrict__ ncolp,
                           const int *__restrict__ rowcap) {
    if ((int)blockIdx.y >= *ncolp) return;
    bperm += blockIdx.y * ld;
    xperm += blockIdx.y * ld;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !skip_row(rowcap, up + i))
        st_x(&xperm[up + i], bperm[up + i] * invd[up + i]);
}

// level-set, thread per row; rows of one local level
template <class T>
__global__ void k_tri_level_thread(const int *__restrict__ rowptr,
                                   const int *__restrict__ col,
                                   const T *__restrict__ val,
                                   const T *__restrict__ invd,
                                   const T *__restrict__ bperm,
                                   const int *__restrict__ rows, int nrows,
                                   int sb, int up, T *xperm, long ld, const int *__restrict__ ncolp,
                                   const int *__restrict__ rowcap) {
    if ((int)blockIdx.y >= *ncolp) return;
    bperm += blockIdx.y * ld;
    xperm += blockIdx.y * ld;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nrows) return;
    const int r = rows[k];
    if (skip_row(rowcap, up + r)) return;
    const int s = rowptr[sb + r], e = rowptr[sb + r + 1];
    T acc = 0;
    for (int j = s; j < e; j++) acc += val[j] * xperm[up + col[j]];
    st_x(&xperm[up + r], (bperm[up + r] - acc) * invd[up + r]);
}

template <class T>
__device__ inline T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ inline double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// level-set, wave per row
template <class T>
__global__ void k_tri_level_wave(const int *__restrict__ rowptr,
                                 const int *__restrict__ col,
                                 const T *__restrict__ val,
                                 const T *__restrict__ invd,
                                 const T *__restrict__ bperm,
                                 const int *__restrict__ rows, int nrows,
                                 int sb, int up, T *xperm, long ld, const int *__restrict__ ncolp,
                                 const int *__restrict__ rowcap) {
    if ((int)blockIdx.y >= *ncolp) return;
    bperm += blockIdx.y * ld;
    xperm += blockIdx.y * ld;
    const int k = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (k >= nrows) return;
    const int r = rows[k];
    if (skip_row(rowcap, up + r)) return;  // the whole wave: one row
    const int s = rowptr[sb + r], e = rowptr[sb + r + 1];
    T acc = 0;
    for (int j = s + lane; j < e; j += 64) acc += val[j] * xperm[up + col[j]];
    acc = wave_sum(acc);
    if (lane == 0) st_x(&xperm[up + r], (bperm[up + r] - acc) * invd[up + r]);
}

// sync-free CSR pull, lane per row, 64 consecutive rows per wave ticket.
// The grid is one-dimensional with the columns interleaved: workgroup i
// solves column i % width, so consecutive workgroups are the columns of one
// tile and every column's frontier is resident at once.  (With a grid of
// (tiles, columns) the device dispatched all of column 0's tiles first;
// they waited on their level chain while later columns waited for slots,
// and a 32-column walk of the batch union ran its columns a few at a time.)
// Within a column the tickets order the tiles, so every wait is on a tile
// that a workgroup already holds.
template <class T>
__global__ void __launch_bounds__(kT)
k_tri_syncfree(const int *__restrict__ rowptr, const int *__restrict__ col,
               const T *__restrict__ val, const T *__restrict__ invd,
               const T *__restrict__ bperm, int sb, int up, int bm,
               int subst, int *ticket, T *xperm, int *status,
               unsigned spin_limit, long ld, int tstride, const int *__restrict__ ncolp,
               const int *__restrict__ rowcap, int width) {
    const int c = (int)(blockIdx.x % (unsigned)width);
    if (c >= *ncolp) return;
    bperm += c * ld;
    xperm += c * ld;
    ticket += c * tstride;
    const int lane = threadIdx.x & 63;
    int tile = 0;
    if (lane == 0) tile = atomicAdd(ticket, 1);
    tile = __shfl(tile, 0, 64);
    const int pos = tile * 64 + lane;
    const int r = subst == 0 ? pos : bm - 1 - pos;
    const bool valid = pos < bm && !(rowcap && c >= rowcap[up + r]);
    int j = 0, e = 0;
    T bv = 0, iv = 0;
    if (valid) {
        j = rowptr[sb + r];
        e = rowptr[sb + r + 1];
        bv = bperm[up + r];
        iv = invd[up + r];
    }
    T acc = 0;
    bool pend = valid;
    const T *xb = xperm + up;
    unsigned spins = 0;
    while (__any(pend)) {
        if (pend) {
            while (j < e) {
                const T xv = ld_x(&xb[col[j]]);
                if (is_pending(xv)) break;
                acc += val[j] * xv;
                j++;
            }
            if (j == e) {
                st_x(&xperm[up + r], (bv - acc) * iv);
                pend = false;
            }
        }
        if (__any(pend)) {
            // watchdog: a dependency that never arrives means corrupt
            // storage; flag it and release the row instead of hanging
            if (++spins > spin_limit) {
                if (pend) {
                    atomicOr(status, 1);
                    st_x(&xperm[up + r], __builtin_bit_cast(T, Bits<T>::kQNaN));
                    pend = false;
                }
            } else {
                __builtin_amdgcn_s_sleep(1);
            }
        }
    }
}

// One workgroup solves the whole triangular block: thread t takes the
// rows at positions t, t + 1024, ... (solve order: the slot order of a
// forward block, reversed for a backward one, which the global level order
// makes a topological order) and keeps a ring of the last kRingRows
// solved rows' x in LDS, each slot tagged with its row's position
// (release store after the value; acquire load before reading it).  A
// row polls only the ring -- a dependency is an LDS read, not a trip
// through device memory as in the sync-free grid kernel, nor a launch per
// level as in the level-set kernels -- so deep, narrow blocks (hundreds of
// short levels) run faster.  The rows of thread t's k-th pass form round
// k; a wave that is done with its round k moves on, but it starts round
// k only once every wave is done with round k - kGroupSlack, and round
// k writes the slots of positions kRingRows back.  So when every in-block
// dependency is at most kRingRows - kGroupSlack * kGroupThreads positions
// back (the host checks the distance), a slot is reused only after every
// row that reads it has finished.  (Without the gate a wave whose rows
// have no in-block dependencies would run rounds ahead and overwrite the
// slots slower rows still wait for.)  Lanes of one wave may depend on
// each other, so a wave loops
// until all its lanes are done, each lane advancing as far as its
// dependencies allow; a row's entries are read ahead in chunks of kChunk,
// before the waits, and summed in order, as by the other kernels.
// (A level-synchronous form -- one workgroup barrier per level, rows read
// 1024 positions ahead -- took ~1.1 us per level on the banded factor
// against this form's ~0.9 us: a wave waits for its newest loads as a
// whole, so the reads ahead for some lanes held the others at every
// level.  Reading a chunk's tags together and its values together after
// one acquire fence, instead of a tag and a value per entry, was slower
// too: 1.47 against 0.91 ms on the banded 100k factor, 5.6 against 4.8 ms
// on the 1341^2 Laplacian; and so was a stalled row re-reading only the
// tag it waits for until it arrives, 1.43 against 0.91 ms.)
template <class T>
__global__ void __launch_bounds__(kGroupThreads)
k_tri_group(const int *__restrict__ rowptr, const int *__restrict__ col,
            const T *__restrict__ val, const T *__restrict__ invd,
            const T *__restrict__ bperm, int sb, int up, int bm, int subst,
            T *xperm, int *status, unsigned spin_limit, long ld, const int *__restrict__ ncolp,
            const int *__restrict__ rowcap) {
    if ((int)blockIdx.y >= *ncolp) return;
    bperm += blockIdx.y * ld;
    xperm += blockIdx.y * ld;
    constexpr int kChunk = 8;
    constexpr int kWaves = kGroupThreads / 64;
    __shared__ int tag[kRingRows];
    __shared__ T xr[kRingRows];
    __shared__ int wround[kWaves];  // last round each wave has finished
    for (int i = threadIdx.x; i < kRingRows; i += kGroupThreads) tag[i] = -1;
    if (threadIdx.x < kWaves) wround[threadIdx.x] = -1;
    __syncthreads();
    bool bad = false;
    const int lane = threadIdx.x & 63;
    int round = 0;
    for (int p0 = threadIdx.x & ~63; p0 < bm; p0 += kGroupThreads, round++) {
        // rounds before kRingRows / kGroupThreads reuse no slot
        if (round >= kRingRows / kGroupThreads) {
            const int need = round - kGroupSlack;
            unsigned gspins = 0;
            while (__any(lane < kWaves &&
                         __hip_atomic_load(&wround[lane < kWaves ? lane : 0],
                                           __ATOMIC_ACQUIRE,
                                           __HIP_MEMORY_SCOPE_WORKGROUP) < need)) {
                if (++gspins > spin_limit) {  // flag it and go on
                    bad = true;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        const int pos = p0 + (threadIdx.x & 63);
        const int r = subst == 0 ? pos : bm - 1 - pos;
        // a skipped row publishes nothing: only rows of its own factor,
        // skipped too, read it
        const bool valid = pos < bm && !skip_row(rowcap, up + r);
        int jc = 0, e = 0;  // chunk base, row end
        T bv = 0, iv = 0;
        if (valid) {
            jc = rowptr[sb + r];
            e = rowptr[sb + r + 1];
            bv = bperm[up + r];
            iv = invd[up + r];
        }
        int cc[kChunk];
        T vv[kChunk];
        auto load_chunk = [&]() {
#pragma unroll
            for (int q = 0; q < kChunk; q++) {
                cc[q] = jc + q < e ? col[jc + q] : 0;
                vv[q] = jc + q < e ? val[jc + q] : T(0);
            }
        };
        load_chunk();
        int done = 0;  // entries of the chunk already summed
        T acc = 0;
        bool pend = valid;
        unsigned spins = 0;
        while (__any(pend)) {
            if (pend) {
                bool stall = false;
                while (true) {
#pragma unroll
                    for (int q = 0; q < kChunk; q++) {
                        if (!stall && q >= done && jc + q < e) {
                            const int c = cc[q];
                            const int pp = subst == 0 ? c : bm - 1 - c;
                            const int k = pp & (kRingRows - 1);
                            if (__hip_atomic_load(&tag[k], __ATOMIC_ACQUIRE,
                                                  __HIP_MEMORY_SCOPE_WORKGROUP) == pp) {
                                acc += vv[q] * xr[k];
                                done = q + 1;
                            } else {
                                stall = true;
                            }
                        }
                    }
                    if (stall || jc + kChunk >= e) break;
                    jc += kChunk;  // the chunk is summed: the next one
                    done = 0;
                    load_chunk();
                }
                if (!stall) {
                    const T x = (bv - acc) * iv;
                    const int k = pos & (kRingRows - 1);
                    xr[k] = x;
                    __hip_atomic_store(&tag[k], pos, __ATOMIC_RELEASE,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                    xperm[up + r] = x;
                    pend = false;
                }
            }
            if (__any(pend)) {
                if (++spins > spin_limit) {
                    // a dependency that never arrives: flag it, release the
                    // row (the epilogue turns x into NaN)
                    if (pend) {
                        bad = true;
                        const int k = pos & (kRingRows - 1);
                        xr[k] = __builtin_bit_cast(T, Bits<T>::kQNaN);
                        __hip_atomic_store(&tag[k], pos, __ATOMIC_RELEASE,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                        pend = false;
                    }
                } else {
                    __builtin_amdgcn_s_sleep(1);
                }
            }
        }
        if (lane == 0)
            __hip_atomic_store(&wround[threadIdx.x >> 6], round, __ATOMIC_RELEASE,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    // a wave that has left holds no other wave back
    if (lane == 0)
        __hip_atomic_store(&wround[threadIdx.x >> 6], 0x7fffffff, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
    if (bad) atomicOr(status, 1);
}

// Square block SpMV, b[up + r] -= A[r, :] x[left:], with G lanes per row
// (G a power of two, 1 .. 64): the group reads its row's entries
// coalesced, then a fixed-order butterfly; rows from the DCSR list (the
// non-empty rows) or all rows of the block; rows longer than 2048 entries
// are left to k_sq_long.  Each thread keeps KC right-hand sides (columns
// blockIdx.y * KC .. + KC - 1 of *ncolp, ld apart), so the block's entries
// are read once per KC columns.
template <class T, int G, bool DCSR, int KC>
__global__ void __launch_bounds__(kT)
k_sq_group(const int *__restrict__ rowptr, const int *__restrict__ col,
           const T *__restrict__ val, const T *__restrict__ xperm,
           const int *__restrict__ rows, int nrows, int sb, int up, int left,
           T *bperm, long ld, const int *__restrict__ ncolp,
           const int *__restrict__ rowcap) {
    const int c0 = blockIdx.y * KC;
    const int ncol = *ncolp;
    if (c0 >= ncol) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = t / G, g = t & (G - 1);
    if (k >= nrows) return;  // whole groups leave together
    const int r = DCSR ? rows[k] : k;
    const int s = rowptr[sb + r], e = rowptr[sb + r + 1];
    if (s == e || e - s > kLongRow) return;
    const int nc = min(KC, min(ncol, rowcap ? rowcap[up + r] : ncol) - c0);
    if (nc <= 0) return;
    const T *x0 = xperm + c0 * ld + left;
    T acc[KC];
#pragma unroll
    for (int c = 0; c < KC; c++) acc[c] = 0;
    for (int j = s + g; j < e; j += G) {
        const T v = val[j];
        const int cj = col[j];
#pragma unroll
        for (int c = 0; c < KC; c++)
            if (c < nc) acc[c] += v * x0[c * ld + cj];
    }
#pragma unroll
    for (int c = 0; c < KC; c++) {
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) acc[c] += __shfl_xor(acc[c], o, G);
        if (g == 0 && c < nc) bperm[(c0 + c) * ld + up + r] -= acc[c];
    }
}

// one workgroup per long row (> 2048 entries) and right-hand side
// (blockIdx.y), fixed-order tree reduction
template <class T>
__global__ void __launch_bounds__(kT)
k_sq_long(const int *__restrict__ rowptr, const int *__restrict__ col,
          const T *__restrict__ val, const T *__restrict__ xperm,
          const int *__restrict__ rows, int sb, int up, int left, T *bperm,
          long ld, const int *__restrict__ ncolp, const int *__restrict__ rowcap) {
    if ((int)blockIdx.y >= *ncolp) return;
    xperm += blockIdx.y * ld;
    bperm += blockIdx.y * ld;
    const int r = rows[blockIdx.x];
    if (skip_row(rowcap, up + r)) return;  // the whole workgroup
    const int s = rowptr[sb + r], e = rowptr[sb + r + 1];
    T acc = 0;
    for (int j = s + threadIdx.x; j < e; j += kT)
        acc += val[j] * xperm[left + col[j]];
    acc = wave_sum(acc);
    __shared__ T sh[kT / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        T t = 0;
        for (int w = 0; w < kT / 64; w++) t += sh[w];
        bperm[up + r] -= t;
    }
}

inline int nblk(long n, int t = kT) { return (int)((n + t - 1) / t); }

// lanes per row for a square block's SpMV: the power of two at or above
// its mean non-empty row length (long rows excluded), 1 .. 64
inline int lanes_per_row(int nnzr) {
    int g = 1;
    while (g < nnzr && g < 64) g <<= 1;
    return g;
}

// columns per thread of the SpMV for ncol right-hand sides
inline int cols_per_thread(int ncol) { return ncol >= 4 ? 4 : ncol >= 2 ? 2 : 1; }

template <class T, bool DCSR, int KC>
void launch_sq_kc(int g, int nrows, int ncol, hipStream_t s, const int *rowptr,
                  const int *col, const T *val, const T *xperm, const int *rows,
                  int sb, int up, int left, T *bperm, long ld, const int *ncolp,
                  const int *rowcap) {
    const long th = (long)nrows * g;
    const dim3 grid(nblk(th), (ncol + KC - 1) / KC);
#define RB_SQ(G) k_sq_group<T, G, DCSR, KC><<<grid, kT, 0, s>>>(rowptr, col, val, xperm, rows, nrows, sb, up, left, bperm, ld, ncolp, rowcap)
    switch (g) {
    case 1: RB_SQ(1); break;
    case 2: RB_SQ(2); break;
    case 4: RB_SQ(4); break;
    case 8: RB_SQ(8); break;
    case 16: RB_SQ(16); break;
    case 32: RB_SQ(32); break;
    default: RB_SQ(64); break;
    }
#undef RB_SQ
}

template <class T, bool DCSR>
void launch_sq(int g, int nrows, int ncol, hipStream_t s, const int *rowptr,
               const int *col, const T *val, const T *xperm, const int *rows,
               int sb, int up, int left, T *bperm, long ld, const int *ncolp,
               const int *rowcap) {
    switch (cols_per_thread(ncol)) {
    case 4: launch_sq_kc<T, DCSR, 4>(g, nrows, ncol, s, rowptr, col, val, xperm, rows, sb, up, left, bperm, ld, ncolp, rowcap); break;
    case 2: launch_sq_kc<T, DCSR, 2>(g, nrows, ncol, s, rowptr, col, val, xperm, rows, sb, up, left, bperm, ld, ncolp, rowcap); break;
    default: launch_sq_kc<T, DCSR, 1>(g, nrows, ncol, s, rowptr, col, val, xperm, rows, sb, up, left, bperm, ld, ncolp, rowcap); break;
    }
}

// ---------------------------------------------------- refinement kernels
// Several right-hand sides at once: column c = blockIdx.y; the refinement
// scratch (r, r32, d32) holds column c at c * m; per column, bmax / rmax
// hold max |b| / max |r| as the bits of a non-negative double (ordered like
// the values), done[c] is set once the column has converged, and the
// columns that are done skip every later pass.

// r = b, x = 0, max |b|
// x = 0, r = b and max |b| of column blockIdx.y: the maxima are folded
// per workgroup, one atomic each (an atomic per wave on the column's one
// word serialised: 55 us on 2.25M rows)
__global__ void __launch_bounds__(kT)
    k_ir_init(const double *__restrict__ b, long ldb, double *x, long ldx,
              double *r, int m, unsigned long long *bmax) {
    __shared__ double wmax[kT / 64];
    const int c = blockIdx.y;
    double mx = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += gridDim.x * blockDim.x) {
        const double v = b[c * ldb + i];
        r[(long)c * m + i] = v;
        x[c * ldx + i] = 0.0;
        mx = fmax(mx, fabs(v));
    }
    mx = wave_max(mx);
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kT / 64; w++) mx = fmax(mx, wmax[w]);
        atomicMax(&bmax[c], (unsigned long long)__double_as_longlong(mx));
    }
}

__global__ void k_ir_to32(const double *__restrict__ r, int m, float *r32,
                          const int *__restrict__ done) {
    const int c = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (!done[c] && i < m) r32[(long)c * m + i] = (float)r[(long)c * m + i];
}

// x += d (FP64 update with the FP32 correction)
__global__ void k_ir_update(const float *__restrict__ d, int m, double *x,
                            long ldx, const int *__restrict__ done) {
    const int c = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (!done[c] && i < m) x[c * ldx + i] += (double)d[(long)c * m + i];
}

// r = b - A x in the original ordering, FP64, over the original matrix as
// CSR (rows, b, r and x all in the caller's order); max |r| per column.  G lanes per row (a power of two at or above the mean row
// length, as for the square blocks' SpMV) read the row's entries coalesced
// and sum them in a fixed order (a thread per row was slow on the
// power-law factor, whose rows reach 512 entries), for KC columns at once
// (blockIdx.y * KC ..), so the matrix is read once per KC columns instead
// of once per column.  The grid is at most 1024 workgroups per column
// group, each looping over rows and folding its maxima into one atomic per
// column: an atomic per wave on a column's one max word serialised (half a
// million of them per correction with 8 lanes per row on 4M rows).
template <int G, int KC>
__global__ void __launch_bounds__(kT)
k_ir_residual(const int *__restrict__ rptr, const int *__restrict__ rcol,
              const double *__restrict__ rval, int m, int ncol, const double *__restrict__ b, long ldb,
              const double *__restrict__ x, long ldx, double *r,
              unsigned long long *rmax, const int *__restrict__ done) {
    const int c0 = blockIdx.y * KC;
    bool live[KC];
    bool any = false;
#pragma unroll
    for (int q = 0; q < KC; q++) {
        live[q] = c0 + q < ncol && !done[c0 + q];
        any = any || live[q];
    }
    if (!any) return;  // the whole workgroup
    __shared__ double wmax[KC][kT / 64];
    const int g = threadIdx.x & (G - 1);
    const long rows_per_pass = (long)gridDim.x * (blockDim.x / G);
    double a[KC];
#pragma unroll
    for (int q = 0; q < KC; q++) a[q] = 0.0;
    // i is the same for the G lanes of a group, so they loop together and
    // their shuffles stay inside the group
    for (long i = ((long)blockIdx.x * blockDim.x + threadIdx.x) / G; i < m;
         i += rows_per_pass) {
        double acc[KC];
#pragma unroll
        for (int q = 0; q < KC; q++) acc[q] = 0.0;
        for (int k = rptr[i] + g; k < rptr[i + 1]; k += G) {
            const double v = rval[k];
            const int cj = rcol[k];
#pragma unroll
            for (int q = 0; q < KC; q++)
                if (live[q]) acc[q] += v * x[(c0 + q) * ldx + cj];
        }
#pragma unroll
        for (int q = 0; q < KC; q++) {
#pragma unroll
            for (int o = G / 2; o > 0; o >>= 1) acc[q] += __shfl_xor(acc[q], o, G);
        }
        if (g == 0) {
#pragma unroll
            for (int q = 0; q < KC; q++)
                if (live[q]) {
                    const int c = c0 + q;
                    const double v = b[c * ldb + i] - acc[q];
                    r[(long)c * m + i] = v;
                    // fmax drops a NaN operand: a NaN residual counts as
                    // +inf, so its column never passes for converged
                    a[q] = fmax(a[q], v == v ? fabs(v) : HUGE_VAL);
                }
        }
    }
#pragma unroll
    for (int q = 0; q < KC; q++) {
        const double w = wave_max(a[q]);
        if ((threadIdx.x & 63) == 0) wmax[q][threadIdx.x >> 6] = w;
    }
    __syncthreads();
    if (threadIdx.x < KC && live[threadIdx.x]) {
        double t = 0.0;
        for (int w = 0; w < kT / 64; w++) t = fmax(t, wmax[threadIdx.x][w]);
        atomicMax(&rmax[c0 + threadIdx.x], (unsigned long long)__double_as_longlong(t));
    }
}

template <int G, int KC>
static void launch_residual(rb_plan *p, int nc, const double *b, long ldb,
                            const double *x, long ldx, hipStream_t s) {
    const int m = p->m;
    const int grid = (int)std::min<long>(1024, nblk((long)m * G));
    k_ir_residual<G, KC><<<dim3(grid, (nc + KC - 1) / KC), kT, 0, s>>>(
        p->d_rptr, p->d_rcol, p->d_rval, m, nc, b, ldb, x,
        ldx, p->d_r, p->d_norms + nc, p->d_irdone);
}

template <int KC>
static void residual_kc(rb_plan *p, int nc, const double *b, long ldb,
                        const double *x, long ldx, hipStream_t s) {
    switch (lanes_per_row(p->m > 0 ? (int)((long)p->nnz / p->m) : 1)) {
    case 1: launch_residual<1, KC>(p, nc, b, ldb, x, ldx, s); break;
    case 2: launch_residual<2, KC>(p, nc, b, ldb, x, ldx, s); break;
    case 4: launch_residual<4, KC>(p, nc, b, ldb, x, ldx, s); break;
    case 8: launch_residual<8, KC>(p, nc, b, ldb, x, ldx, s); break;
    case 16: launch_residual<16, KC>(p, nc, b, ldb, x, ldx, s); break;
    case 32: launch_residual<32, KC>(p, nc, b, ldb, x, ldx, s); break;
    default: launch_residual<64, KC>(p, nc, b, ldb, x, ldx, s); break;
    }
}

static void residual(rb_plan *p, int nc, const double *b, long ldb,
                     const double *x, long ldx, hipStream_t s) {
    switch (cols_per_thread(nc)) {
    case 4: residual_kc<4>(p, nc, b, ldb, x, ldx, s); break;
    case 2: residual_kc<2>(p, nc, b, ldb, x, ldx, s); break;
    default: residual_kc<1>(p, nc, b, ldb, x, ldx, s); break;
    }
}

// after correction `it`: a column whose max |r| <= tol * max |b| is done;
// its relative residual and correction count are kept, rmax is cleared
// for the next residual
__global__ void k_ir_check(const unsigned long long *__restrict__ bmax,
                           unsigned long long *rmax, int ncol, double tol,
                           int *done, int *iters, double *rel, int it,
                           int *left) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncol || done[c]) return;
    const double bn = __longlong_as_double((long long)bmax[c]);
    const double rn = __longlong_as_double((long long)rmax[c]);
    const double q = bn > 0.0 ? rn / bn : 0.0;
    rel[c] = q;
    iters[c] = it;
    if (q <= tol) {
        done[c] = 1;
        atomicSub(left, 1);
    }
    rmax[c] = 0;
}

__global__ void k_ir_reset(unsigned long long *norms, int *done, int *iters,
                           double *rel, int ncol, int *left) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c == 0) *left = ncol;
    if (c < ncol) {
        norms[c] = 0;
        norms[ncol + c] = 0;
        done[c] = 0;
        iters[c] = 0;
        rel[c] = 0.0;
    }
}

// asynchronous calls: a column that missed tol marks the plan (bit 8)
__global__ void k_ir_flag(const int *__restrict__ done, int ncol, int *status) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncol && !done[c]) atomicOr(status, 8);
}

// largest in-block dependency distance |row - column| of every triangular
// block (blockIdx.y = triangular block; sb / bm its slots and rows)
__global__ void k_tri_maxdist(const int *__restrict__ rowptr,
                              const int *__restrict__ col,
                              const int *__restrict__ sbs,
                              const int *__restrict__ bms, int *out) {
    const int t = blockIdx.y, sb = sbs[t], bm = bms[t];
    int mx = 0;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < bm;
         r += gridDim.x * blockDim.x)
        for (int j = rowptr[sb + r]; j < rowptr[sb + r + 1]; j++)
            mx = max(mx, abs(r - col[j]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(&out[t], mx);
}

// before a timed run of a triangular block: its x slots back to the
// sentinel and its ticket to zero
__global__ void k_fill_int(int *p, long n, int v) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_block_reset(double *x, int n, int *ticket) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ((unsigned long long *)x)[i] = kPending;
    if (i == 0) *ticket = 0;
}

}  // namespace

// One block of the walk in value type T (the kernels of B.kern / B.lanes)
// on a grid of ncol columns, solving the first *nc of them: column c of
// the permuted b and x at c * ld, its tickets at c * ntri.
template <class T>
static void launch_block(rb_plan *p, int i, const T *val, const T *invd,
                         T *bperm, T *xperm, int *tickets, int ncol, long ld,
                         const int *nc, hipStream_t s) {
    const Block &B = p->blocks[i];
    if (B.blk_m == 0 || ncol <= 0) return;
    if (B.kind == 0) {
        if (B.method == 0) {
            k_tri_diag<T><<<dim3(nblk(B.blk_m), ncol), kT, 0, s>>>(
                invd, bperm, B.up, B.blk_m, xperm, ld, nc, p->d_rowcap);
        } else if (B.kern == kTriGroup) {
            k_tri_group<T><<<dim3(1, ncol), kGroupThreads, 0, s>>>(
                p->d_rowptr, p->d_col, val, invd, bperm, B.slot_base, B.up,
                B.blk_m, p->subst, xperm, p->d_status, p->spin_limit, ld, nc,
                p->d_rowcap);
        } else if (B.kern == kTriRef && B.method == 2) {
            for (int L = 0; L < B.nlv; L++) {
                const int a = B.lvptr_h[L], c = B.lvptr_h[L + 1];
                const int nr = c - a;
                if (nr <= 0) continue;
                const int *rows = p->d_lvitem + B.lv_base + a;
                if (L == 0 || B.nnzlv_h[L] / nr <= 15)
                    k_tri_level_thread<T><<<dim3(nblk(nr), ncol), kT, 0, s>>>(
                        p->d_rowptr, p->d_col, val, invd, bperm, rows, nr,
                        B.slot_base, B.up, xperm, ld, nc, p->d_rowcap);
                else
                    k_tri_level_wave<T><<<dim3(nblk((long)nr * 64), ncol), kT, 0, s>>>(
                        p->d_rowptr, p->d_col, val, invd, bperm, rows, nr,
                        B.slot_base, B.up, xperm, ld, nc, p->d_rowcap);
            }
        } else {  // methods 1 and 3, or kTriSyncfree
            k_tri_syncfree<T><<<(unsigned)nblk((long)B.blk_m, 64) * ncol, 64, 0, s>>>(
                p->d_rowptr, p->d_col, val, invd, bperm, B.slot_base,
                B.up, B.blk_m, p->subst, tickets + i / 2, xperm,
                p->d_status, p->spin_limit, ld, p->ntri, nc, p->d_rowcap, ncol);
        }
        return;
    }
    // The reference's method picks CSR (0, 2) or DCSR (1, 3) rows; lanes
    // per row are the block's own (the reference's thread / warp split is
    // 1 or a whole wave).
    const int *rows = p->d_dcsr_rows + B.dcsr_base;
    const bool dcsr = B.method == 1 || B.method == 3;
    const int nr = dcsr ? B.m_new : B.blk_m;
    if (B.method < 0 || nr == 0) return;  // empty block
    if (dcsr)
        launch_sq<T, true>(B.lanes, nr, ncol, s, p->d_rowptr, p->d_col, val,
                           xperm, rows, B.slot_base, B.up, B.left, bperm, ld, nc,
                           p->d_rowcap);
    else
        launch_sq<T, false>(B.lanes, nr, ncol, s, p->d_rowptr, p->d_col, val,
                            xperm, rows, B.slot_base, B.up, B.left, bperm, ld, nc,
                            p->d_rowcap);
    if (B.longrow > 0)
        k_sq_long<T><<<dim3(B.longrow, ncol), kT, 0, s>>>(
            p->d_rowptr, p->d_col, val, xperm, p->d_long_rows + B.long_base,
            B.slot_base, B.up, B.left, bperm, ld, nc, p->d_rowcap);
}

// The block walk of one solve in value type T: kernels on plan-owned
// buffers only, so it can be captured once into a graph and replayed.
template <class T>
static void launch_blocks(rb_plan *p, const T *val, const T *invd, T *bperm,
                          T *xperm, int *tickets, int ncol, long ld,
                          const int *nc, hipStream_t s) {
    for (int i = 0; i < p->nblocks; i++)
        launch_block<T>(p, i, val, invd, bperm, xperm, tickets, ncol, ld, nc, s);
}

// Capture the walk on a grid of ncol columns (a power of two) into a
// graph, once per value type and width; a replay solves the first
// *d_ncol columns, which the prologue sets.  The scratch the graphs hold
// grows by powers of two (at most log2 of the column cap times), so a
// stream of solves of varying widths replays at most six graphs per type.
// Returns false when capture is not possible; the walk is then launched
// kernel by kernel.
template <class T>
static bool capture_blocks(rb_plan *p, int ncol, void **exec) {
    constexpr int ty = sizeof(T) == 8 ? 0 : 1;
    *exec = nullptr;
    if (p->graph_failed) return false;
    for (auto &e : p->graphs[ty])
        if (e.first == ncol) {
            *exec = e.second;
            return true;
        }
    const T *val = ty == 0 ? (const T *)p->d_val : (const T *)p->d_val32;
    const T *invd = ty == 0 ? (const T *)p->d_invd : (const T *)p->d_invd32;
    T *bperm = ty == 0 ? (T *)p->d_bperm : (T *)p->d_bperm32;
    T *xperm = ty == 0 ? (T *)p->d_xperm : (T *)p->d_xperm32;
    hipStream_t cs = nullptr;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    bool ok = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) == hipSuccess;
    if (ok) ok = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (ok) {
        launch_blocks<T>(p, val, invd, bperm, xperm, p->d_tickets, ncol, p->m,
                         p->d_ncol, cs);
        ok = hipStreamEndCapture(cs, &g) == hipSuccess && g != nullptr;
    }
    if (ok) ok = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) == hipSuccess;
    if (g) (void)hipGraphDestroy(g);
    if (cs) (void)hipStreamDestroy(cs);
    (void)hipGetLastError();
    if (!ok) {
        p->graph_failed = true;
        return false;
    }
    p->graphs[ty].push_back({ncol, (void *)ge});
    *exec = ge;
    return true;
}

void free_solve_scratch(rb_plan *p) {
    if (p->kcols > 1) {
        (void)hipFree(p->d_bperm);
        (void)hipFree(p->d_xperm);
    }
    if (p->kcols32 > 1) {
        (void)hipFree(p->d_bperm32);
        (void)hipFree(p->d_xperm32);
    }
    if (p->kticket > 1) (void)hipFree(p->d_tickets);
    p->kcols = p->kcols32 = p->kticket = 1;
    free_ir(p);
}

static void drop_graphs(rb_plan *p, int ty) {
    for (auto &e : p->graphs[ty]) (void)hipGraphExecDestroy((hipGraphExec_t)e.second);
    p->graphs[ty].clear();
}

void destroy_graph(rb_plan *p) {
    drop_graphs(p, 0);
    drop_graphs(p, 1);
}

void rowcap_for_call(rb_plan *p, hipStream_t s) {
    if (!p->d_rowcap) return;
    if (p->rowcap_set) {
        p->rowcap_set = false;
        return;
    }
    if (!p->rowcap_all && p->m > 0) {
        k_fill_int<<<nblk(p->m), kT, 0, s>>>(p->d_rowcap, p->m, 0x7fffffff);
        p->rowcap_all = true;
    }
}

// Right-hand sides one walk takes: up to 32, within 4 GiB of permuted b
// and x (32 columns up to 8M rows).
static int walk_cols(const rb_plan *p) {
    const long per = 16L * std::max(1, p->m);
    return (int)std::max(1L, std::min(32L, (4L << 30) / per));
}

// Grid width of a walk of ncol columns: the next power of two, at most
// the cap (itself sized for the scratch).
static int walk_width(const rb_plan *p, int ncol) {
    int w = 1;
    while (w < ncol) w *= 2;
    return std::min(w, walk_cols(p));
}

// Permuted b / x and tickets for ncol right-hand sides (value type T);
// growing them drops the graphs that hold the old buffers.
template <class T>
static rb_status ensure_cols(rb_plan *p, int ncol, hipStream_t s) {
    constexpr int ty = sizeof(T) == 8 ? 0 : 1;
    int &cap = ty == 0 ? p->kcols : p->kcols32;
    if (ncol <= cap) return RB_OK;
    RB_HIP(hipStreamSynchronize(s));
    T *&bp = ty == 0 ? (T *&)p->d_bperm : (T *&)p->d_bperm32;
    T *&xp = ty == 0 ? (T *&)p->d_xperm : (T *&)p->d_xperm32;
    const size_t n = (size_t)ncol * p->m;
    T *nb = nullptr, *nx = nullptr;
    int *nt = nullptr;
    if (hipMalloc(&nb, sizeof(T) * n) != hipSuccess ||
        hipMalloc(&nx, sizeof(T) * n) != hipSuccess ||
        (ncol > p->kticket &&
         hipMalloc(&nt, sizeof(int) * (size_t)ncol * std::max(1, p->ntri)) != hipSuccess)) {
        (void)hipGetLastError();
        if (nb) (void)hipFree(nb);
        if (nx) (void)hipFree(nx);
        set_error(RB_ERR_NO_MEMORY, "hipMalloc failed (right-hand side scratch)");
        return RB_ERR_NO_MEMORY;
    }
    drop_graphs(p, 0);
    drop_graphs(p, 1);
    // the first buffers belong to the plan's arena; later ones to the plan
    if (cap > 1) {
        (void)hipFree(bp);
        (void)hipFree(xp);
    }
    bp = nb;
    xp = nx;
    if (nt) {
        if (p->kticket > 1) (void)hipFree(p->d_tickets);
        p->d_tickets = nt;
        p->kticket = ncol;
    }
    cap = ncol;
    return RB_OK;
}

template <class T>
static rb_status walk(rb_plan *p, int ncol, const T *b, long ldb, T *x,
                      long ldx, hipStream_t s, const int *left = nullptr) {
    constexpr int ty = sizeof(T) == 8 ? 0 : 1;
    const int m = p->m;
    const int width = walk_width(p, ncol);
    RB_TRY(ensure_cols<T>(p, width, s));
    const T *val = ty == 0 ? (const T *)p->d_val : (const T *)p->d_val32;
    const T *invd = ty == 0 ? (const T *)p->d_invd : (const T *)p->d_invd32;
    T *bperm = ty == 0 ? (T *)p->d_bperm : (T *)p->d_bperm32;
    T *xperm = ty == 0 ? (T *)p->d_xperm : (T *)p->d_xperm32;
    void *exec = nullptr;
    const bool graph = capture_blocks<T>(p, width, &exec);
    const int pro = nblk(m > p->ntri ? m : p->ntri);
    k_prologue<T><<<dim3(pro, ncol), kT, 0, s>>>(b, ldb, p->d_levelItem, m, m,
                                                 bperm, xperm, p->d_tickets,
                                                 p->ntri, p->d_ncol, left);
    mark_start(p, s);
    if (graph)
        RB_HIP(hipGraphLaunch((hipGraphExec_t)exec, s));
    else
        launch_blocks<T>(p, val, invd, bperm, xperm, p->d_tickets, width, m,
                         p->d_ncol, s);
    mark_end(p, s);
    k_epilogue<T><<<dim3(nblk(m), ncol), kT, 0, s>>>(xperm, m, p->d_levelItem,
                                                     m, x, ldx, p->d_status, left,
                                                     p->d_rowcap);
    RB_HIP(hipGetLastError());
    return RB_OK;
}

rb_status solve_blocked(rb_plan *p, const double *d_b, int ldb, double *d_x,
                        int ldx, int nrhs, hipStream_t s) {
    if (p->m == 0) return RB_OK;
    const int cap = walk_cols(p);
    for (int k = 0; k < nrhs; k += cap)
        RB_TRY(walk<double>(p, std::min(cap, nrhs - k), d_b + (size_t)k * ldb,
                            ldb, d_x + (size_t)k * ldx, ldx, s));
    return RB_OK;
}

// The block walk on caller-provided per-solve buffers, launched kernel by
// kernel (a captured graph holds the plan's own buffers): ws holds the
// permuted b and x (m values each) and the triangular blocks' tickets;
// one right-hand side at a time.
rb_status solve_blocked_ws(rb_plan *p, const double *d_b, int ldb,
                           double *d_x, int ldx, int nrhs, hipStream_t s,
                           void *ws, size_t ws_bytes) {
    if (p->m == 0) return RB_OK;
    const int m = p->m;
    if (blocked_ws_bytes(p) > ws_bytes) {
        set_error(RB_ERR_INVALID_ARG, "workspace too small for the plan");
        return RB_ERR_INVALID_ARG;
    }
    double *bperm = (double *)ws;
    double *xperm = bperm + m;
    int *tickets = (int *)(xperm + m);
    for (int k = 0; k < nrhs; k++) {
        const int pro = nblk(m > p->ntri ? m : p->ntri);
        k_prologue<double><<<pro, kT, 0, s>>>(d_b + (size_t)k * ldb, ldb,
                                              p->d_levelItem, m, m, bperm,
                                              xperm, tickets, p->ntri, nullptr,
                                              nullptr);
        mark_start(p, s);
        launch_blocks<double>(p, p->d_val, p->d_invd, bperm, xperm, tickets,
                              1, m, p->d_ncol + 1, s);
        mark_end(p, s);
        k_epilogue<double><<<nblk(m), kT, 0, s>>>(
            xperm, m, p->d_levelItem, m, d_x + (size_t)k * ldx, ldx,
            p->d_status, nullptr, p->d_rowcap);
        RB_HIP(hipGetLastError());
    }
    return RB_OK;
}

size_t blocked_ws_bytes(const rb_plan *p) {
    return 2 * sizeof(double) * (size_t)p->m + sizeof(int) * (size_t)p->ntri;
}

// Kernel choice per block by timing.  Triangular blocks try the reference
// selector's kernel, the sync-free grid kernel and, when every in-block
// dependency is close enough for its ring, the one-workgroup kernel;
// square blocks try powers of two from a quarter to four times the rule's
// lanes per row (1 .. 64).  Each candidate runs
// twice on the plan's own buffers (b = 0) and the faster time counts; the
// block keeps the fastest candidate.  mode 1 keeps the reference
// selector's kernels and the lanes-per-row rule; modes 2 / 3 put every
// triangular block on the sync-free / one-workgroup kernel (where it
// applies).  The captured block walk is dropped so the next solve captures
// the chosen kernels.
rb_status tune_blocks(rb_plan *p, int mode, hipStream_t s, double budget_ms,
                      bool *cut) {
    if (cut) *cut = false;
    const int nb = p->nblocks, ntri = p->ntri;
    for (int i = 0; i < nb; i++) {
        Block &B = p->blocks[i];
        B.kern = kTriRef;
        B.lanes = B.kind == 1 ? lanes_per_row(B.nnzr) : 0;
        B.tuned_ms = 0.0;
        B.maxdist = 0;
    }
    destroy_graph(p);
    if (mode == 1 || p->m == 0) return RB_OK;
    // in-block dependency distances of the triangular blocks
    {
        std::vector<int> hs(ntri), hm(ntri), hd(ntri, 0);
        int mx = 1;
        for (int t = 0; t < ntri; t++) {
            hs[t] = p->blocks[2 * t].slot_base;
            hm[t] = p->blocks[2 * t].blk_m;
            mx = std::max(mx, hm[t]);
        }
        int *d = nullptr;
        RB_HIP(hipMalloc(&d, sizeof(int) * 3 * (size_t)ntri));
        rb_status st = RB_OK;
        if (hipMemcpyAsync(d, hs.data(), sizeof(int) * ntri, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(d + ntri, hm.data(), sizeof(int) * ntri, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemsetAsync(d + 2 * ntri, 0, sizeof(int) * ntri, s) != hipSuccess)
            st = RB_ERR_HIP;
        if (st == RB_OK) {
            k_tri_maxdist<<<dim3(std::min(nblk(mx), 64), ntri), kT, 0, s>>>(
                p->d_rowptr, p->d_col, d, d + ntri, d + 2 * ntri);
            if (hipMemcpyAsync(hd.data(), d + 2 * ntri, sizeof(int) * ntri,
                               hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess)
                st = RB_ERR_HIP;
        }
        (void)hipFree(d);
        if (st != RB_OK) return status_of_hip(hipGetLastError(), "tune_blocks");
        for (int t = 0; t < ntri; t++) p->blocks[2 * t].maxdist = hd[t];
    }
    auto group_ok = [&](const Block &B) {
        return B.maxdist <= kRingRows - kGroupSlack * kGroupThreads;
    };
    if (mode == 2 || mode == 3) {
        for (int t = 0; t < ntri; t++) {
            Block &B = p->blocks[2 * t];
            if (B.method == 0 || B.blk_m == 0) continue;
            if (mode == 2) B.kern = kTriSyncfree;
            else if (group_ok(B)) B.kern = kTriGroup;
        }
        return RB_OK;
    }
    // mode 0: time the candidates
    hipEvent_t e0, e1;
    RB_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return status_of_hip(hipGetLastError(), "hipEventCreate");
    }
    rb_status st = RB_OK;
    if (hipMemsetAsync(p->d_bperm, 0, sizeof(double) * (size_t)p->m, s) != hipSuccess ||
        hipMemsetAsync(p->d_xperm, 0, sizeof(double) * (size_t)p->m, s) != hipSuccess)
        st = RB_ERR_HIP;
    auto time_block = [&](int i) -> double {
        const Block &B = p->blocks[i];
        double best = 1e30;
        for (int rep = 0; rep < 2 && st == RB_OK; rep++) {
            if (B.kind == 0)
                k_block_reset<<<nblk(B.blk_m), kT, 0, s>>>(p->d_xperm + B.up, B.blk_m,
                                                           p->d_tickets + i / 2);
            if (hipEventRecord(e0, s) != hipSuccess) st = RB_ERR_HIP;
            launch_block<double>(p, i, p->d_val, p->d_invd, p->d_bperm, p->d_xperm,
                                 p->d_tickets, 1, p->m, p->d_ncol + 1, s);
            float ms = 0.f;
            if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                st = RB_ERR_HIP;
            best = std::min(best, (double)ms);
        }
        // a candidate whose bounded waits gave up is never picked (the
        // flag is cleared for the next one)
        int trip = 0;
        if (st == RB_OK &&
            hipMemcpy(&trip, p->d_status, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
            st = RB_ERR_HIP;
        if (trip != 0) {
            const int z = 0;
            if (hipMemcpy(p->d_status, &z, sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
                st = RB_ERR_HIP;
            return 2e30;
        }
        return best;
    };
    // blocks below this many entries keep the rule (their launches cost
    // more than any difference between kernels)
    constexpr int kTuneMinEntries = 1 << 14;
    double lower = 0.0;  // sum of the timed blocks' best times
    for (int i = 0; i < nb && st == RB_OK; i++) {
        if (budget_ms > 0.0 && lower > budget_ms) {
            if (cut) *cut = true;
            break;
        }
        Block &B = p->blocks[i];
        if (B.blk_m == 0 || B.blk_nnz < kTuneMinEntries) continue;
        std::vector<std::pair<int, int>> cand;  // (kern, lanes)
        if (B.kind == 0) {
            if (B.method == 0) continue;
            cand.push_back({kTriRef, 0});
            if (B.method == 2) cand.push_back({kTriSyncfree, 0});
            if (group_ok(B)) cand.push_back({kTriGroup, 0});
        } else {
            const int nr = (B.method == 1 || B.method == 3) ? B.m_new : B.blk_m;
            if (B.method < 0 || nr == 0) continue;
            // the rule's lanes per row and two powers of two either side
            const int g0 = lanes_per_row(B.nnzr);
            for (int g = std::max(1, g0 / 4); g <= std::min(64, g0 * 4); g <<= 1)
                cand.push_back({0, g});
        }
        if (cand.size() < 2) continue;
        double best = 1e30;  // a tripped candidate times 2e30
        std::pair<int, int> pick = cand[0];
        for (const auto &c : cand) {
            B.kern = c.first;
            if (B.kind == 1) B.lanes = c.second;
            const double t = time_block(i);
            if (t < best) {
                best = t;
                pick = c;
            }
        }
        B.kern = pick.first;
        if (B.kind == 1) B.lanes = pick.second;
        B.tuned_ms = best;
        if (best < 1e30) lower += best;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // a tripped watchdog during tuning would poison the plan's status
    int z = 0;
    if (st == RB_OK && hipMemcpy(p->d_status, &z, sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
        st = RB_ERR_HIP;
    if (st != RB_OK) return status_of_hip(hipGetLastError(), "tune_blocks");
    return RB_OK;
}

// Refinement scratch for ncol right-hand sides: r (FP64), r32 and d32
// (FP32), m per column, and the per-column norms, flags, counts and
// residuals; growing it frees the previous scratch.
static rb_status ensure_ir(rb_plan *p, int ncol) {
    if (ncol <= p->kir) return RB_OK;
    const size_t n = (size_t)ncol * std::max(1, p->m);
    void *r = nullptr, *r32 = nullptr, *d32 = nullptr, *meta = nullptr;
    const size_t mb = (sizeof(unsigned long long) * 2 + sizeof(int) * 2 +
                       sizeof(double)) * (size_t)ncol + sizeof(int);
    if (hipMalloc(&r, sizeof(double) * n) != hipSuccess ||
        hipMalloc(&r32, sizeof(float) * n) != hipSuccess ||
        hipMalloc(&d32, sizeof(float) * n) != hipSuccess ||
        hipMalloc(&meta, mb) != hipSuccess) {
        (void)hipGetLastError();
        for (void *q : {r, r32, d32, meta})
            if (q) (void)hipFree(q);
        set_error(RB_ERR_NO_MEMORY, "hipMalloc failed (refinement scratch)");
        return RB_ERR_NO_MEMORY;
    }
    free_ir(p);
    p->d_r = (double *)r;
    p->d_r32 = (float *)r32;
    p->d_d32 = (float *)d32;
    p->d_norms = (unsigned long long *)meta;
    p->d_irrel = (double *)(p->d_norms + 2 * (size_t)ncol);
    p->d_irdone = (int *)(p->d_irrel + ncol);
    p->d_iriters = p->d_irdone + ncol;
    p->d_irleft = p->d_iriters + ncol;
    p->kir = ncol;
    return RB_OK;
}

void free_ir(rb_plan *p) {
    if (p->kir == 0) return;
    (void)hipFree(p->d_r);
    (void)hipFree(p->d_r32);
    (void)hipFree(p->d_d32);
    (void)hipFree(p->d_norms);
    p->d_r = nullptr;
    p->d_r32 = p->d_d32 = nullptr;
    p->d_norms = nullptr;
    p->kir = 0;
}

// FP32 solves refined in FP64 (rb_solve_f32ir), all right-hand sides of a
// walk at once: x = 0, r = b; repeat { d = solve32((float) r); x += d;
// r = b - A x (FP64) } per column until max|r| <= tol * max|b| or
// max_iter corrections.  The loop runs on the device; the host reads the
// columns' flags after the second correction and each later one (one
// synchronisation for the usual two corrections), or never when the
// caller asks for no report (iters and relres both null): then max_iter
// corrections run (columns that are done skip their passes) and a column
// that missed tol sets the plan's status (rb_plan_check).
rb_status solve_f32ir(rb_plan *p, const double *d_b, int ldb, double *d_x,
                      int ldx, int nrhs, double tol, int max_iter,
                      hipStream_t s, int *iters, double *relres) {
    const int m = p->m;
    const bool report = iters || relres;
    const int cap = walk_cols(p);
    RB_TRY(ensure_ir(p, cap));
    int it_max = 0;
    double rel_max = 0.0;
    bool bad = false;
    std::vector<int> hdone(cap), hit(cap);
    std::vector<double> hrel(cap);
    for (int k0 = 0; k0 < nrhs && m > 0; k0 += cap) {
        const int nc = std::min(cap, nrhs - k0);
        const double *b = d_b + (size_t)k0 * ldb;
        double *x = d_x + (size_t)k0 * ldx;
        const dim3 g2(nblk(m), nc);
        k_ir_reset<<<nblk(nc), kT, 0, s>>>(p->d_norms, p->d_irdone,
                                           p->d_iriters, p->d_irrel, nc,
                                           p->d_irleft);
        k_ir_init<<<dim3(std::min(nblk(m), 1024), nc), kT, 0, s>>>(
            b, ldb, x, ldx, p->d_r, m, p->d_norms);
        // max |b| / max |r| of column c: d_norms[c] / d_norms[nc + c]
        for (int it = 1; it <= max_iter; it++) {
            // once every column has converged the solve is skipped (the
            // walk's prologue sets the width its kernels read to zero; the
            // fused kernels return at once)
            if (p->schedule == RB_SCHED_FUSED && p->fused) {
                // rounding of r, the solve and x += d in the fused passes
                RB_TRY(solve_fused32(p, p->d_r, m, x, ldx, nc, s, p->d_irleft,
                                     p->d_irdone));
            } else {
                k_ir_to32<<<g2, kT, 0, s>>>(p->d_r, m, p->d_r32, p->d_irdone);
                RB_TRY(walk<float>(p, nc, p->d_r32, m, p->d_d32, m, s, p->d_irleft));
                k_ir_update<<<g2, kT, 0, s>>>(p->d_d32, m, x, ldx, p->d_irdone);
            }
            residual(p, nc, b, ldb, x, ldx, s);
            k_ir_check<<<nblk(nc), kT, 0, s>>>(p->d_norms, p->d_norms + nc, nc,
                                               tol, p->d_irdone, p->d_iriters,
                                               p->d_irrel, it, p->d_irleft);
            RB_HIP(hipGetLastError());
            if (report && (it >= 2 || it == max_iter)) {
                RB_HIP(hipMemcpyAsync(hdone.data(), p->d_irdone, sizeof(int) * nc,
                                      hipMemcpyDeviceToHost, s));
                RB_HIP(hipStreamSynchronize(s));
                bool all = true;
                for (int c = 0; c < nc; c++) all = all && hdone[c];
                if (all) break;
            }
        }
        if (report) {
            RB_HIP(hipMemcpyAsync(hdone.data(), p->d_irdone, sizeof(int) * nc,
                                  hipMemcpyDeviceToHost, s));
            RB_HIP(hipMemcpyAsync(hit.data(), p->d_iriters, sizeof(int) * nc,
                                  hipMemcpyDeviceToHost, s));
            RB_HIP(hipMemcpyAsync(hrel.data(), p->d_irrel, sizeof(double) * nc,
                                  hipMemcpyDeviceToHost, s));
            RB_HIP(hipStreamSynchronize(s));
            for (int c = 0; c < nc; c++) {
                if (!hdone[c]) bad = true;
                it_max = std::max(it_max, hit[c]);
                if (hrel[c] != hrel[c] || hrel[c] > rel_max) rel_max = hrel[c];
            }
        } else {
            k_ir_flag<<<nblk(nc), kT, 0, s>>>(p->d_irdone, nc, p->d_status);
            RB_HIP(hipGetLastError());
        }
    }
    if (iters) *iters = it_max;
    if (relres) *relres = rel_max;
    if (bad) {
        set_error(RB_ERR_NOT_CONVERGED,
                  "FP32 refinement did not reach the requested residual");
        return RB_ERR_NOT_CONVERGED;
    }
    return RB_OK;
}

}  // namespace rb
