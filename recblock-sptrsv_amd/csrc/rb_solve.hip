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
// Every sum runs in a fixed order, so a solve is bitwise reproducible.
//
// The kernels are templates on the value type: the FP64 walk is the plain
// solve; the FP32 walk (FP32 copies of the slot values) is the inner solver
// of the FP32 + FP64 iterative refinement at the end of this file.
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

template <class T>
__global__ void k_prologue(const T *__restrict__ b, const int *__restrict__ lf,
                           int m, T *bperm, T *xperm, int *tickets, int ntri) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) {
        bperm[i] = b[lf[i]];
        ((typename Bits<T>::U *)xperm)[i] = Bits<T>::kPend;
    }
    if (i < ntri) tickets[i] = 0;
}

// x back to the caller's order; after a watchdog trip (status set) every
// entry is a quiet NaN, so no solve of a tripped plan hands out a finite x
template <class T>
__global__ void k_epilogue(const T *__restrict__ xperm,
                           const int *__restrict__ lf, int m, T *x,
                           const int *status) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m)
        x[lf[i]] = *status ? __builtin_bit_cast(T, Bits<T>::kQNaN) : xperm[i];
}

template <class T>
__global__ void k_tri_diag(const T *__restrict__ invd,
                           const T *__restrict__ bperm, int up, int n,
                           T *xperm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) st_x(&xperm[up + i], bperm[up + i] * invd[up + i]);
}

// level-set, thread per row; rows of one local level
template <class T>
__global__ void k_tri_level_thread(const int *__restrict__ rowptr,
                                   const int *__restrict__ col,
                                   const T *__restrict__ val,
                                   const T *__restrict__ invd,
                                   const T *__restrict__ bperm,
                                   const int *__restrict__ rows, int nrows,
                                   int sb, int up, T *xperm) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nrows) return;
    const int r = rows[k];
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
                                 int sb, int up, T *xperm) {
    const int k = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (k >= nrows) return;
    const int r = rows[k];
    const int s = rowptr[sb + r], e = rowptr[sb + r + 1];
    T acc = 0;
    for (int j = s + lane; j < e; j += 64) acc += val[j] * xperm[up + col[j]];
    acc = wave_sum(acc);
    if (lane == 0) st_x(&xperm[up + r], (bperm[up + r] - acc) * invd[up + r]);
}

// sync-free CSR pull, lane per row, 64 consecutive rows per wave ticket
template <class T>
__global__ void __launch_bounds__(kT)
k_tri_syncfree(const int *__restrict__ rowptr, const int *__restrict__ col,
               const T *__restrict__ val, const T *__restrict__ invd,
               const T *__restrict__ bperm, int sb, int up, int bm,
               int subst, int *ticket, T *xperm, int *status,
               unsigned spin_limit) {
    const int lane = threadIdx.x & 63;
    int tile = 0;
    if (lane == 0) tile = atomicAdd(ticket, 1);
    tile = __shfl(tile, 0, 64);
    const int pos = tile * 64 + lane;
    const bool valid = pos < bm;
    const int r = subst == 0 ? pos : bm - 1 - pos;
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
// short levels) run faster.  Correct when every in-block dependency is at
// most kRingRows - kGroupThreads positions back: all rows in flight lie
// within 1024 positions of the oldest unsolved one, so a slot is reused
// only after every row that reads it has finished (the host checks the
// distance).  Lanes of one wave may depend on each other, so a wave loops
// until all its lanes are done, each lane advancing as far as its
// dependencies allow; a row's entries are read ahead in chunks of kChunk,
// before the waits, and summed in order, as by the other kernels.
// (A level-synchronous form -- one workgroup barrier per level, rows read
// 1024 positions ahead -- took ~1.1 us per level on the banded factor
// against this form's ~0.9 us: a wave waits for its newest loads as a
// whole, so the reads ahead for some lanes held the others at every
// level.)
template <class T>
__global__ void __launch_bounds__(kGroupThreads)
k_tri_group(const int *__restrict__ rowptr, const int *__restrict__ col,
            const T *__restrict__ val, const T *__restrict__ invd,
            const T *__restrict__ bperm, int sb, int up, int bm, int subst,
            T *xperm, int *status, unsigned spin_limit) {
    constexpr int kChunk = 8;
    __shared__ int tag[kRingRows];
    __shared__ T xr[kRingRows];
    for (int i = threadIdx.x; i < kRingRows; i += kGroupThreads) tag[i] = -1;
    __syncthreads();
    bool bad = false;
    for (int p0 = threadIdx.x & ~63; p0 < bm; p0 += kGroupThreads) {
        const int pos = p0 + (threadIdx.x & 63);
        const bool valid = pos < bm;
        const int r = subst == 0 ? pos : bm - 1 - pos;
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
    }
    if (bad) atomicOr(status, 1);
}

// square block SpMV, CSR, thread per row; b[up + r] -= A[r,:] x[left:]
template <class T>
__global__ void k_sq_csr_thread(const int *__restrict__ rowptr,
                                const int *__restrict__ col,
                                const T *__restrict__ val,
                                const T *__restrict__ xperm, int sb,
                                int up, int left, int bm, T *bperm) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= bm) return;
    const int s = rowptr[sb + r], e = rowptr[sb + r + 1];
    if (s == e || e - s > kLongRow) return;
    T acc = 0;
    for (int j = s; j < e; j++) acc += val[j] * xperm[left + col[j]];
    bperm[up + r] -= acc;
}

template <class T>
__global__ void k_sq_csr_wave(const int *__restrict__ rowptr,
                              const int *__restrict__ col,
                              const T *__restrict__ val,
                              const T *__restrict__ xperm, int sb,
                              int up, int left, int bm, T *bperm) {
    const int r = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (r >= bm) return;
    const int s = rowptr[sb + r], e = rowptr[sb + r + 1];
    if (s == e || e - s > kLongRow) return;
    T acc = 0;
    for (int j = s + lane; j < e; j += 64) acc += val[j] * xperm[left + col[j]];
    acc = wave_sum(acc);
    if (lane == 0) bperm[up + r] -= acc;
}

template <class T>
__global__ void k_sq_dcsr_thread(const int *__restrict__ rowptr,
                                 const int *__restrict__ col,
                                 const T *__restrict__ val,
                                 const T *__restrict__ xperm,
                                 const int *__restrict__ rows, int nrows,
                                 int sb, int up, int left, T *bperm) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nrows) return;
    const int r = rows[k];
    const int s = rowptr[sb + r], e = rowptr[sb + r + 1];
    if (e - s > kLongRow) return;
    T acc = 0;
    for (int j = s; j < e; j++) acc += val[j] * xperm[left + col[j]];
    bperm[up + r] -= acc;
}

template <class T>
__global__ void k_sq_dcsr_wave(const int *__restrict__ rowptr,
                               const int *__restrict__ col,
                               const T *__restrict__ val,
                               const T *__restrict__ xperm,
                               const int *__restrict__ rows, int nrows,
                               int sb, int up, int left, T *bperm) {
    const int k = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (k >= nrows) return;
    const int r = rows[k];
    const int s = rowptr[sb + r], e = rowptr[sb + r + 1];
    if (e - s > kLongRow) return;
    T acc = 0;
    for (int j = s + lane; j < e; j += 64) acc += val[j] * xperm[left + col[j]];
    acc = wave_sum(acc);
    if (lane == 0) bperm[up + r] -= acc;
}

// square block SpMV with G lanes per row (G a power of two, 2..32): the
// group reads its row's entries coalesced, then a fixed-order butterfly;
// rows from the DCSR list (non-empty rows) or all rows of the block
template <class T, int G, bool DCSR>
__global__ void __launch_bounds__(kT)
k_sq_group(const int *__restrict__ rowptr, const int *__restrict__ col,
           const T *__restrict__ val, const T *__restrict__ xperm,
           const int *__restrict__ rows, int nrows, int sb, int up, int left,
           T *bperm) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = t / G, g = t & (G - 1);
    if (k >= nrows) return;  // whole groups leave together
    const int r = DCSR ? rows[k] : k;
    const int s = rowptr[sb + r], e = rowptr[sb + r + 1];
    if (s == e || e - s > kLongRow) return;
    T acc = 0;
    for (int j = s + g; j < e; j += G) acc += val[j] * xperm[left + col[j]];
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, G);
    if (g == 0) bperm[up + r] -= acc;
}

// one workgroup per long row (> 2048 entries), fixed-order tree reduction
template <class T>
__global__ void __launch_bounds__(kT)
k_sq_long(const int *__restrict__ rowptr, const int *__restrict__ col,
          const T *__restrict__ val, const T *__restrict__ xperm,
          const int *__restrict__ rows, int sb, int up, int left, T *bperm) {
    const int r = rows[blockIdx.x];
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

template <class T, bool DCSR>
void launch_sq(int g, int nrows, hipStream_t s, const int *rowptr,
               const int *col, const T *val, const T *xperm, const int *rows,
               int sb, int up, int left, T *bperm) {
    const long th = (long)nrows * g;
    switch (g) {
    case 2: k_sq_group<T, 2, DCSR><<<nblk(th), kT, 0, s>>>(rowptr, col, val, xperm, rows, nrows, sb, up, left, bperm); break;
    case 4: k_sq_group<T, 4, DCSR><<<nblk(th), kT, 0, s>>>(rowptr, col, val, xperm, rows, nrows, sb, up, left, bperm); break;
    case 8: k_sq_group<T, 8, DCSR><<<nblk(th), kT, 0, s>>>(rowptr, col, val, xperm, rows, nrows, sb, up, left, bperm); break;
    case 16: k_sq_group<T, 16, DCSR><<<nblk(th), kT, 0, s>>>(rowptr, col, val, xperm, rows, nrows, sb, up, left, bperm); break;
    default: k_sq_group<T, 32, DCSR><<<nblk(th), kT, 0, s>>>(rowptr, col, val, xperm, rows, nrows, sb, up, left, bperm); break;
    }
}

// ---------------------------------------------------- refinement kernels
__global__ void k_to_f32(const double *__restrict__ a, int m, float *o) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) o[i] = (float)a[i];
}

// x += d (FP64 update with the FP32 correction)
__global__ void k_add_f32(const float *__restrict__ d, int m, double *x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) x[i] += (double)d[i];
}

// max |v| over the vector, as the bits of a non-negative double (ordered
// like the values) into *out (cleared by the caller)
__global__ void k_absmax(const double *__restrict__ v, int m,
                         unsigned long long *out) {
    double mx = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += gridDim.x * blockDim.x)
        mx = fmax(mx, fabs(v[i]));
    mx = wave_max(mx);
    if ((threadIdx.x & 63) == 0)
        atomicMax(out, (unsigned long long)__double_as_longlong(mx));
}

// r = b - A x in the original ordering, FP64: row i of the level-ordered
// CSR is original row lf[i], its columns hold original ids; also max |r|
__global__ void k_residual(const int *__restrict__ rptr,
                           const int *__restrict__ rcol,
                           const double *__restrict__ rval,
                           const int *__restrict__ lf, int m,
                           const double *__restrict__ b,
                           const double *__restrict__ x, double *r,
                           unsigned long long *rmax) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0.0;
    if (i < m) {
        const int o = lf[i];
        double acc = 0.0;
        for (int k = rptr[i]; k < rptr[i + 1]; k++) acc += rval[k] * x[rcol[k]];
        const double v = b[o] - acc;
        r[o] = v;
        a = fabs(v);
    }
    a = wave_max(a);
    if ((threadIdx.x & 63) == 0)
        atomicMax(rmax, (unsigned long long)__double_as_longlong(a));
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
__global__ void k_block_reset(double *x, int n, int *ticket) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ((unsigned long long *)x)[i] = kPending;
    if (i == 0) *ticket = 0;
}

}  // namespace

// One block of the walk in value type T (the kernels of B.kern / B.lanes).
template <class T>
static void launch_block(rb_plan *p, int i, const T *val, const T *invd,
                         T *bperm, T *xperm, int *tickets, hipStream_t s) {
    const Block &B = p->blocks[i];
    if (B.blk_m == 0) return;
    if (B.kind == 0) {
        if (B.method == 0) {
            k_tri_diag<T><<<nblk(B.blk_m), kT, 0, s>>>(invd, bperm, B.up,
                                                       B.blk_m, xperm);
        } else if (B.kern == kTriGroup) {
            k_tri_group<T><<<1, kGroupThreads, 0, s>>>(
                p->d_rowptr, p->d_col, val, invd, bperm, B.slot_base, B.up,
                B.blk_m, p->subst, xperm, p->d_status, p->spin_limit);
        } else if (B.kern == kTriRef && B.method == 2) {
            for (int L = 0; L < B.nlv; L++) {
                const int a = B.lvptr_h[L], c = B.lvptr_h[L + 1];
                const int nr = c - a;
                if (nr <= 0) continue;
                const int *rows = p->d_lvitem + B.lv_base + a;
                if (L == 0 || B.nnzlv_h[L] / nr <= 15)
                    k_tri_level_thread<T><<<nblk(nr), kT, 0, s>>>(
                        p->d_rowptr, p->d_col, val, invd, bperm, rows, nr,
                        B.slot_base, B.up, xperm);
                else
                    k_tri_level_wave<T><<<nblk((long)nr * 64), kT, 0, s>>>(
                        p->d_rowptr, p->d_col, val, invd, bperm, rows, nr,
                        B.slot_base, B.up, xperm);
            }
        } else {  // methods 1 and 3, or kTriSyncfree
            k_tri_syncfree<T><<<nblk((long)B.blk_m, 64), 64, 0, s>>>(
                p->d_rowptr, p->d_col, val, invd, bperm, B.slot_base,
                B.up, B.blk_m, p->subst, tickets + i / 2, xperm,
                p->d_status, p->spin_limit);
        }
        return;
    }
    // The reference's method picks CSR (0, 2) or DCSR (1, 3) rows; lanes
    // per row are the block's own (the reference's thread / warp split is
    // 1 or a whole wave).
    const int *rows = p->d_dcsr_rows + B.dcsr_base;
    const bool dcsr = B.method == 1 || B.method == 3;
    const int nr = dcsr ? B.m_new : B.blk_m;
    const int g = B.lanes;
    if (B.method < 0 || nr == 0) return;  // empty block
    if (g == 1) {
        if (dcsr)
            k_sq_dcsr_thread<T><<<nblk(nr), kT, 0, s>>>(
                p->d_rowptr, p->d_col, val, xperm, rows, nr, B.slot_base,
                B.up, B.left, bperm);
        else
            k_sq_csr_thread<T><<<nblk(nr), kT, 0, s>>>(
                p->d_rowptr, p->d_col, val, xperm, B.slot_base, B.up, B.left,
                nr, bperm);
    } else if (g == 64) {
        if (dcsr)
            k_sq_dcsr_wave<T><<<nblk((long)nr * 64), kT, 0, s>>>(
                p->d_rowptr, p->d_col, val, xperm, rows, nr, B.slot_base,
                B.up, B.left, bperm);
        else
            k_sq_csr_wave<T><<<nblk((long)nr * 64), kT, 0, s>>>(
                p->d_rowptr, p->d_col, val, xperm, B.slot_base, B.up, B.left,
                nr, bperm);
    } else if (dcsr) {
        launch_sq<T, true>(g, nr, s, p->d_rowptr, p->d_col, val, xperm, rows,
                           B.slot_base, B.up, B.left, bperm);
    } else {
        launch_sq<T, false>(g, nr, s, p->d_rowptr, p->d_col, val, xperm,
                            rows, B.slot_base, B.up, B.left, bperm);
    }
    if (B.longrow > 0)
        k_sq_long<T><<<B.longrow, kT, 0, s>>>(
            p->d_rowptr, p->d_col, val, xperm, p->d_long_rows + B.long_base,
            B.slot_base, B.up, B.left, bperm);
}

// The block walk of one solve in value type T: kernels on plan-owned
// buffers only, so it can be captured once into a graph and replayed.
template <class T>
static void launch_blocks(rb_plan *p, const T *val, const T *invd, T *bperm,
                          T *xperm, int *tickets, hipStream_t s) {
    for (int i = 0; i < p->nblocks; i++)
        launch_block<T>(p, i, val, invd, bperm, xperm, tickets, s);
}

// Capture a block walk into a graph (once per plan and value type).
// Returns false when capture is not possible; the walk is then launched
// kernel by kernel.
template <class T>
static bool capture_blocks(rb_plan *p, void **exec, const T *val,
                           const T *invd, T *bperm, T *xperm) {
    if (p->graph_failed) return false;
    if (*exec) return true;
    hipStream_t cs = nullptr;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    bool ok = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) == hipSuccess;
    if (ok) ok = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (ok) {
        launch_blocks<T>(p, val, invd, bperm, xperm, p->d_tickets, cs);
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
    *exec = ge;
    return true;
}

void destroy_graph(rb_plan *p) {
    if (p->graph_exec) (void)hipGraphExecDestroy((hipGraphExec_t)p->graph_exec);
    if (p->graph_exec32) (void)hipGraphExecDestroy((hipGraphExec_t)p->graph_exec32);
    p->graph_exec = nullptr;
    p->graph_exec32 = nullptr;
}

template <class T>
static rb_status walk(rb_plan *p, void **exec, const T *val, const T *invd,
                      T *bperm, T *xperm, const T *b, T *x, hipStream_t s) {
    const int m = p->m;
    const bool graph = capture_blocks<T>(p, exec, val, invd, bperm, xperm);
    const int pro = nblk(m > p->ntri ? m : p->ntri);
    k_prologue<T><<<pro, kT, 0, s>>>(b, p->d_levelItem, m, bperm, xperm,
                                     p->d_tickets, p->ntri);
    mark_start(p, s);
    if (graph)
        RB_HIP(hipGraphLaunch((hipGraphExec_t)*exec, s));
    else
        launch_blocks<T>(p, val, invd, bperm, xperm, p->d_tickets, s);
    mark_end(p, s);
    k_epilogue<T><<<nblk(m), kT, 0, s>>>(xperm, p->d_levelItem, m, x,
                                         p->d_status);
    RB_HIP(hipGetLastError());
    return RB_OK;
}

rb_status solve_blocked(rb_plan *p, const double *d_b, int ldb, double *d_x,
                        int ldx, int nrhs, hipStream_t s) {
    if (p->m == 0) return RB_OK;
    for (int k = 0; k < nrhs; k++)
        RB_TRY(walk<double>(p, &p->graph_exec, p->d_val, p->d_invd,
                            p->d_bperm, p->d_xperm, d_b + (size_t)k * ldb,
                            d_x + (size_t)k * ldx, s));
    return RB_OK;
}

// The block walk on caller-provided per-solve buffers, launched kernel by
// kernel (a captured graph holds the plan's own buffers): ws holds the
// permuted b and x (m values each) and the triangular blocks' tickets.
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
        k_prologue<double><<<pro, kT, 0, s>>>(d_b + (size_t)k * ldb,
                                              p->d_levelItem, m, bperm, xperm,
                                              tickets, p->ntri);
        mark_start(p, s);
        launch_blocks<double>(p, p->d_val, p->d_invd, bperm, xperm, tickets,
                              s);
        mark_end(p, s);
        k_epilogue<double><<<nblk(m), kT, 0, s>>>(
            xperm, p->d_levelItem, m, d_x + (size_t)k * ldx, p->d_status);
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
rb_status tune_blocks(rb_plan *p, int mode, hipStream_t s) {
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
        return B.maxdist <= kRingRows - kGroupThreads;
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
                                 p->d_tickets, s);
            float ms = 0.f;
            if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                st = RB_ERR_HIP;
            best = std::min(best, (double)ms);
        }
        return best;
    };
    // blocks below this many entries keep the rule (their launches cost
    // more than any difference between kernels)
    constexpr int kTuneMinEntries = 1 << 14;
    for (int i = 0; i < nb && st == RB_OK; i++) {
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
        double best = 1e30;
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

// FP32 solves refined in FP64 (rb_solve_f32ir)// FP32 solves refined in FP64 (rb_solve_f32ir): per right-hand side,
// x = 0, r = b; repeat { d = solve32((float) r); x += d; r = b - A x (FP64) }
// until max|r| <= tol * max|b| or max_iter corrections.
rb_status solve_f32ir(rb_plan *p, const double *d_b, int ldb, double *d_x,
                      int ldx, int nrhs, double tol, int max_iter,
                      hipStream_t s, int *iters, double *relres) {
    const int m = p->m;
    int it_max = 0;
    double rel_max = 0.0;
    bool bad = false;
    for (int k = 0; k < nrhs && m > 0; k++) {
        const double *b = d_b + (size_t)k * ldb;
        double *x = d_x + (size_t)k * ldx;
        unsigned long long *nrm = p->d_norms;
        RB_HIP(hipMemsetAsync(nrm, 0, 2 * sizeof(unsigned long long), s));
        k_absmax<<<std::min(nblk(m), 1024), kT, 0, s>>>(b, m, nrm);
        RB_HIP(hipMemsetAsync(x, 0, sizeof(double) * (size_t)m, s));
        RB_HIP(hipMemcpyAsync(p->d_r, b, sizeof(double) * (size_t)m,
                              hipMemcpyDeviceToDevice, s));
        unsigned long long hb = 0;
        RB_HIP(hipMemcpyAsync(&hb, nrm, sizeof(hb), hipMemcpyDeviceToHost, s));
        RB_HIP(hipStreamSynchronize(s));
        const double bn = __builtin_bit_cast(double, hb);
        double rel = 0.0;
        int it = 0;
        while (bn > 0.0 && it < max_iter) {
            it++;
            k_to_f32<<<nblk(m), kT, 0, s>>>(p->d_r, m, p->d_r32);
            RB_TRY(walk<float>(p, &p->graph_exec32, p->d_val32, p->d_invd32,
                               p->d_bperm32, p->d_xperm32, p->d_r32, p->d_d32,
                               s));
            k_add_f32<<<nblk(m), kT, 0, s>>>(p->d_d32, m, x);
            RB_HIP(hipMemsetAsync(nrm + 1, 0, sizeof(unsigned long long), s));
            k_residual<<<nblk(m), kT, 0, s>>>(p->d_rptr, p->d_rcol, p->d_rval,
                                              p->d_levelItem, m, b, x, p->d_r,
                                              nrm + 1);
            RB_HIP(hipGetLastError());
            unsigned long long hr = 0;
            RB_HIP(hipMemcpyAsync(&hr, nrm + 1, sizeof(hr),
                                  hipMemcpyDeviceToHost, s));
            RB_HIP(hipStreamSynchronize(s));
            rel = __builtin_bit_cast(double, hr) / bn;
            if (rel <= tol) break;
        }
        if (!(rel <= tol)) bad = true;
        it_max = std::max(it_max, it);
        if (rel != rel || rel > rel_max) rel_max = rel;
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
