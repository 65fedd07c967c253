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
//   method 0/2 CSR thread / wave per row, 1/3 DCSR thread / wave per
//   non-empty row (R/utils_spmv_cuda.h:60-203), plus a workgroup per row
//   longer than 2048 entries (the long-row pass :30-58).
// Every sum runs in a fixed order, so a solve is bitwise reproducible.
#include "rb_internal.h"

namespace rb {

namespace {

constexpr int kT = 256;
constexpr unsigned kSpinLimit = 1u << 22;  // about seconds of polling

__device__ inline double ld_x(const double *p) {
    return __longlong_as_double((long long)__hip_atomic_load(
        (const unsigned long long *)p, __ATOMIC_RELAXED,
        __HIP_MEMORY_SCOPE_AGENT));
}

__device__ inline void st_x(double *p, double v) {
    __hip_atomic_store((unsigned long long *)p,
                       (unsigned long long)__double_as_longlong(v),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_prologue(const double *__restrict__ b,
                           const int *__restrict__ lf, int m, double *bperm,
                           double *xperm, int *tickets, int ntri) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) {
        bperm[i] = b[lf[i]];
        ((unsigned long long *)xperm)[i] = kPending;
    }
    if (i < ntri) tickets[i] = 0;
}

__global__ void k_epilogue(const double *__restrict__ xperm,
                           const int *__restrict__ lf, int m, double *x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) x[lf[i]] = xperm[i];
}

__global__ void k_tri_diag(const double *__restrict__ invd,
                           const double *__restrict__ bperm, int up, int n,
                           double *xperm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) st_x(&xperm[up + i], bperm[up + i] * invd[up + i]);
}

// level-set, thread per row; rows of one local level
__global__ void k_tri_level_thread(const int *__restrict__ rowptr,
                                   const int *__restrict__ col,
                                   const double *__restrict__ val,
                                   const double *__restrict__ invd,
                                   const double *__restrict__ bperm,
                                   const int *__restrict__ rows, int nrows,
                                   int sb, int up, double *xperm) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nrows) return;
    const int r = rows[k];
    const int s = rowptr[sb + r], e = rowptr[sb + r + 1];
    double acc = 0.0;
    for (int j = s; j < e; j++) acc += val[j] * xperm[up + col[j]];
    st_x(&xperm[up + r], (bperm[up + r] - acc) * invd[up + r]);
}

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// level-set, wave per row
__global__ void k_tri_level_wave(const int *__restrict__ rowptr,
                                 const int *__restrict__ col,
                                 const double *__restrict__ val,
                                 const double *__restrict__ invd,
                                 const double *__restrict__ bperm,
                                 const int *__restrict__ rows, int nrows,
                                 int sb, int up, double *xperm) {
    const int k = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (k >= nrows) return;
    const int r = rows[k];
    const int s = rowptr[sb + r], e = rowptr[sb + r + 1];
    double acc = 0.0;
    for (int j = s + lane; j < e; j += 64) acc += val[j] * xperm[up + col[j]];
    acc = wave_sum(acc);
    if (lane == 0) st_x(&xperm[up + r], (bperm[up + r] - acc) * invd[up + r]);
}

// sync-free CSR pull, lane per row, 64 consecutive rows per wave ticket
__global__ void __launch_bounds__(kT)
k_tri_syncfree(const int *__restrict__ rowptr, const int *__restrict__ col,
               const double *__restrict__ val,
               const double *__restrict__ invd,
               const double *__restrict__ bperm, int sb, int up, int bm,
               int subst, int *ticket, double *xperm, int *status) {
    const int lane = threadIdx.x & 63;
    int tile = 0;
    if (lane == 0) tile = atomicAdd(ticket, 1);
    tile = __shfl(tile, 0, 64);
    const int pos = tile * 64 + lane;
    const bool valid = pos < bm;
    const int r = subst == 0 ? pos : bm - 1 - pos;
    int j = 0, e = 0;
    double bv = 0.0, iv = 0.0;
    if (valid) {
        j = rowptr[sb + r];
        e = rowptr[sb + r + 1];
        bv = bperm[up + r];
        iv = invd[up + r];
    }
    double acc = 0.0;
    bool pend = valid;
    const double *xb = xperm + up;
    unsigned spins = 0;
    while (__any(pend)) {
        if (pend) {
            while (j < e) {
                const double xv = ld_x(&xb[col[j]]);
                if (__double_as_longlong(xv) == (long long)kPending) break;
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
            if (++spins > kSpinLimit) {
                if (pend) {
                    atomicOr(status, 1);
                    st_x(&xperm[up + r], __longlong_as_double(0x7FF8000000000000ll));
                    pend = false;
                }
            } else {
                __builtin_amdgcn_s_sleep(1);
            }
        }
    }
}

// square block SpMV, CSR, thread per row; b[up + r] -= A[r,:] x[left:]
__global__ void k_sq_csr_thread(const int *__restrict__ rowptr,
                                const int *__restrict__ col,
                                const double *__restrict__ val,
                                const double *__restrict__ xperm, int sb,
                                int up, int left, int bm, double *bperm) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= bm) return;
    const int s = rowptr[sb + r], e = rowptr[sb + r + 1];
    if (s == e || e - s > kLongRow) return;
    double acc = 0.0;
    for (int j = s; j < e; j++) acc += val[j] * xperm[left + col[j]];
    bperm[up + r] -= acc;
}

__global__ void k_sq_csr_wave(const int *__restrict__ rowptr,
                              const int *__restrict__ col,
                              const double *__restrict__ val,
                              const double *__restrict__ xperm, int sb,
                              int up, int left, int bm, double *bperm) {
    const int r = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (r >= bm) return;
    const int s = rowptr[sb + r], e = rowptr[sb + r + 1];
    if (s == e || e - s > kLongRow) return;
    double acc = 0.0;
    for (int j = s + lane; j < e; j += 64) acc += val[j] * xperm[left + col[j]];
    acc = wave_sum(acc);
    if (lane == 0) bperm[up + r] -= acc;
}

__global__ void k_sq_dcsr_thread(const int *__restrict__ rowptr,
                                 const int *__restrict__ col,
                                 const double *__restrict__ val,
                                 const double *__restrict__ xperm,
                                 const int *__restrict__ rows, int nrows,
                                 int sb, int up, int left, double *bperm) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nrows) return;
    const int r = rows[k];
    const int s = rowptr[sb + r], e = rowptr[sb + r + 1];
    if (e - s > kLongRow) return;
    double acc = 0.0;
    for (int j = s; j < e; j++) acc += val[j] * xperm[left + col[j]];
    bperm[up + r] -= acc;
}

__global__ void k_sq_dcsr_wave(const int *__restrict__ rowptr,
                               const int *__restrict__ col,
                               const double *__restrict__ val,
                               const double *__restrict__ xperm,
                               const int *__restrict__ rows, int nrows,
                               int sb, int up, int left, double *bperm) {
    const int k = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (k >= nrows) return;
    const int r = rows[k];
    const int s = rowptr[sb + r], e = rowptr[sb + r + 1];
    if (e - s > kLongRow) return;
    double acc = 0.0;
    for (int j = s + lane; j < e; j += 64) acc += val[j] * xperm[left + col[j]];
    acc = wave_sum(acc);
    if (lane == 0) bperm[up + r] -= acc;
}

// one workgroup per long row (> 2048 entries), fixed-order tree reduction
__global__ void __launch_bounds__(kT)
k_sq_long(const int *__restrict__ rowptr, const int *__restrict__ col,
          const double *__restrict__ val, const double *__restrict__ xperm,
          const int *__restrict__ rows, int sb, int up, int left,
          double *bperm) {
    const int r = rows[blockIdx.x];
    const int s = rowptr[sb + r], e = rowptr[sb + r + 1];
    double acc = 0.0;
    for (int j = s + threadIdx.x; j < e; j += kT)
        acc += val[j] * xperm[left + col[j]];
    acc = wave_sum(acc);
    __shared__ double sh[kT / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kT / 64; w++) t += sh[w];
        bperm[up + r] -= t;
    }
}

inline int nblk(long n, int t = kT) { return (int)((n + t - 1) / t); }

}  // namespace

// The block walk of one solve: kernels on plan-owned buffers only, so it
// can be captured once into a graph and replayed.
static void launch_blocks(rb_plan *p, hipStream_t s) {
        for (int i = 0; i < p->nblocks; i++) {
            const Block &B = p->blocks[i];
            if (B.blk_m == 0) continue;
            if (B.kind == 0) {
                switch (B.method) {
                case 0:
                    k_tri_diag<<<nblk(B.blk_m), kT, 0, s>>>(
                        p->d_invd, p->d_bperm, B.up, B.blk_m, p->d_xperm);
                    break;
                case 2:
                    for (int L = 0; L < B.nlv; L++) {
                        const int a = B.lvptr_h[L], c = B.lvptr_h[L + 1];
                        const int nr = c - a;
                        if (nr <= 0) continue;
                        const int *rows = p->d_lvitem + B.lv_base + a;
                        if (L == 0 || B.nnzlv_h[L] / nr <= 15)
                            k_tri_level_thread<<<nblk(nr), kT, 0, s>>>(
                                p->d_rowptr, p->d_col, p->d_val, p->d_invd,
                                p->d_bperm, rows, nr, B.slot_base, B.up,
                                p->d_xperm);
                        else
                            k_tri_level_wave<<<nblk((long)nr * 64), kT, 0, s>>>(
                                p->d_rowptr, p->d_col, p->d_val, p->d_invd,
                                p->d_bperm, rows, nr, B.slot_base, B.up,
                                p->d_xperm);
                    }
                    break;
                default:  // 1 and 3
                    k_tri_syncfree<<<nblk((long)B.blk_m, 64), 64, 0, s>>>(
                        p->d_rowptr, p->d_col, p->d_val, p->d_invd,
                        p->d_bperm, B.slot_base, B.up, B.blk_m, p->subst,
                        p->d_tickets + i / 2, p->d_xperm, p->d_status);
                    break;
                }
            } else {
                const int *rows = p->d_dcsr_rows + B.dcsr_base;
                switch (B.method) {
                case -1:
                    break;
                case 0:
                    k_sq_csr_thread<<<nblk(B.blk_m), kT, 0, s>>>(
                        p->d_rowptr, p->d_col, p->d_val, p->d_xperm,
                        B.slot_base, B.up, B.left, B.blk_m, p->d_bperm);
                    break;
                case 2:
                    k_sq_csr_wave<<<nblk((long)B.blk_m * 64), kT, 0, s>>>(
                        p->d_rowptr, p->d_col, p->d_val, p->d_xperm,
                        B.slot_base, B.up, B.left, B.blk_m, p->d_bperm);
                    break;
                case 1:
                    k_sq_dcsr_thread<<<nblk(B.m_new), kT, 0, s>>>(
                        p->d_rowptr, p->d_col, p->d_val, p->d_xperm, rows,
                        B.m_new, B.slot_base, B.up, B.left, p->d_bperm);
                    break;
                default:
                    k_sq_dcsr_wave<<<nblk((long)B.m_new * 64), kT, 0, s>>>(
                        p->d_rowptr, p->d_col, p->d_val, p->d_xperm, rows,
                        B.m_new, B.slot_base, B.up, B.left, p->d_bperm);
                    break;
                }
                if (B.method >= 0 && B.longrow > 0)
                    k_sq_long<<<B.longrow, kT, 0, s>>>(
                        p->d_rowptr, p->d_col, p->d_val, p->d_xperm,
                        p->d_long_rows + B.long_base, B.slot_base, B.up,
                        B.left, p->d_bperm);
            }
        }
}

// Capture the block walk into a graph (once per plan).  Returns false when
// capture is not possible; the walk is then launched kernel by kernel.
static bool capture_blocks(rb_plan *p) {
    if (p->graph_failed) return false;
    if (p->graph_exec) return true;
    hipStream_t cs = nullptr;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    bool ok = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) == hipSuccess;
    if (ok) ok = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (ok) {
        launch_blocks(p, cs);
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
    p->graph_exec = ge;
    return true;
}

void destroy_graph(rb_plan *p) {
    if (p->graph_exec) (void)hipGraphExecDestroy((hipGraphExec_t)p->graph_exec);
    p->graph_exec = nullptr;
}

rb_status solve_blocked(rb_plan *p, const double *d_b, int ldb, double *d_x,
                        int ldx, int nrhs, hipStream_t s) {
    const int m = p->m;
    if (m == 0) return RB_OK;
    const bool graph = capture_blocks(p);
    for (int k = 0; k < nrhs; k++) {
        const double *b = d_b + (size_t)k * ldb;
        double *x = d_x + (size_t)k * ldx;
        const int pro = nblk(m > p->ntri ? m : p->ntri);
        k_prologue<<<pro, kT, 0, s>>>(b, p->d_levelItem, m, p->d_bperm,
                                      p->d_xperm, p->d_tickets, p->ntri);
        mark_start(p, s);
        if (graph)
            RB_HIP(hipGraphLaunch((hipGraphExec_t)p->graph_exec, s));
        else
            launch_blocks(p, s);
        mark_end(p, s);
        k_epilogue<<<nblk(m), kT, 0, s>>>(p->d_xperm, p->d_levelItem, m, x);
        RB_HIP(hipGetLastError());
    }
    return RB_OK;
}

}  // namespace rb
