// Batched solves of many independent factors of uneven size.
//
// The factors are laid out block-diagonally as one matrix (factor f owns
// rows and columns [off_f, off_f + m_f)) and preprocessed as one plan: the
// level analysis of a block-diagonal matrix is the union of the factors'
// level sets, so one persistent fused launch (or one captured block walk)
// advances every factor at once, and the fused schedule's waves for the
// factors run side by side on the compute units.  Right-hand sides are
// columns of a union block of kmax = max_f nrhs_f columns: a gather pass
// copies factor f's columns into its row range (zeros past nrhs_f), the
// union plan solves all kmax columns (several per launch on the fused
// schedule; the blocked walk skips each factor's rows past nrhs_f through
// the per-row column counts written for the call), and a scatter pass
// copies the solutions back.
#include <algorithm>
#include <cstring>
#include <vector>

#include "rb_internal.h"

struct rb_batch {
    int nf = 0, n_total = 0, nnz_total = 0;
    std::vector<int> m, off;
    rb_plan *plan = nullptr;
    rb::Arena arena;
    int *d_fid = nullptr;   // factor of each union row
    int *d_rowcap = nullptr;  // per permuted row: its factor's column count
    int *d_off = nullptr;   // [nf] row offsets
    // per-call tables (device, filled from pinned host staging)
    const double **d_bp = nullptr;
    double **d_xp = nullptr;
    int *d_ld = nullptr;    // [3 * nf]: ldb, ldx, nrhs
    void *h_stage = nullptr;
    size_t stage_bytes = 0;
    hipEvent_t staged = nullptr;  // the last table upload has been read
    // union right-hand sides / solutions, n_total x kalloc column-major
    double *d_B = nullptr, *d_X = nullptr;
    int kalloc = 0;
    ~rb_batch() {
        if (d_B) (void)hipFree(d_B);
        if (d_X) (void)hipFree(d_X);
        if (h_stage) (void)hipHostFree(h_stage);
        if (staged) (void)hipEventDestroy(staged);
    }
};

namespace rb {
namespace {

constexpr int kT = 256;
inline int nblk(long n) { return (int)((n + kT - 1) / kT); }

// factor f's CSC into the union at row/column offset off (entry offset eo)
__global__ void k_place(const int *__restrict__ cp, const int *__restrict__ ri,
                        const double *__restrict__ v, int m, int nnz, int off,
                        int eo, int *ucp, int *uri, double *uv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) ucp[off + i] = eo + cp[i];
    if (i < nnz) {
        uri[eo + i] = off + ri[i];
        uv[eo + i] = v[i];
    }
}

__global__ void k_fill_fid(int *fid, int off, int m, int f) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) fid[off + i] = f;
}

// union column k, row i <- factor fid[i]'s column k (0 past its nrhs)
__global__ void k_gather(const int *__restrict__ fid, const int *__restrict__ off,
                         const double *const *__restrict__ bp,
                         const int *__restrict__ ld, int nf, long n, int kmax,
                         double *B) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * kmax) return;
    const long k = t / n, i = t - k * n;
    const int f = fid[i];
    const int nr = ld[2 * nf + f];
    B[t] = k < nr ? bp[f][k * (long)ld[f] + (i - off[f])] : 0.0;
}

// the union plan's row caps for one call: permuted row i (original row
// lf[i]) solves as many columns as its factor has
__global__ void k_rowcap(const int *__restrict__ lf, const int *__restrict__ fid,
                         const int *__restrict__ nrhs, int n, int *rowcap) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) rowcap[i] = nrhs[fid[lf[i]]];
}

__global__ void k_scatter(const int *__restrict__ fid, const int *__restrict__ off,
                          double *const *__restrict__ xp,
                          const int *__restrict__ ld, int nf, long n, int kmax,
                          const double *__restrict__ X) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * kmax) return;
    const long k = t / n, i = t - k * n;
    const int f = fid[i];
    if (k < ld[2 * nf + f]) xp[f][k * (long)ld[nf + f] + (i - off[f])] = X[t];
}

}  // namespace
}  // namespace rb

using namespace rb;

static rb_status done() {
    set_error(RB_OK, "");
    return RB_OK;
}

extern "C" {

rb_status rb_batch_create(int nf, const int *const *d_colptr,
                          const int *const *d_rowidx,
                          const double *const *d_val, const int *m,
                          const int *nnz, int substitution, int lv,
                          const rb_opts *opts, void *stream, rb_batch **out) {
    if (!out) {
        set_error(RB_ERR_INVALID_ARG, "batch is NULL");
        return RB_ERR_INVALID_ARG;
    }
    *out = nullptr;
    if (nf < 1 || !d_colptr || !d_rowidx || !d_val || !m || !nnz) {
        set_error(RB_ERR_INVALID_ARG, "invalid batch arguments");
        return RB_ERR_INVALID_ARG;
    }
    long n_total = 0, nnz_total = 0;
    for (int f = 0; f < nf; f++) {
        if (m[f] < 0 || nnz[f] < 0 || !d_colptr[f] ||
            (nnz[f] > 0 && (!d_rowidx[f] || !d_val[f]))) {
            set_error(RB_ERR_INVALID_ARG, "invalid factor " + std::to_string(f));
            return RB_ERR_INVALID_ARG;
        }
        n_total += m[f];
        nnz_total += nnz[f];
    }
    if (n_total > 0x7fffffffL - 1 || nnz_total > 0x7fffffffL) {
        set_error(RB_ERR_INVALID_ARG, "batch exceeds 32-bit sizes");
        return RB_ERR_INVALID_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    rb_batch *B = new rb_batch();
    auto fail = [&](rb_status st) {
        (void)hipStreamSynchronize(s);
        if (B->plan) rb_plan_destroy(B->plan);
        B->arena.release();
        delete B;
        return st;
    };
    B->nf = nf;
    B->n_total = (int)n_total;
    B->nnz_total = (int)nnz_total;
    B->m.assign(m, m + nf);
    B->off.resize(nf);
    // union CSC (temporary) and the row -> factor table
    int *ucp = nullptr, *uri = nullptr;
    double *uv = nullptr;
    Arena tmp;
    rb_status st = RB_OK;
    if ((st = tmp.alloc(&ucp, n_total + 1)) != RB_OK ||
        (st = tmp.alloc(&uri, nnz_total)) != RB_OK ||
        (st = tmp.alloc(&uv, nnz_total)) != RB_OK ||
        (st = B->arena.alloc(&B->d_fid, n_total)) != RB_OK ||
        (st = B->arena.alloc(&B->d_rowcap, std::max(1L, n_total))) != RB_OK ||
        (st = B->arena.alloc(&B->d_off, nf)) != RB_OK ||
        (st = B->arena.alloc(&B->d_bp, nf)) != RB_OK ||
        (st = B->arena.alloc(&B->d_xp, nf)) != RB_OK ||
        (st = B->arena.alloc(&B->d_ld, 3 * nf)) != RB_OK) {
        tmp.release();
        return fail(st);
    }
    long o = 0, eo = 0;
    for (int f = 0; f < nf; f++) {
        B->off[f] = (int)o;
        const int mm = m[f], zz = nnz[f];
        const long work = std::max(mm, zz);
        if (work > 0)
            k_place<<<nblk(work), kT, 0, s>>>(d_colptr[f], d_rowidx[f],
                                              d_val[f], mm, zz, (int)o, (int)eo,
                                              ucp, uri, uv);
        if (mm > 0) k_fill_fid<<<nblk(mm), kT, 0, s>>>(B->d_fid, (int)o, mm, f);
        o += mm;
        eo += zz;
    }
    const int last = (int)nnz_total;
    if (hipMemcpyAsync(ucp + n_total, &last, sizeof(int), hipMemcpyHostToDevice,
                       s) != hipSuccess ||
        hipMemcpyAsync(B->d_off, B->off.data(), sizeof(int) * nf,
                       hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        tmp.release();
        return fail(status_of_hip(hipGetLastError(), "batch assembly"));
    }
    st = rb_plan_create(ucp, uri, uv, (int)n_total, (int)n_total,
                        (int)nnz_total, substitution, lv, opts, stream,
                        &B->plan);
    (void)hipStreamSynchronize(s);
    tmp.release();
    if (st != RB_OK) return fail(st);
    // the blocked walk skips the rows of factors with fewer columns than
    // the batch's widest (graphs captured while tuning hold no caps)
    B->plan->d_rowcap = B->d_rowcap;
    B->plan->rowcap_all = false;
    destroy_graph(B->plan);
    B->stage_bytes = sizeof(void *) * 2 * nf + sizeof(int) * 3 * nf;
    if (hipHostMalloc(&B->h_stage, B->stage_bytes) != hipSuccess ||
        hipEventCreateWithFlags(&B->staged, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        set_error(RB_ERR_NO_MEMORY, "pinned staging buffer");
        return fail(RB_ERR_NO_MEMORY);
    }
    *out = B;
    return done();
}

rb_plan *rb_batch_plan(rb_batch *B) { return B ? B->plan : nullptr; }

// Upload the per-call tables and make sure the union buffers hold kmax
// columns.  Returns kmax through *kmax.
static rb_status batch_stage(rb_batch *B, const double *const *d_b,
                             const int *ldb, double *const *d_x,
                             const int *ldx, const int *nrhs, hipStream_t s,
                             int *kmax) {
    const int nf = B->nf;
    int k = 0;
    for (int f = 0; f < nf; f++) {
        if (nrhs[f] < 0 || (nrhs[f] > 0 && (!d_b[f] || !d_x[f] ||
                                            ldb[f] < B->m[f] || ldx[f] < B->m[f]))) {
            set_error(RB_ERR_INVALID_ARG, "invalid solve arguments for factor " +
                                              std::to_string(f));
            return RB_ERR_INVALID_ARG;
        }
        k = std::max(k, nrhs[f]);
    }
    *kmax = k;
    if (k == 0 || B->n_total == 0) return RB_OK;
    if (k > B->kalloc) {
        RB_HIP(hipStreamSynchronize(s));
        if (B->d_B) RB_HIP(hipFree(B->d_B));
        if (B->d_X) RB_HIP(hipFree(B->d_X));
        B->d_B = B->d_X = nullptr;
        B->kalloc = 0;
        const size_t bytes = sizeof(double) * (size_t)B->n_total * k;
        if (hipMalloc(&B->d_B, bytes) != hipSuccess ||
            hipMalloc(&B->d_X, bytes) != hipSuccess) {
            (void)hipGetLastError();
            set_error(RB_ERR_NO_MEMORY, "hipMalloc failed (batch columns)");
            return RB_ERR_NO_MEMORY;
        }
        B->kalloc = k;
    }
    // the previous upload must have been read before the staging buffer
    // is rewritten
    RB_HIP(hipEventSynchronize(B->staged));
    char *h = (char *)B->h_stage;
    std::memcpy(h, d_b, sizeof(void *) * nf);
    std::memcpy(h + sizeof(void *) * nf, d_x, sizeof(void *) * nf);
    int *hl = (int *)(h + sizeof(void *) * 2 * nf);
    std::memcpy(hl, ldb, sizeof(int) * nf);
    std::memcpy(hl + nf, ldx, sizeof(int) * nf);
    std::memcpy(hl + 2 * nf, nrhs, sizeof(int) * nf);
    RB_HIP(hipMemcpyAsync(B->d_bp, h, sizeof(void *) * nf,
                          hipMemcpyHostToDevice, s));
    RB_HIP(hipMemcpyAsync(B->d_xp, h + sizeof(void *) * nf,
                          sizeof(void *) * nf, hipMemcpyHostToDevice, s));
    RB_HIP(hipMemcpyAsync(B->d_ld, hl, sizeof(int) * 3 * nf,
                          hipMemcpyHostToDevice, s));
    RB_HIP(hipEventRecord(B->staged, s));
    const long tot = (long)B->n_total * k;
    k_gather<<<nblk(tot), kT, 0, s>>>(B->d_fid, B->d_off, B->d_bp, B->d_ld, nf,
                                      B->n_total, k, B->d_B);
    k_rowcap<<<nblk(B->n_total), kT, 0, s>>>(B->plan->d_levelItem, B->d_fid,
                                             B->d_ld + 2 * nf, B->n_total,
                                             B->d_rowcap);
    RB_HIP(hipGetLastError());
    B->plan->rowcap_set = true;
    B->plan->rowcap_all = false;
    return RB_OK;
}

static rb_status batch_unstage(rb_batch *B, int k, hipStream_t s) {
    const long tot = (long)B->n_total * k;
    k_scatter<<<nblk(tot), kT, 0, s>>>(B->d_fid, B->d_off, B->d_xp, B->d_ld,
                                       B->nf, B->n_total, k, B->d_X);
    RB_HIP(hipGetLastError());
    return RB_OK;
}

rb_status rb_batch_solve(rb_batch *B, const double *const *d_b,
                         const int *ldb, double *const *d_x, const int *ldx,
                         const int *nrhs, void *stream) {
    if (!B || !d_b || !ldb || !d_x || !ldx || !nrhs) {
        set_error(RB_ERR_INVALID_ARG, "null argument");
        return RB_ERR_INVALID_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    int k = 0;
    RB_TRY(batch_stage(B, d_b, ldb, d_x, ldx, nrhs, s, &k));
    if (k == 0 || B->n_total == 0) return done();
    RB_TRY(rb_solve(B->plan, B->d_B, B->n_total, B->d_X, B->n_total, k, stream));
    RB_TRY(batch_unstage(B, k, s));
    return done();
}

rb_status rb_batch_solve_f32ir(rb_batch *B, const double *const *d_b,
                               const int *ldb, double *const *d_x,
                               const int *ldx, const int *nrhs, double tol,
                               int max_iter, void *stream, int *iters,
                               double *relres) {
    if (!B || !d_b || !ldb || !d_x || !ldx || !nrhs) {
        set_error(RB_ERR_INVALID_ARG, "null argument");
        return RB_ERR_INVALID_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    int k = 0;
    if (iters) *iters = 0;
    if (relres) *relres = 0.0;
    RB_TRY(batch_stage(B, d_b, ldb, d_x, ldx, nrhs, s, &k));
    if (k == 0 || B->n_total == 0) return done();
    const rb_status st = rb_solve_f32ir(B->plan, B->d_B, B->n_total, B->d_X,
                                        B->n_total, k, tol, max_iter, stream,
                                        iters, relres);
    if (st != RB_OK && st != RB_ERR_NOT_CONVERGED) return st;
    RB_TRY(batch_unstage(B, k, s));
    if (st != RB_OK) {
        set_error(st, "FP32 refinement did not reach the requested residual");
        return st;
    }
    return done();
}

rb_status rb_batch_destroy(rb_batch *B) {
    if (!B) return done();
    (void)hipDeviceSynchronize();
    if (B->plan) rb_plan_destroy(B->plan);
    B->arena.release();
    delete B;
    return done();
}

}  // extern "C"
