// C ABI of librecblock: plan API, inspection, and the two legacy entry
// points with the reference's signatures.
#include <chrono>
#include <cmath>
#include <cstring>

#include "rb_internal.h"

namespace rb {

namespace {
thread_local rb_status g_status = RB_OK;
thread_local std::string g_msg;
int g_bench_repeat = 200;  // BENCH_REPEAT, R/common.h:13-15
}  // namespace

void set_error(rb_status st, const std::string &msg) {
    g_status = st;
    g_msg = msg;
}

rb_status status_of_hip(hipError_t e, const char *what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    rb_status st = e == hipErrorOutOfMemory ? RB_ERR_NO_MEMORY : RB_ERR_HIP;
    set_error(st, m);
    return st;
}

static rb_status ok() {
    g_status = RB_OK;
    g_msg.clear();
    return RB_OK;
}

static int lv_auto(int m, int device_id) {
    // R/main.cu:245-254
    int li;
    const double thr = device_id == 0 ? 92160.0 : 58880.0;
    for (li = 1; li <= 100; li++)
        if ((double)m / std::pow(2.0, (double)(li + 1)) < thr) break;
    return li;
}

// Kernel choice per block (tune_blocks, rb_opts.block_kernels) and the
// schedule.  BLOCKED and FUSED are taken as asked (a fused-only plan keeps
// the reference selector's block kernels without timing them); AUTO times
// warm solves of each candidate on scratch right-hand sides (tune_rhs
// columns per call, default 1) and keeps the faster (the plan is built for
// both).  The fused schedule is timed first: with one column, the block
// kernel timing stops as soon as the timed blocks alone take more than
// twice the fused solve, and the blocked schedule is then not timed at all;
// otherwise its first timed solve settles the choice when it is more than
// twice the fused time (its other two are then not run).
static rb_status choose_schedule(rb_plan *p, int block_mode, hipStream_t s) {
    if (p->want_schedule == RB_SCHED_BLOCKED || !p->fused) {
        p->schedule = RB_SCHED_BLOCKED;
        return tune_blocks(p, block_mode, s);
    }
    if (p->want_schedule == RB_SCHED_FUSED) {
        p->schedule = RB_SCHED_FUSED;
        return tune_blocks(p, 1, s);
    }
    const int m = p->m;
    const int k = std::max(1, std::min(p->tune_rhs, 64));
    const size_t bytes = sizeof(double) * (size_t)m * k;
    double *db = nullptr, *dx = nullptr;
    RB_HIP(hipMalloc(&db, bytes));
    if (hipMalloc(&dx, bytes) != hipSuccess) {
        (void)hipFree(db);
        return status_of_hip(hipGetLastError(), "hipMalloc(tune)");
    }
    rb_status st = RB_OK;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    // the work of a solve does not depend on the values: b = 0
    if (hipMemsetAsync(db, 0, bytes, s) != hipSuccess) st = RB_ERR_HIP;
    auto timed = [&](int c, int reps) {
        float t = 0.f;
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < reps && st == RB_OK; r++)
            st = c == 0 ? solve_blocked(p, db, m, dx, m, k, s)
                        : solve_fused(p, db, m, dx, m, k, s);
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        (void)hipEventElapsedTime(&t, e0, e1);
        return (double)t;
    };
    // a schedule whose solve tripped the watchdog is never chosen
    auto tripped = [&]() {
        int d = 0;
        if (hipMemcpy(&d, p->d_status, sizeof(int), hipMemcpyDeviceToHost) !=
            hipSuccess)
            st = RB_ERR_HIP;
        if (d != 0) {
            const int z = 0;
            (void)hipMemcpy(p->d_status, &z, sizeof(int), hipMemcpyHostToDevice);
            fused_rearm(p);
        }
        return d != 0;
    };
    double ms[2] = {0.0, 0.0};
    bool ok[2] = {true, true}, cut = false;
    if (st == RB_OK) {
        (void)timed(1, 1);  // warm
        ms[1] = timed(1, 3) / 3.0;
        ok[1] = !tripped();
    }
    if (st == RB_OK)
        st = tune_blocks(p, block_mode, s, k == 1 && ok[1] ? 2.0 * ms[1] : 0.0,
                         &cut);
    if (st == RB_OK && !cut) {
        (void)timed(0, 1);  // warm (captures the walk)
        const double t1 = timed(0, 1);
        ms[0] = t1 > 2.0 * ms[1] && ok[1] ? t1 : (t1 + timed(0, 2)) / 3.0;
        ok[0] = !tripped();
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(db);
    (void)hipFree(dx);
    if (st != RB_OK) return st;
    p->tune_ms_blocked = cut ? 0.0 : ms[0];
    p->tune_ms_fused = ms[1];
    p->schedule = ok[1] && (cut || !ok[0] || ms[1] < ms[0]) ? RB_SCHED_FUSED
                                                            : RB_SCHED_BLOCKED;
    return RB_OK;
}

}  // namespace rb

using namespace rb;

extern "C" {

rb_status rb_last_status(void) { return g_status; }

const char *rb_last_error(void) { return g_msg.c_str(); }

int rb_set_bench_repeat(int repeat) {
    int old = g_bench_repeat;
    if (repeat >= 1) g_bench_repeat = repeat;
    return old;
}

int rb_lv_auto(int m, int device_id) { return lv_auto(m, device_id); }

void rb_opts_default(rb_opts *opts) {
    if (!opts) return;
    std::memset(opts, 0, sizeof(*opts));
    opts->schedule = RB_SCHED_AUTO;
}

rb_status rb_plan_create(const int *d_colptr, const int *d_rowidx,
                         const double *d_val, int m, int n, int nnz,
                         int substitution, int lv, const rb_opts *opts,
                         void *stream, rb_plan **plan) {
    if (!plan) {
        set_error(RB_ERR_INVALID_ARG, "plan is NULL");
        return RB_ERR_INVALID_ARG;
    }
    *plan = nullptr;
    if (m != n) {
        set_error(RB_ERR_NOT_SQUARE, "m != n");
        return RB_ERR_NOT_SQUARE;
    }
    if (m < 0 || nnz < 0 || !d_colptr || (nnz > 0 && (!d_rowidx || !d_val)) ||
        (substitution != 0 && substitution != 1)) {
        set_error(RB_ERR_INVALID_ARG,
                  "invalid argument (sizes, null pointers or substitution)");
        return RB_ERR_INVALID_ARG;
    }
    rb_opts o;
    rb_opts_default(&o);
    if (opts) o = *opts;
    if (lv == -1) lv = lv_auto(m, o.device_id);
    if (lv < 0 || lv > 20) {
        set_error(RB_ERR_INVALID_ARG, "lv must be -1 or in [0, 20]");
        return RB_ERR_INVALID_ARG;
    }
    rb_plan *p = new rb_plan();
    p->m = m;
    p->nnz = nnz;
    p->subst = substitution;
    p->lv = lv;
    (void)hipGetDevice(&p->device);
    p->schedule = RB_SCHED_BLOCKED;
    p->want_schedule = o.schedule;
    p->fused_waves = o.fused_waves;
    p->graph_failed = o.no_graph != 0;  // eager launches only
    p->mixed = o.mixed != 0;
    p->tune_rhs = o.tune_rhs;
    if (o.watchdog_spins > 0) p->spin_limit = (unsigned)o.watchdog_spins;
    rb_status st = RB_OK;
    Arena canon;  // the canonical copy lives until the plan is built
    struct CanonGuard {
        Arena &a;
        ~CanonGuard() { a.release(); }
    } canon_guard{canon};
    if (o.canonicalize) {
        int *cp = nullptr, *ri = nullptr;
        double *cv = nullptr;
        int cnnz = 0;
        st = canonicalize_csc(d_colptr, d_rowidx, d_val, m, nnz, canon, &cp,
                              &ri, &cv, &cnnz, (hipStream_t)stream);
        if (st == RB_OK) {
            d_colptr = cp;
            d_rowidx = ri;
            d_val = cv;
            p->nnz = cnnz;
        }
    }
    if (st == RB_OK)
        st = build_plan(p, d_colptr, d_rowidx, d_val, (hipStream_t)stream);
    if (st == RB_OK) {
        auto t0 = std::chrono::steady_clock::now();
        // kernels per block for the blocked walk (not timed when only the
        // fused schedule will run)
        st = choose_schedule(p, o.block_kernels, (hipStream_t)stream);
        p->phase_ms[3] = std::chrono::duration<double, std::milli>(
                             std::chrono::steady_clock::now() - t0)
                             .count();
    }
    if (st != RB_OK) {
        (void)hipStreamSynchronize((hipStream_t)stream);
        destroy_graph(p);
        fused_destroy(p);
        p->arena.release();
        delete p;
        return st;
    }
    *plan = p;
    return ok();
}

rb_status rb_solve(rb_plan *p, const double *d_b, int ldb, double *d_x,
                   int ldx, int nrhs, void *stream) {
    if (!p || nrhs < 0 || (nrhs > 0 && (!d_b || !d_x)) ||
        (nrhs > 0 && (ldb < p->m || ldx < p->m))) {
        set_error(RB_ERR_INVALID_ARG, "invalid solve arguments");
        return RB_ERR_INVALID_ARG;
    }
    if (nrhs == 0) return ok();
    p->mark_first = true;
    rowcap_for_call(p, (hipStream_t)stream);
    rb_status st = p->schedule == RB_SCHED_FUSED
                       ? solve_fused(p, d_b, ldb, d_x, ldx, nrhs,
                                     (hipStream_t)stream)
                       : solve_blocked(p, d_b, ldb, d_x, ldx, nrhs,
                                       (hipStream_t)stream);
    if (st != RB_OK) return st;
    return ok();
}

rb_status rb_plan_check(rb_plan *p, void *stream) {
    if (!p) {
        set_error(RB_ERR_INVALID_ARG, "null plan");
        return RB_ERR_INVALID_ARG;
    }
    RB_HIP(hipStreamSynchronize((hipStream_t)stream));
    int st = 0;
    RB_HIP(hipMemcpy(&st, p->d_status, sizeof(int), hipMemcpyDeviceToHost));
    if (st != 0) {
        // clear the flag and re-arm the per-solve state (an aborted fused
        // solve leaves slots of both x buffers un-armed)
        const int z = 0;
        RB_HIP(hipMemcpy(p->d_status, &z, sizeof(int), hipMemcpyHostToDevice));
        if (st & 7) {
            fused_rearm(p);
            set_error(RB_ERR_INTERNAL,
                      "device watchdog: a solve waited on a value that never "
                      "arrived (x was written as NaN by every solve since)");
            return RB_ERR_INTERNAL;
        }
        set_error(RB_ERR_NOT_CONVERGED,
                  "an asynchronous FP32 refinement missed its tolerance");
        return RB_ERR_NOT_CONVERGED;
    }
    return ok();
}

rb_status rb_solve_f32ir(rb_plan *p, const double *d_b, int ldb, double *d_x,
                         int ldx, int nrhs, double tol, int max_iter,
                         void *stream, int *iters, double *relres) {
    if (!p || nrhs < 0 || (nrhs > 0 && (!d_b || !d_x)) ||
        (nrhs > 0 && (ldb < p->m || ldx < p->m)) || max_iter < 1 ||
        !(tol > 0.0)) {
        set_error(RB_ERR_INVALID_ARG, "invalid refinement arguments");
        return RB_ERR_INVALID_ARG;
    }
    if (!p->mixed) {
        set_error(RB_ERR_INVALID_ARG,
                  "plan was created without rb_opts.mixed (no FP32 copies)");
        return RB_ERR_INVALID_ARG;
    }
    if (iters) *iters = 0;
    if (relres) *relres = 0.0;
    if (nrhs == 0) return ok();
    p->mark_first = true;
    rowcap_for_call(p, (hipStream_t)stream);
    rb_status st = solve_f32ir(p, d_b, ldb, d_x, ldx, nrhs, tol, max_iter,
                               (hipStream_t)stream, iters, relres);
    if (st != RB_OK) return st;
    return ok();
}

rb_status rb_plan_set_timing(rb_plan *p, int enable) {
    if (!p) {
        set_error(RB_ERR_INVALID_ARG, "null plan");
        return RB_ERR_INVALID_ARG;
    }
    if (enable && !p->ev0) {
        RB_HIP(hipEventCreate(&p->ev0));
        RB_HIP(hipEventCreate(&p->ev1));
    }
    p->timing = enable != 0;
    return ok();
}

rb_status rb_plan_kernel_ms(rb_plan *p, double *ms) {
    if (!p || !ms) {
        set_error(RB_ERR_INVALID_ARG, "null argument");
        return RB_ERR_INVALID_ARG;
    }
    if (!p->ev0) {
        set_error(RB_ERR_INVALID_ARG, "timing was never enabled on this plan");
        return RB_ERR_INVALID_ARG;
    }
    RB_HIP(hipEventSynchronize(p->ev1));
    float f = 0.0f;
    RB_HIP(hipEventElapsedTime(&f, p->ev0, p->ev1));
    *ms = f;
    return ok();
}

rb_status rb_plan_destroy(rb_plan *p) {
    if (!p) return ok();
    (void)hipDeviceSynchronize();
    if (p->ev0) (void)hipEventDestroy(p->ev0);
    if (p->ev1) (void)hipEventDestroy(p->ev1);
    destroy_graph(p);
    fused_destroy(p);
    free_solve_scratch(p);
    p->arena.release();
    delete p;
    return ok();
}

rb_status rb_plan_block_kernels(const rb_plan *p, int *kern, int *lanes,
                                double *tuned_ms) {
    if (!p) {
        set_error(RB_ERR_INVALID_ARG, "null plan");
        return RB_ERR_INVALID_ARG;
    }
    for (int i = 0; i < p->nblocks; i++) {
        const Block &B = p->blocks[i];
        if (kern) kern[i] = B.kind == 0 ? B.kern : 0;
        if (lanes) lanes[i] = B.kind == 1 ? B.lanes : 0;
        if (tuned_ms) tuned_ms[i] = B.tuned_ms;
    }
    return ok();
}

rb_status rb_plan_info_get(const rb_plan *p, rb_plan_info *info) {
    if (!p || !info) {
        set_error(RB_ERR_INVALID_ARG, "null argument");
        return RB_ERR_INVALID_ARG;
    }
    info->m = p->m;
    info->nnz = p->nnz;
    info->substitution = p->subst;
    info->lv = p->lv;
    info->nblocks = p->nblocks;
    info->nlevel = p->nlevel;
    info->schedule = p->schedule;
    info->preprocess_ms = p->preprocess_ms;
    info->device_bytes = p->arena.bytes;
    for (int i = 0; i < 4; i++) info->phase_ms[i] = p->phase_ms[i];
    info->tune_ms[0] = p->tune_ms_blocked;
    info->tune_ms[1] = p->tune_ms_fused;
    return ok();
}

rb_status rb_plan_blocks(const rb_plan *p, int *blk_m, int *blk_n,
                         int *blk_nnz, int *up, int *down, int *left,
                         int *right, int *method, int *nlv, int *nnzr,
                         int *m_new, int *lenmax, int *longrow, int *longlen,
                         double *empty_ratio) {
    if (!p) {
        set_error(RB_ERR_INVALID_ARG, "null plan");
        return RB_ERR_INVALID_ARG;
    }
    for (int i = 0; i < p->nblocks; i++) {
        const Block &b = p->blocks[i];
        if (blk_m) blk_m[i] = b.blk_m;
        if (blk_n) blk_n[i] = b.blk_n;
        if (blk_nnz) blk_nnz[i] = b.blk_nnz;
        if (up) up[i] = b.up;
        if (down) down[i] = b.down;
        if (left) left[i] = b.left;
        if (right) right[i] = b.right;
        if (method) method[i] = b.method;
        if (nlv) nlv[i] = b.kind == 0 ? b.nlv : 0;
        if (nnzr) nnzr[i] = b.nnzr;
        if (m_new) m_new[i] = b.kind == 1 ? b.m_new : 0;
        if (lenmax) lenmax[i] = b.kind == 1 ? b.lenmax : 0;
        if (longrow) longrow[i] = b.kind == 1 ? b.longrow : 0;
        if (longlen) longlen[i] = b.kind == 1 ? b.longlen : 0;
        if (empty_ratio) empty_ratio[i] = b.kind == 1 ? b.empty_ratio : 0.0;
    }
    return ok();
}

rb_status rb_plan_levels(const rb_plan *p, int *levelItem, int *levelPtr) {
    if (!p) {
        set_error(RB_ERR_INVALID_ARG, "null plan");
        return RB_ERR_INVALID_ARG;
    }
    if (levelItem && p->m > 0)
        RB_HIP(hipMemcpy(levelItem, p->d_levelItem, sizeof(int) * p->m,
                         hipMemcpyDeviceToHost));
    if (levelPtr)
        RB_HIP(hipMemcpy(levelPtr, p->d_levelPtr,
                         sizeof(int) * (size_t)(p->nlevel + 1),
                         hipMemcpyDeviceToHost));
    return ok();
}

rb_status rb_plan_tri_levels(const rb_plan *p, int t, int *lvptr,
                             int *lvitem, int *nnz_lv) {
    if (!p || t < 0 || t >= p->ntri) {
        set_error(RB_ERR_INVALID_ARG, "bad plan or block");
        return RB_ERR_INVALID_ARG;
    }
    const Block &b = p->blocks[2 * t];
    if (lvptr)
        RB_HIP(hipMemcpy(lvptr, p->d_lvptr + b.lvptr_base,
                         sizeof(int) * (size_t)(b.nlv + 1),
                         hipMemcpyDeviceToHost));
    if (lvitem && b.blk_m > 0)
        RB_HIP(hipMemcpy(lvitem, p->d_lvitem + b.lv_base,
                         sizeof(int) * (size_t)b.blk_m,
                         hipMemcpyDeviceToHost));
    if (nnz_lv)
        RB_HIP(hipMemcpy(nnz_lv, p->d_nnzlv + b.lvptr_base,
                         sizeof(int) * (size_t)b.nlv, hipMemcpyDeviceToHost));
    return ok();
}

rb_status rb_plan_block_csr(const rb_plan *p, int i, int *rowptr,
                            int *colidx, double *val) {
    if (!p || i < 0 || i >= p->nblocks) {
        set_error(RB_ERR_INVALID_ARG, "bad plan or block");
        return RB_ERR_INVALID_ARG;
    }
    const Block &b = p->blocks[i];
    std::vector<int> rp(b.blk_m + 1);
    RB_HIP(hipMemcpy(rp.data(), p->d_rowptr + b.slot_base,
                     sizeof(int) * (size_t)(b.blk_m + 1),
                     hipMemcpyDeviceToHost));
    const int base = rp[0], noff = rp[b.blk_m] - rp[0];
    std::vector<int> c(noff > 0 ? noff : 1);
    std::vector<double> v(noff > 0 ? noff : 1);
    std::vector<double> invd(b.kind == 0 ? (b.blk_m > 0 ? b.blk_m : 1) : 1);
    if (noff > 0) {
        RB_HIP(hipMemcpy(c.data(), p->d_col + base, sizeof(int) * noff,
                         hipMemcpyDeviceToHost));
        RB_HIP(hipMemcpy(v.data(), p->d_val + base, sizeof(double) * noff,
                         hipMemcpyDeviceToHost));
    }
    if (b.kind == 0 && b.blk_m > 0)
        RB_HIP(hipMemcpy(invd.data(), p->d_invd + b.up,
                         sizeof(double) * b.blk_m, hipMemcpyDeviceToHost));
    // assemble, inserting the diagonal of triangular blocks in column order
    int o = 0;
    if (rowptr) rowptr[0] = 0;
    for (int r = 0; r < b.blk_m; r++) {
        bool diag_done = b.kind != 0;
        for (int q = rp[r] - base; q < rp[r + 1] - base; q++) {
            if (!diag_done && c[q] > r) {
                if (colidx) colidx[o] = r;
                if (val) val[o] = 1.0 / invd[r];
                o++;
                diag_done = true;
            }
            if (colidx) colidx[o] = c[q];
            if (val) val[o] = v[q];
            o++;
        }
        if (!diag_done) {
            if (colidx) colidx[o] = r;
            if (val) val[o] = 1.0 / invd[r];
            o++;
        }
        if (rowptr) rowptr[r + 1] = o;
    }
    return ok();
}

rb_status rb_plan_permuted_csc(const rb_plan *p, int *colptr, int *rowidx,
                               double *val) {
    if (!p) {
        set_error(RB_ERR_INVALID_ARG, "null plan");
        return RB_ERR_INVALID_ARG;
    }
    if (colptr)
        RB_HIP(hipMemcpy(colptr, p->d_pcolptr, sizeof(int) * (size_t)(p->m + 1),
                         hipMemcpyDeviceToHost));
    if (rowidx && p->nnz > 0)
        RB_HIP(hipMemcpy(rowidx, p->d_prowidx, sizeof(int) * (size_t)p->nnz,
                         hipMemcpyDeviceToHost));
    if (val && p->nnz > 0)
        RB_HIP(hipMemcpy(val, p->d_pval, sizeof(double) * (size_t)p->nnz,
                         hipMemcpyDeviceToHost));
    return ok();
}

// --------------------------------------------------------------- legacy

void recblocking_solver_cuda(int *d_cscColPtrTR, int *d_cscRowIdxTR,
                             double *d_cscValTR, int m, int n, int nnzTR,
                             double *d_x, double *d_b, int substitution,
                             int lv, double *cal_time,
                             double *preprocess_time) {
    rb_plan *p = nullptr;
    rb_opts o;
    rb_opts_default(&o);
    auto t0 = std::chrono::steady_clock::now();
    rb_status st = rb_plan_create(d_cscColPtrTR, d_cscRowIdxTR, d_cscValTR,
                                  m, n, nnzTR, substitution, lv, &o, nullptr,
                                  &p);
    auto t1 = std::chrono::steady_clock::now();
    if (st != RB_OK) return;
    if (preprocess_time)
        *preprocess_time =
            std::chrono::duration<double, std::milli>(t1 - t0).count();
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        set_error(RB_ERR_HIP, "hipEventCreate failed");
        rb_plan_destroy(p);
        return;
    }
    // one untimed solve, then BENCH_REPEAT timed ones (mean, as the
    // reference reports, R/recblocking_calculate.h:37-167)
    st = rb_solve(p, d_b, m, d_x, m, 1, nullptr);
    const int rep = g_bench_repeat;
    if (st == RB_OK) {
        (void)hipEventRecord(e0, nullptr);
        for (int r = 0; r < rep && st == RB_OK; r++)
            st = rb_solve(p, d_b, m, d_x, m, 1, nullptr);
        (void)hipEventRecord(e1, nullptr);
    }
    hipError_t he = hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st == RB_OK && he != hipSuccess) st = status_of_hip(he, "solve");
    if (st == RB_OK) st = rb_plan_check(p, nullptr);
    if (cal_time) *cal_time = st == RB_OK ? (double)ms / rep : 0.0;
    rb_status keep = g_status;
    std::string msg = g_msg;
    rb_plan_destroy(p);
    if (st != RB_OK) set_error(keep, msg);
}

void recblocking_solver(int *cscColPtr, int *cscRowIdx, double *cscVal, int m,
                        int n, int nnz, double *b, double *x, double *x_ref,
                        int rhs, int lv, int substitution, double *cal_time) {
    (void)x_ref;
    if (m != n || m < 0 || nnz < 0 || rhs < 0 || !cscColPtr) {
        set_error(m != n ? RB_ERR_NOT_SQUARE : RB_ERR_INVALID_ARG,
                  "invalid argument");
        return;
    }
    int *dcp = nullptr, *dri = nullptr;
    double *dv = nullptr, *db = nullptr, *dx = nullptr;
    const size_t nb = (size_t)m * (rhs > 0 ? rhs : 1);
    auto cleanup = [&]() {
        if (dcp) (void)hipFree(dcp);
        if (dri) (void)hipFree(dri);
        if (dv) (void)hipFree(dv);
        if (db) (void)hipFree(db);
        if (dx) (void)hipFree(dx);
    };
    if (hipMalloc(&dcp, sizeof(int) * (m + 1)) != hipSuccess ||
        hipMalloc(&dri, sizeof(int) * (nnz > 0 ? nnz : 1)) != hipSuccess ||
        hipMalloc(&dv, sizeof(double) * (nnz > 0 ? nnz : 1)) != hipSuccess ||
        hipMalloc(&db, sizeof(double) * (nb > 0 ? nb : 1)) != hipSuccess ||
        hipMalloc(&dx, sizeof(double) * (nb > 0 ? nb : 1)) != hipSuccess) {
        (void)hipGetLastError();
        cleanup();
        set_error(RB_ERR_NO_MEMORY, "hipMalloc failed");
        return;
    }
    (void)hipMemcpy(dcp, cscColPtr, sizeof(int) * (m + 1), hipMemcpyHostToDevice);
    if (nnz > 0) {
        (void)hipMemcpy(dri, cscRowIdx, sizeof(int) * nnz, hipMemcpyHostToDevice);
        (void)hipMemcpy(dv, cscVal, sizeof(double) * nnz, hipMemcpyHostToDevice);
    }
    if (rhs > 0)
        (void)hipMemcpy(db, b, sizeof(double) * nb, hipMemcpyHostToDevice);
    rb_plan *p = nullptr;
    rb_opts o;
    rb_opts_default(&o);
    rb_status st = rb_plan_create(dcp, dri, dv, m, n, nnz, substitution, lv,
                                  &o, nullptr, &p);
    if (st == RB_OK && rhs > 0) {
        hipEvent_t e0, e1;
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        st = rb_solve(p, db, m, dx, m, rhs, nullptr);
        const int rep = g_bench_repeat;
        (void)hipEventRecord(e0, nullptr);
        for (int r = 0; r < rep && st == RB_OK; r++)
            st = rb_solve(p, db, m, dx, m, rhs, nullptr);
        (void)hipEventRecord(e1, nullptr);
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (st == RB_OK) st = rb_plan_check(p, nullptr);
        if (cal_time) *cal_time = st == RB_OK ? (double)ms / rep : 0.0;
        if (st == RB_OK &&
            hipMemcpy(x, dx, sizeof(double) * nb, hipMemcpyDeviceToHost) !=
                hipSuccess)
            st = status_of_hip(hipGetLastError(), "hipMemcpy(x)");
    }
    rb_status keep = g_status;
    std::string msg = g_msg;
    if (p) rb_plan_destroy(p);
    cleanup();
    if (st != RB_OK) set_error(keep, msg);
    else ok();
}

}  // extern "C"
