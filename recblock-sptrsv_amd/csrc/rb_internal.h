// Internal declarations of librecblock (gfx950).  Not installed.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/recblock.h"

namespace rb {

// Thread-local error reporting shared by all entry points.
void set_error(rb_status st, const std::string &msg);
rb_status status_of_hip(hipError_t e, const char *what);

#define RB_HIP(call)                                                        \
    do {                                                                    \
        hipError_t _e = (call);                                             \
        if (_e != hipSuccess) return ::rb::status_of_hip(_e, #call);        \
    } while (0)

#define RB_TRY(expr)                                                        \
    do {                                                                    \
        rb_status _s = (expr);                                              \
        if (_s != RB_OK) return _s;                                         \
    } while (0)

// Sentinel bit pattern of an x slot that is not solved yet: a signalling
// NaN payload that IEEE arithmetic never produces (results are quiet NaNs).
constexpr uint64_t kPending = 0xFFF4A5A55A5A0001ull;
constexpr uint32_t kPending32 = 0xFFA5A001u;  // the same for FP32 slots
constexpr int kLongRow = 2048;  // LONGROW_THRESHOLD, R/utils_spmv_cuda.h:9

// Device allocation helper that tracks bytes held by a plan.
struct Arena {
    std::vector<void *> ptrs;
    size_t bytes = 0;
    // chunk > 0: buffers are carved (256-byte aligned) out of chunks of at
    // least `chunk` bytes, each next chunk twice the last, so a build's
    // temporaries cost a few hipMalloc / hipFree calls instead of one each
    // (every hipFree waits for the device)
    size_t chunk = 0;
    char *cur = nullptr, *end = nullptr;
    Arena() = default;
    explicit Arena(size_t first_chunk) : chunk(first_chunk) {}
    template <class T>
    rb_status alloc(T **p, size_t count) {
        *p = nullptr;
        if (count == 0) count = 1;
        const size_t need = (count * sizeof(T) + 255) & ~(size_t)255;
        if (chunk > 0 && cur && (size_t)(end - cur) >= need) {
            *p = (T *)cur;
            cur += need;
            bytes += count * sizeof(T);
            return RB_OK;
        }
        const size_t got = chunk > 0 ? std::max(need, chunk) : count * sizeof(T);
        hipError_t e = hipMalloc((void **)p, got);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            set_error(RB_ERR_NO_MEMORY, "hipMalloc failed");
            return RB_ERR_NO_MEMORY;
        }
        ptrs.push_back(*p);
        bytes += count * sizeof(T);
        if (chunk > 0) {
            cur = (char *)*p + need;
            end = (char *)*p + got;
            chunk *= 2;
        }
        return RB_OK;
    }
    void release() {
        for (void *p : ptrs) (void)hipFree(p);
        ptrs.clear();
        bytes = 0;
        cur = end = nullptr;
    }
};

// Device-wide primitives (rb_prim.hip, hipCUB).
rb_status exclusive_scan_i32(const int *d_in, int *d_out, int n,
                             hipStream_t s);
rb_status inclusive_max_i32(const int *d_in, int *d_out, int n,
                            hipStream_t s);
rb_status sort_pairs_i32(const int *keys_in, int *keys_out,
                         const int *vals_in, int *vals_out, int n, int bits,
                         hipStream_t s);

// Level analysis: Kahn BFS in the reference's discovery order
// (rb_levels.hip).  colptr/rowidx: square CSC whose columns hold the
// diagonal; rowcnt[v]: entries of row v including the diagonal.
// Writes levelItem[m], levelPtr[nlevel+1] (device) and *nlevel (host).
rb_status bfs_levels(const int *d_colptr, const int *d_rowidx,
                     const int *d_rowcnt, int m, int *d_levelItem,
                     int *d_levelPtr, int *nlevel, hipStream_t s);

// The same level analysis for every triangular block of a block-diagonal
// CSC at once, one workgroup per block: block t owns rows [ups[t],
// ups[t] + bms[t]); its rows in discovery order go to item[ups[t] ..]
// (permuted ids), its block-local level starts to lvptr[lvbase[t] ..]
// (level 0 at 0) and its level count to nlv[t] (device arrays).
rb_status bfs_blocks(const int *d_colptr, const int *d_rowidx,
                     const int *d_rowcnt, int m, int nblocks, const int *d_ups,
                     const int *d_bms, const int *d_lvbase, int *d_item,
                     int *d_lvptr, int *d_nlv, hipStream_t s);

// Per-block descriptor kept on the host (and mirrored on the device for the
// fused executor).
struct Block {
    int kind;  // 0 triangular, 1 square
    int up, down, left, right;
    int blk_m, blk_n, blk_nnz;
    int slot_base;  // first slot (row) of this block in the slot space
    int method;
    int nlv, nnzr, m_new, lenmax, longrow, longlen;
    double empty_ratio;
    int dcsr_base;   // first entry in dcsr_rows
    int long_base;   // first entry in long_rows
    int lv_base;     // triangular: first entry in lvptr / lvitem arrays
    int lvptr_base;
    std::vector<int> lvptr_h;   // method 2: local level boundaries
    std::vector<int> nnzlv_h;   // method 2: find_level_nnz
    // execution choice of the blocked walk (the selector's `method` above
    // is kept as the reference decides it): triangular blocks run kern
    // (kTriRef: the kernel of `method`), square blocks `lanes` lanes per
    // row; set by the rule of the reference selector and then by timing
    // (tune_blocks)
    int kern = 0;
    int lanes = 1;
    int maxdist = 0;        // triangular: largest in-block dependency distance
    double tuned_ms = 0.0;  // time of the chosen kernel when tuned
};

// triangular block kernels of the blocked walk
constexpr int kTriRef = 0;       // method 0 / 2 / (1, 3) as the reference picks
constexpr int kTriSyncfree = 1;  // sync-free CSR pull over the whole grid
constexpr int kTriGroup = 2;     // one workgroup, x ring in LDS
constexpr int kRingRows = 8192;  // kTriGroup's ring (positions)
constexpr int kGroupThreads = 1024;
constexpr int kGroupSlack = 2;   // kTriGroup: rounds a wave may run ahead

struct FusedPlan;

}  // namespace rb

// The plan object.
struct rb_plan {
    int m = 0, nnz = 0, subst = 0, lv = 0, nblocks = 0, ntri = 0;
    int step = 0;
    int nlevel = 0;
    int device = 0;
    int schedule = RB_SCHED_BLOCKED;
    double preprocess_ms = 0.0;
    // plan build phases (host clock, ms): level analysis, reorder to
    // selector and storage, fused schedule build, schedule tuning
    double phase_ms[4] = {0.0, 0.0, 0.0, 0.0};
    rb::Arena arena;

    // permutation and global levels
    int *d_levelItem = nullptr;  // final permutation for b / x
    int *d_levelPtr = nullptr;   // BFS level boundaries (nlevel+1)
    // permuted CSC (kept for inspection)
    int *d_pcolptr = nullptr, *d_prowidx = nullptr;
    double *d_pval = nullptr;

    // slot space: one slot per (block, block-local row)
    int nslots = 0;
    int *d_rowptr = nullptr;   // nslots+1 entry offsets (off-diagonals only)
    int *d_col = nullptr;      // block-local column of each entry
    double *d_val = nullptr;
    double *d_invd = nullptr;  // 1 / diagonal per permuted row
    int *d_dcsr_rows = nullptr;
    int *d_long_rows = nullptr;
    int ndcsr = 0, nlong = 0;
    // triangular local levels
    int *d_lvptr = nullptr;    // concatenated, sum(blk_m + 1) over tri
    int *d_lvitem = nullptr;   // concatenated local level order (m)
    int *d_nnzlv = nullptr;    // concatenated find_level_nnz (lvptr layout)

    std::vector<rb::Block> blocks;

    // per-solve scratch
    double *d_bperm = nullptr;  // m
    double *d_xperm = nullptr;  // m
    int *d_tickets = nullptr;   // one per triangular block
    int *d_status = nullptr;    // device watchdog flags (0 = healthy)
    int *d_ncol = nullptr;      // [0] columns of the running walk, [1] = 1
    // batch plans: per permuted row, the columns its factor has in the
    // current batch solve (the walk skips the row in later columns); the
    // entry points other than the batch's reset it to "all" before a solve
    int *d_rowcap = nullptr;
    bool rowcap_set = false;    // the batch has filled d_rowcap for this solve
    bool rowcap_all = false;    // d_rowcap holds "every column" everywhere
    unsigned spin_limit = 1u << 22;  // polls before a bounded wait gives up
    int scratch_rhs = 0;

    // fused schedule (built when requested or for the tuner)
    rb::FusedPlan *fused = nullptr;
    int want_schedule = RB_SCHED_AUTO;
    int fused_waves = 0;        // wave cap for the fused partition (0 = auto)
    double tune_ms_blocked = 0.0, tune_ms_fused = 0.0;
    int tune_rhs = 1;           // right-hand sides per tuning solve
    // blocked schedule: the block walk captured as graphs, per value type
    // (FP64, FP32) and grid width (powers of two)
    std::vector<std::pair<int, void *>> graphs[2];
    bool graph_failed = false;
    // columns the permuted b / x (FP64, FP32) and tickets are sized for
    int kcols = 1, kcols32 = 1, kticket = 1;
    // kernel timing (rb_plan_set_timing)
    bool timing = false;
    bool mark_first = false;  // the next main launch opens the timed span
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    // FP32 solve + FP64 refinement (rb_opts.mixed)
    bool mixed = false;
    float *d_val32 = nullptr, *d_invd32 = nullptr;   // slot values in FP32
    float *d_bperm32 = nullptr, *d_xperm32 = nullptr;
    int *d_rptr = nullptr, *d_rcol = nullptr;        // FP64 matrix for the
    double *d_rval = nullptr;                        // residual: CSR, original order
    // refinement scratch for kir right-hand sides (allocated on first use)
    int kir = 0;
    double *d_r = nullptr;                           // residual, FP64
    float *d_r32 = nullptr, *d_d32 = nullptr;        // residual / correction
    unsigned long long *d_norms = nullptr;           // max |b|, max |r|
    double *d_irrel = nullptr;                       // relative residuals
    int *d_irdone = nullptr, *d_iriters = nullptr;   // converged, corrections
    int *d_irleft = nullptr;                         // columns not converged
};

namespace rb {
// Sorted, merged copy of a CSC (rb_canon.hip), allocated in `out`.
rb_status canonicalize_csc(const int *d_colptr, const int *d_rowidx,
                           const double *d_val, int n, int nnz, Arena &out,
                           int **c_colptr, int **c_rowidx, double **c_val,
                           int *c_nnz, hipStream_t s);

rb_status build_plan(rb_plan *p, const int *d_colptr, const int *d_rowidx,
                     const double *d_val, hipStream_t s);
rb_status solve_blocked(rb_plan *p, const double *d_b, int ldb, double *d_x,
                        int ldx, int nrhs, hipStream_t s);
rb_status solve_f32ir(rb_plan *p, const double *d_b, int ldb, double *d_x,
                      int ldx, int nrhs, double tol, int max_iter,
                      hipStream_t s, int *iters, double *relres);
// per-block kernel choice of the blocked walk by timing each candidate on
// the plan's own buffers (rb_solve.hip); mode: rb_opts.block_kernels
// budget_ms > 0: stop timing once the timed blocks' best times add up to
// more than budget_ms (the blocked walk cannot be faster; the remaining
// blocks keep the rule) and set *cut
rb_status tune_blocks(rb_plan *p, int mode, hipStream_t s,
                      double budget_ms = 0.0, bool *cut = nullptr);
// fused schedule (rb_fused.hip): build_fused_start copies the permuted
// CSR structure (d_prptr, d_ccol) and the level pointers to the host and
// starts the host passes on a thread; build_fused_finish(job, true) then
// uploads the schedule on the same stream (d_cval, the plan's levelItem
// and invd must be ready by then) and joins the thread, or with false
// abandons it.  Between the two calls the thread makes no device call and
// reads only the host copies and plan fields that stay fixed (m, nnz,
// nlevel, subst, want_schedule, tune_rhs); its allocations from the plan's
// arena and its work on the stream happen inside build_fused_finish, while
// the caller waits.
struct FusedJob;
rb_status build_fused_start(rb_plan *p, const int *d_prptr, const int *d_ccol,
                            const double *d_cval, int Wmax, hipStream_t s,
                            FusedJob **job);
rb_status build_fused_finish(FusedJob *job, bool upload);
rb_status solve_fused(rb_plan *p, const double *d_b, int ldb, double *d_x,
                      int ldx, int nrhs, hipStream_t s);
void fused_destroy(rb_plan *p);
// the same solves on caller-provided per-solve buffers (ws, ws_bytes)
// instead of the plan's: the blocked walk needs blocked_ws_bytes(p) (one
// column at a time, eager launches), the fused schedule 16 bytes per slot
// (fused_slots(p)) and column of a launch; neither allocates
size_t blocked_ws_bytes(const rb_plan *p);
size_t fused_slots(const rb_plan *p);
rb_status solve_blocked_ws(rb_plan *p, const double *d_b, int ldb,
                           double *d_x, int ldx, int nrhs, hipStream_t s,
                           void *ws, size_t ws_bytes);
rb_status solve_fused_ws(rb_plan *p, const double *d_b, int ldb, double *d_x,
                         int ldx, int nrhs, hipStream_t s, double *ws,
                         size_t ws_bytes);
// One correction of the FP32 refinement on the fused schedule (FP32
// copies of its tables, made on first use): FP32 solves of the FP64
// residual d_r (rounded to FP32 on the way in), added to the FP64 d_x in
// every column c with done[c] == 0; does nothing once *left == 0 (left may
// be null)
rb_status solve_fused32(rb_plan *p, const double *d_r, int ldr, double *d_x,
                        int ldx, int nrhs, hipStream_t s, const int *left,
                        const int *done);
// events around a solve's main launches when timing is on: from the start
// of the first (the entry point sets mark_first) to the end of the last
inline void mark_start(rb_plan *p, hipStream_t s) {
    if (p->timing && p->mark_first) (void)hipEventRecord(p->ev0, s);
    p->mark_first = false;
}
inline void mark_end(rb_plan *p, hipStream_t s) {
    if (p->timing) (void)hipEventRecord(p->ev1, s);
}
void destroy_graph(rb_plan *p);
// at the start of a solve call: a batch plan's row caps back to "every
// column" unless the batch has just set them for this call
void rowcap_for_call(rb_plan *p, hipStream_t s);
// frees the right-hand-side and refinement scratch grown beyond the plan's
// own (one column)
void free_solve_scratch(rb_plan *p);
void free_ir(rb_plan *p);
// after a watchdog trip: every slot-order x buffer of the fused schedule is
// filled with the sentinel again before its next use
void fused_rearm(rb_plan *p);
int fused_max_waves();
}  // namespace rb
