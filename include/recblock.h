/*
 * recblock.h -- C ABI of the MI355X (gfx950) recursive-block SpTRSV library
 * (librecblock.so).
 *
 * All pointers named d_* are device pointers (hipMalloc'd, device 0 unless
 * the caller made another device current); all others are host pointers.
 * Indices are 0-based int32; values are IEEE double.  A matrix is square
 * (m == n) compressed sparse column (CSC):
 *   forward  (substitution 0): lower-triangular L, row ids ascending in each
 *            column, so the diagonal is the first entry of every column;
 *   backward (substitution 1): upper-triangular U, row ids ascending, the
 *            diagonal is the last entry of every column.
 * No function aborts: every entry point returns an rb_status (the legacy
 * entry points, which return void as the reference's do, leave their status
 * in rb_last_status() / rb_last_error()).
 */
#ifndef RECBLOCK_H
#define RECBLOCK_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    RB_OK = 0,
    RB_ERR_INVALID_ARG = 1,   /* null pointer, bad size, lv out of range */
    RB_ERR_NOT_SQUARE = 2,    /* m != n */
    RB_ERR_BAD_STRUCTURE = 3, /* colptr not monotone, row id out of range,
                                 entry outside the triangle, unsorted or
                                 duplicate row ids in a column */
    RB_ERR_MISSING_DIAG = 4,  /* a column without its diagonal entry (the
                                 reference's level analysis hangs here) */
    RB_ERR_ZERO_DIAG = 5,     /* a zero diagonal value */
    RB_ERR_HIP = 6,           /* a HIP runtime call failed */
    RB_ERR_NO_MEMORY = 7,     /* device allocation failed */
    RB_ERR_INTERNAL = 8,
    RB_ERR_NOT_CONVERGED = 9   /* FP32 refinement missed its tolerance */
} rb_status;

#define RB_SUBSTITUTION_FORWARD 0
#define RB_SUBSTITUTION_BACKWARD 1

/* Execution schedules (rb_opts.schedule). */
#define RB_SCHED_AUTO 0      /* the library's tuner picks (default) */
#define RB_SCHED_BLOCKED 1   /* per-block kernels in the reference selector's
                                order: TRSV, SpMV, TRSV, ... */
#define RB_SCHED_FUSED 2     /* one persistent dataflow launch per solve */

typedef struct {
    int schedule;      /* RB_SCHED_* */
    int device_id;     /* device whose lv auto-threshold applies (lv = -1) */
    int fused_waves;   /* fused schedule: cap on consumer waves (0 = auto) */
    int no_graph;      /* blocked schedule: 1 = launch the block walk kernel
                          by kernel instead of replaying a captured graph */
    int mixed;         /* 1 = also keep FP32 copies of the block values and
                          the FP64 matrix for residuals: enables
                          rb_solve_f32ir */
    int tune_rhs;      /* RB_SCHED_AUTO: right-hand sides per call the
                          schedule is tuned for (0 or 1: single) */
    int canonicalize;  /* 1 = accept row ids in any order inside a column
                          and repeated entries (summed in input order): the
                          plan is built from the sorted, merged matrix and
                          its nnz is the merged count.  0 (default) =
                          RB_ERR_BAD_STRUCTURE for such input, as before.
                          No counterpart in the reference, which needs
                          sorted, unique row ids. */
    int watchdog_spins; /* 0 = default bound (about 4M polls, seconds); > 0:
                          every bounded wait of a solve gives up after this
                          many polls.  A small value lets a test drive the
                          failure path (x all NaN, rb_plan_check reporting
                          RB_ERR_INTERNAL) while every kernel exits
                          normally.  No counterpart in the reference, whose
                          waits are unbounded. */
    int block_kernels;  /* blocked schedule, kernel per block: 0 (default) =
                          time the candidates on each block at plan build
                          and keep the fastest (triangular: the reference
                          selector's kernel, the sync-free grid kernel, the
                          one-workgroup kernel when its dependencies fit its
                          LDS ring; square: 1..64 lanes per row); 1 = the
                          reference selector's kernels (method) and the
                          lanes-per-row rule; 2 / 3 = every triangular block
                          with the sync-free / one-workgroup kernel (where
                          it applies).  Every choice gives the same x, bit
                          for bit.  The reference's selector
                          (R/recblocking_solver_cuda.h:176-367,441-546) fixes
                          the kernel by thresholds; its decisions are still
                          reported by rb_plan_blocks. */
} rb_opts;

typedef struct rb_plan rb_plan;

/* ---------------------------------------------------------------------
 * Legacy entry points (same C signatures and argument meaning as the
 * reference; VALUE_TYPE is double).
 * ------------------------------------------------------------------- */

/* Replaces recblocking_solver_cuda (R/recblocking_solver_cuda.h:21-32,
 * called at R/main.cu:278).  Device CSC of L (forward) or U (backward);
 * solves once and then BENCH_REPEAT (200, see rb_set_bench_repeat) more
 * times; *cal_time = mean solve time in ms, *preprocess_time = plan build
 * time in ms (set for both directions; the reference sets it for forward
 * only).  lv >= 1 as in the reference; lv = -1 applies the auto rule of
 * R/main.cu:245-254 for device 0.  d_x receives x in the original order.
 * Difference: d_b is left unchanged (the reference overwrites it with the
 * permuted, partially updated right-hand side). */
void recblocking_solver_cuda(int *d_cscColPtrTR, int *d_cscRowIdxTR,
                             double *d_cscValTR, int m, int n, int nnzTR,
                             double *d_x, double *d_b, int substitution,
                             int lv, double *cal_time,
                             double *preprocess_time);

/* Replaces recblocking_solver (R/recblocking_solver.h:16-28).  Host CSC,
 * host b (m x rhs, column-major) and x (n x rhs).  x_ref is unused, as in
 * the reference.  All rhs columns are solved (the reference sizes buffers
 * for rhs but solves one).  *cal_time = mean ms per solve of all rhs. */
void recblocking_solver(int *cscColPtr, int *cscRowIdx, double *cscVal,
                        int m, int n, int nnz, double *b, double *x,
                        double *x_ref, int rhs, int lv, int substitution,
                        double *cal_time);

/* Status of the last legacy call on this thread, and a message. */
rb_status rb_last_status(void);
const char *rb_last_error(void);

/* Number of timed repeats of the legacy entry points (reference default
 * BENCH_REPEAT = 200, R/common.h:13-15).  Returns the previous value. */
int rb_set_bench_repeat(int repeat);

/* The lv auto rule of R/main.cu:245-254. */
int rb_lv_auto(int m, int device_id);

/* ---------------------------------------------------------------------
 * Plan API (split of the legacy calls into preprocessing and solve;
 * the reference's internal split is mat_preprocessing / L_preprocessing /
 * U_preprocessing then L_calculate / U_calculate,
 * R/recblocking_preprocess.h:15,298,387,969 and
 * R/recblocking_calculate.h:15,170).
 * ------------------------------------------------------------------- */

void rb_opts_default(rb_opts *opts);

/* Build a plan from a device CSC.  Runs the level analysis, level-order
 * permutation, recursive partition, block extraction, selector statistics
 * and storage build on the device; the inputs are only read.  lv >= 0, or
 * -1 for the auto rule.  stream may be 0 (the null stream).  The call
 * synchronises on stream. */
rb_status rb_plan_create(const int *d_colptr, const int *d_rowidx,
                         const double *d_val, int m, int n, int nnz,
                         int substitution, int lv, const rb_opts *opts,
                         void *stream, rb_plan **plan);

/* Solve op(T) X = B for nrhs right-hand sides: d_b is m x nrhs
 * column-major with leading dimension ldb >= m, d_x likewise with ldx.
 * Stream-ordered and asynchronous (no host synchronisation); d_b is not
 * modified.  One plan must not run two solves at once; distinct plans may. */
rb_status rb_solve(rb_plan *plan, const double *d_b, int ldb, double *d_x,
                   int ldx, int nrhs, void *stream);

rb_status rb_plan_destroy(rb_plan *plan);

/* Synchronise stream and report the device-side watchdog: RB_OK, or
 * RB_ERR_INTERNAL when a solve's bounded spin-wait gave up.  From the solve
 * whose wait gave up until this call, every solve on the plan writes quiet
 * NaN to all of x (never a finite wrong answer); this call then clears the
 * flag and re-arms the plan's per-solve state, so later solves run
 * normally.  Otherwise RB_ERR_NOT_CONVERGED when an rb_solve_f32ir call
 * without a report left a column short of its tolerance (the flag is
 * cleared).  No counterpart in the reference, whose spin loops have no
 * bound. */
rb_status rb_plan_check(rb_plan *plan, void *stream);

/* Plan summary for inspection and tests. */
typedef struct {
    int m, nnz, substitution, lv;
    int nblocks;          /* 2^(lv+1) - 1 */
    int nlevel;           /* global level count */
    int schedule;         /* schedule a solve uses (RB_SCHED_BLOCKED/FUSED) */
    double preprocess_ms; /* plan build time (phases 0-2 below) */
    size_t device_bytes;  /* device memory held by the plan */
    double phase_ms[4];   /* host-clock ms of the build's phases: level
                             analysis; reorder, partition, blocks, selector
                             and storage; fused schedule build; schedule
                             tuning (timed solves, not in preprocess_ms) */
    double tune_ms[2];    /* warm solve time of the blocked / fused schedule
                             at tune_rhs columns when RB_SCHED_AUTO timed
                             both, else 0 */
} rb_plan_info;

rb_status rb_plan_info_get(const rb_plan *plan, rb_plan_info *info);

/* Per-block selector inputs and decisions, host arrays of nblocks entries
 * (any pointer may be NULL).  Triangular blocks are even indices, square
 * blocks odd, in the reference's block order.  method: triangular 0..3,
 * square 0..3 or -1 for an empty block (same meaning as the reference's
 * SpTRSV_block.method / SpMV_block.method).  For triangular blocks nnzr =
 * blk_nnz / blk_m; for square blocks (blk_nnz - longlen) / m_new, -1 if
 * m_new == 0. */
rb_status rb_plan_blocks(const rb_plan *plan, int *blk_m, int *blk_n,
                         int *blk_nnz, int *up, int *down, int *left,
                         int *right, int *method, int *nlv, int *nnzr,
                         int *m_new, int *lenmax, int *longrow, int *longlen,
                         double *empty_ratio);

/* Kernels the blocked schedule runs per block (host arrays of nblocks,
 * any may be NULL): kern (triangular: 0 the reference selector's kernel,
 * 1 sync-free grid, 2 one workgroup; square: 0), lanes (square: lanes per
 * row; triangular: 0), tuned_ms (the chosen kernel's time when tuned, else
 * 0).  No counterpart in the reference. */
rb_status rb_plan_block_kernels(const rb_plan *plan, int *kern, int *lanes,
                                double *tuned_ms);

/* Global level analysis (host arrays): levelItem[m] is the permutation that
 * maps level-order position i to original row levelItem[i] (reversed for
 * backward, as the reference leaves it), levelPtr[nlevel+1] the BFS level
 * boundaries (BFS order, before any reversal). */
rb_status rb_plan_levels(const rb_plan *plan, int *levelItem, int *levelPtr);

/* Local level analysis of triangular block t (0 <= t < 2^lv): lvptr[nlv+1],
 * lvitem[blk_m] (block-local ids in BFS discovery order) and
 * nnz_lv[nlv] (find_level_nnz semantics, R/utils_cuda.h:781-796). */
rb_status rb_plan_tri_levels(const rb_plan *plan, int t, int *lvptr,
                             int *lvitem, int *nnz_lv);

/* Contents of block i as a block-local CSR (rows 0..blk_m-1, columns
 * relative to left[i]); the diagonal of a triangular block is included.
 * rowptr[blk_m+1], colidx[blk_nnz], val[blk_nnz] (host). */
rb_status rb_plan_block_csr(const rb_plan *plan, int i, int *rowptr,
                            int *colidx, double *val);

/* Trace of the fused schedule's last solve, recorded when the environment
 * variable RB_FUSED_TRACE is set at plan build: 16 counters per consumer
 * wave (wall-clock ticks of 10 ns): [1] consumer waiting for global
 * dependencies, [2] consumer total, [3] steps, [7] resolver waiting for
 * FIFO space, [8] polling, [9] batches, [10] re-poll rounds (others 0).
 * *nwaves receives the wave count (0 when the plan has no fused schedule).
 * counters may be NULL.  No counterpart in the reference (which times
 * whole solves only). */
rb_status rb_plan_trace(const rb_plan *plan, unsigned long long *counters,
                        int *nwaves);

/* FP32 solves with FP64 iterative refinement (plan created with
 * rb_opts.mixed = 1).  Per right-hand side: x = 0, r = b; repeat
 * { d = solve((float) r) in FP32 on the plan's schedule (the blocked
 * walk's kernels on FP32 copies of the block values, or the fused kernel's
 * FP32 form, whose FP32 tables and buffers are allocated on the first
 * call); x += d in FP64; r = b - A x in FP64 on the device } until
 * max_i |r_i| <= tol * max_i |b_i| or max_iter corrections.
 * FP64 is used only for the residual and the update of x.  d_b / d_x as in
 * rb_solve.  The loop runs on the device, up to 32 columns at a time;
 * each column stops correcting once it has converged.  With iters or
 * relres given, *iters receives the most corrections any column needed and
 * *relres the largest final max|r| / max|b|; the call synchronises on
 * stream after the second correction and each later one (once for the
 * usual two) and returns RB_ERR_NOT_CONVERGED when a column missed tol (x
 * then holds the last iterate; a column whose residual holds a NaN never
 * converges, and its relres is reported as inf).  With both NULL the call does not
 * synchronise: max_iter corrections are queued and a column that missed
 * tol sets the plan's status, which rb_plan_check reports as
 * RB_ERR_NOT_CONVERGED.  No counterpart in the reference (FP64 only). */
rb_status rb_solve_f32ir(rb_plan *plan, const double *d_b, int ldb,
                         double *d_x, int ldx, int nrhs, double tol,
                         int max_iter, void *stream, int *iters,
                         double *relres);

/* rb_solve without device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate -- code under stream or graph
 * capture, serving loops and harnesses that forbid hipMalloc.  Same
 * arguments, results and stream ordering as rb_solve (results are bitwise
 * equal), but the per-solve scratch (permuted b and x and tickets of the
 * blocked schedule, slot-order x and b/diag buffers of the fused schedule)
 * lives in one fixed device array of rb_noalloc_limit() bytes (64 MiB) per
 * device, so a plan whose single column needs more is rejected with
 * RB_ERR_INVALID_ARG; columns go through in groups that fit.  The call
 * allocates nothing and never synchronises.  Every call on a device shares
 * that array: two calls that overlap on one device (other streams, other
 * threads) race and must be ordered by the caller. */
rb_status rb_solve_noalloc(rb_plan *plan, const double *d_b, int ldb,
                           double *d_x, int ldx, int nrhs, void *stream);
size_t rb_noalloc_limit(void);

/* ---------------------------------------------------------------------
 * Batches: many independent factors of uneven size, preprocessed as one
 * block-diagonal plan, so that one persistent launch (fused schedule) or
 * one captured block walk solves every factor and right-hand side.  No
 * counterpart in the reference (one factor, one right-hand side).
 * ------------------------------------------------------------------- */
typedef struct rb_batch rb_batch;

/* nfactors CSC factors (host arrays of device pointers and sizes), all of
 * the same substitution; lv and opts apply to the union (lv = -1: auto
 * rule on the total size).  Synchronises on stream. */
rb_status rb_batch_create(int nfactors, const int *const *d_colptr,
                          const int *const *d_rowidx,
                          const double *const *d_val, const int *m,
                          const int *nnz, int substitution, int lv,
                          const rb_opts *opts, void *stream, rb_batch **batch);

/* Factor f solves nrhs[f] right-hand sides d_b[f] (m[f] x nrhs[f],
 * leading dimension ldb[f]) into d_x[f] (ldx[f]); nrhs[f] may be 0.
 * Host arrays of device pointers.  Stream-ordered; the per-call tables are
 * staged through pinned memory, so the arrays may be reused on return. */
rb_status rb_batch_solve(rb_batch *batch, const double *const *d_b,
                         const int *ldb, double *const *d_x, const int *ldx,
                         const int *nrhs, void *stream);

/* The same with FP32 solves and FP64 refinement (opts->mixed = 1 at
 * creation); see rb_solve_f32ir. */
rb_status rb_batch_solve_f32ir(rb_batch *batch, const double *const *d_b,
                               const int *ldb, double *const *d_x,
                               const int *ldx, const int *nrhs, double tol,
                               int max_iter, void *stream, int *iters,
                               double *relres);

/* The union plan (for rb_plan_info_get, timing and inspection); owned by
 * the batch. */
rb_plan *rb_batch_plan(rb_batch *batch);

rb_status rb_batch_destroy(rb_batch *batch);

/* Kernel timing for measurement harnesses.  While enabled, every solve
 * records HIP events on its stream around its main launches: the fused
 * kernel (fused schedule) or the captured block walk (blocked schedule),
 * without the passes that move b and x between orders, from the start of
 * the solve's first main launch to the end of its last (a solve of more
 * columns than one launch or walk takes makes several; the passes between
 * them are then inside the span; for rb_solve_f32ir the span covers all of
 * its walks).  rb_plan_kernel_ms waits for the last solve's span and
 * returns its duration in milliseconds.  No counterpart in the
 * reference. */
rb_status rb_plan_set_timing(rb_plan *plan, int enable);
rb_status rb_plan_kernel_ms(rb_plan *plan, double *ms);

/* The level-ordered (permuted) matrix as CSC, host arrays of m+1 / nnz. */
rb_status rb_plan_permuted_csc(const rb_plan *plan, int *colptr, int *rowidx,
                               double *val);

/* ---- Matrix Market input and the benchmark driver's factor (host only;
 * no device is touched).  These restate the reference program's host side,
 * so that a maintainer's harness can read the same .mtx files and build the
 * same triangular factor before calling the solver. */

/* Reader, the reference's mmio_info / mmio_data
 * (recblock-sptrsv/mmio_highlevel.h:7-137, 139-296): coordinate files of
 * real / integer / pattern / complex (real part) values; symmetric and
 * hermitian files are mirrored.  rb_mmio_info gives the CSR sizes (nnz counts
 * mirrored entries); rb_mmio_data fills caller-allocated CSR arrays (m + 1,
 * nnz, nnz).  Columns inside a row keep the file's order (mirrored pairs
 * interleaved as read), exactly as the reference builds them.  Returns 0,
 * -1 (cannot open / null argument), -2 (banner: not a coordinate matrix of a
 * known type), -4 (size line), -5 (an entry is missing, malformed or out of
 * range -- where the reference writes out of bounds). */
int rb_mmio_info(int *m, int *n, int *nnz, int *is_symmetric, const char *path);
int rb_mmio_data(int *csr_rowptr, int *csr_colidx, double *csr_val,
                 const char *path);

/* The triangular factor and right-hand sides of the reference's driver
 * (recblock-sptrsv/main.cu:150-228): from the CSR of A (m rows), forward
 * keeps entries with column < row (row order) and appends the diagonal;
 * backward puts the diagonal first and keeps entries with column > row.
 * reference_values = 1 reproduces the driver's rewrite: every kept entry is
 * rand() % 10 + 1 and the diagonal 1.0 (C library rand() after
 * srand(seed)); otherwise the file's values are kept and the row's diagonal
 * entries (summed) must exist (RB_ERR_MISSING_DIAG).  Then, for nrhs > 0,
 * x_ref[r*m + i] = rand() % 10 + 1 and b = T * x_ref per column (T's row
 * order).  Call with tri_colidx = NULL first to get tri_nnz (tri_rowptr has
 * m + 1 entries); x_ref and b have m * nrhs entries.  Output is CSR: pass it
 * through a transposition for the solver's CSC. */
rb_status rb_mmio_triangle(int m, const int *rowptr, const int *colidx,
                           const double *val, int substitution,
                           int reference_values, unsigned seed, int nrhs,
                           int *tri_rowptr, int *tri_colidx, double *tri_val,
                           int *tri_nnz, double *x_ref, double *b);

#ifdef __cplusplus
}
#endif

#endif /* RECBLOCK_H */
