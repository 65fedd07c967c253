/*
 * rb_cpu.c -- the CPU baseline: the reference's recursive-block SpTRSV
 * solve run on the host cores with OpenMP.
 *
 * TEST INFRASTRUCTURE ONLY, like rb_oracle.c: only bench.py's cpu_baseline
 * leg and tests/ load it.  The product library never links or calls it.
 *
 * The reference solves on the GPU only (L_calculate / U_calculate,
 * R/recblocking_calculate.h:15-325); its host side only preprocesses.  This
 * file times that same algorithm on the CPU: the level-ordered matrix and
 * partition from the oracle (rb_oracle.c), blocks in the reference's order
 * tri0, sq1, tri1, ..., a triangular block solved level by level with the
 * rows of a level shared out over the threads (the level-set method of
 * R/utils_sptrsv_cuda.h:116-200 as a CPU loop), or row by row in position
 * order when its levels are too narrow for a barrier per level to pay, and
 * a square block as a row-parallel SpMV with b -= y fused
 * (R/utils_spmv_cuda.h:60-212).  Each block keeps its own compact CSR
 * (block-local columns), so the solve streams the matrix once.
 *
 * Built -O3 -march=native -fopenmp by oracle/Makefile.
 */
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* levels narrower than this on average: the block runs serially (one OpenMP
 * barrier per level costs more than a level's work) */
#define CPU_MIN_LEVEL_WIDTH 256

typedef struct {
    int kind;      /* 0 triangular, 1 square */
    int up, down, left, right;
    int *rp;       /* blk_m + 1, entries of the block's rows (off-diagonal) */
    int *ci;       /* block-local column ids */
    double *v;
    double *invd;  /* triangular: 1 / diagonal per block row */
    int nlv;       /* triangular: local levels (0 = run serially) */
    int *lvptr, *lvitem;
} cpu_block;

typedef struct {
    int m, nblk, subst;
    cpu_block *b;
    int *levelItem;
    double *bp, *xp;
} cpu_plan;

/* Build from the level-ordered CSC (orc_reorder), the partition
 * (orc_partition) and the triangular blocks' local level tables
 * (orc_analyze, concatenated: lvptr blocks of blk_m + 1 entries, lvitem
 * blocks of blk_m).  Returns NULL on a missing diagonal. */
cpu_plan *cpu_plan_create(const int *colptr, const int *rowidx,
                          const double *val, int m, int lv, int subst,
                          const int *up, const int *down, const int *left,
                          const int *right, const int *nlv,
                          const int *lvptr_cat, const int *lvitem_cat,
                          const int *levelItem)
{
    int nblk = 2 * (1 << lv) - 1, i, j, k;
    cpu_plan *p = (cpu_plan *)calloc(1, sizeof(cpu_plan));
    int *rp = (int *)calloc((size_t)m + 1, sizeof(int));
    int *ci, *fill;
    double *vv;
    int nnz = colptr[m];
    long lvp_off = 0, item_off = 0;
    p->m = m;
    p->nblk = nblk;
    p->subst = subst;
    p->b = (cpu_block *)calloc((size_t)nblk, sizeof(cpu_block));
    p->levelItem = (int *)malloc(sizeof(int) * (size_t)(m > 0 ? m : 1));
    memcpy(p->levelItem, levelItem, sizeof(int) * (size_t)m);
    p->bp = (double *)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    p->xp = (double *)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    /* CSR of the whole level-ordered matrix, columns ascending per row */
    ci = (int *)malloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
    vv = (double *)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
    fill = (int *)malloc(sizeof(int) * (size_t)(m > 0 ? m : 1));
    for (k = 0; k < nnz; k++) rp[rowidx[k] + 1]++;
    for (i = 0; i < m; i++) rp[i + 1] += rp[i];
    memcpy(fill, rp, sizeof(int) * (size_t)m);
    for (j = 0; j < m; j++)
        for (k = colptr[j]; k < colptr[j + 1]; k++) {
            int r = rowidx[k], q = fill[r]++;
            ci[q] = j;
            vv[q] = val[k];
        }
    for (i = 0; i < nblk; i++) {
        cpu_block *B = &p->b[i];
        int bm = down[i] - up[i], q = 0;
        B->kind = i % 2;
        B->up = up[i];
        B->down = down[i];
        B->left = left[i];
        B->right = right[i];
        B->rp = (int *)malloc(sizeof(int) * (size_t)(bm + 1));
        /* count, then copy the entries with columns in [left, right) (for a
         * triangular block: its own columns without the diagonal) */
        B->rp[0] = 0;
        for (j = 0; j < bm; j++) {
            int r = up[i] + j, c = 0;
            for (k = rp[r]; k < rp[r + 1]; k++)
                if (ci[k] >= left[i] && ci[k] < right[i] && ci[k] != r) c++;
            B->rp[j + 1] = B->rp[j] + c;
        }
        B->ci = (int *)malloc(sizeof(int) * (size_t)(B->rp[bm] > 0 ? B->rp[bm] : 1));
        B->v = (double *)malloc(sizeof(double) * (size_t)(B->rp[bm] > 0 ? B->rp[bm] : 1));
        for (j = 0; j < bm; j++) {
            int r = up[i] + j;
            for (k = rp[r]; k < rp[r + 1]; k++)
                if (ci[k] >= left[i] && ci[k] < right[i] && ci[k] != r) {
                    B->ci[q] = ci[k] - left[i];
                    B->v[q] = vv[k];
                    q++;
                }
        }
        if (B->kind == 0) {
            B->invd = (double *)malloc(sizeof(double) * (size_t)(bm > 0 ? bm : 1));
            for (j = 0; j < bm; j++) {
                int r = up[i] + j;
                double d = 0.0;
                for (k = rp[r]; k < rp[r + 1]; k++)
                    if (ci[k] == r) d = vv[k];
                if (d == 0.0) {
                    free(ci); free(vv); free(fill); free(rp);
                    return NULL;  /* caller frees nothing: leaks on error */
                }
                B->invd[j] = 1.0 / d;
            }
            B->nlv = nlv[i];
            if (bm > 0 && B->nlv > 0 && bm / B->nlv >= CPU_MIN_LEVEL_WIDTH) {
                B->lvptr = (int *)malloc(sizeof(int) * (size_t)(B->nlv + 1));
                B->lvitem = (int *)malloc(sizeof(int) * (size_t)bm);
                memcpy(B->lvptr, lvptr_cat + lvp_off, sizeof(int) * (size_t)(B->nlv + 1));
                memcpy(B->lvitem, lvitem_cat + item_off, sizeof(int) * (size_t)bm);
            } else {
                B->nlv = 0;  /* serial */
            }
            lvp_off += bm + 1;
            item_off += bm;
        }
    }
    free(ci);
    free(vv);
    free(fill);
    free(rp);
    return p;
}

void cpu_plan_destroy(cpu_plan *p)
{
    int i;
    if (!p) return;
    for (i = 0; i < p->nblk; i++) {
        free(p->b[i].rp);
        free(p->b[i].ci);
        free(p->b[i].v);
        free(p->b[i].invd);
        free(p->b[i].lvptr);
        free(p->b[i].lvitem);
    }
    free(p->b);
    free(p->levelItem);
    free(p->bp);
    free(p->xp);
    free(p);
}

static inline double row_dot(const cpu_block *B, int j, const double *x)
{
    double s = 0.0;
    int k;
    for (k = B->rp[j]; k < B->rp[j + 1]; k++) s += B->v[k] * x[B->ci[k]];
    return s;
}

/* One solve of T x = b (original order in and out) on nthreads threads
 * (0 = the OpenMP default). */
void cpu_plan_solve(cpu_plan *p, const double *b, double *x, int nthreads)
{
    const int m = p->m;
    double *bp = p->bp, *xp = p->xp;
    const int *li = p->levelItem;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        int i, j, L;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (i = 0; i < m; i++) bp[i] = b[li[i]];
        for (i = 0; i < p->nblk; i++) {
            const cpu_block *B = &p->b[i];
            const int bm = B->down - B->up;
            if (B->kind == 1) {
                const double *xs = xp + B->left;
                double *bs = bp + B->up;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
                for (j = 0; j < bm; j++) bs[j] -= row_dot(B, j, xs);
            } else if (B->nlv > 0) {
                const double *bs = bp + B->up;
                double *xs = xp + B->up;
                for (L = 0; L < B->nlv; L++) {
                    int a;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
                    for (a = B->lvptr[L]; a < B->lvptr[L + 1]; a++) {
                        const int r = B->lvitem[a];
                        xs[r] = (bs[r] - row_dot(B, r, xs)) * B->invd[r];
                    }
                }
            } else {
#ifdef _OPENMP
#pragma omp single
#endif
                {
                    const double *bs = bp + B->up;
                    double *xs = xp + B->up;
                    if (p->subst == 0)
                        for (j = 0; j < bm; j++)
                            xs[j] = (bs[j] - row_dot(B, j, xs)) * B->invd[j];
                    else
                        for (j = bm - 1; j >= 0; j--)
                            xs[j] = (bs[j] - row_dot(B, j, xs)) * B->invd[j];
                }
            }
        }
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (i = 0; i < m; i++) x[li[i]] = xp[i];
    }
}

/* Threads an OpenMP region gets by default on this host. */
int cpu_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
