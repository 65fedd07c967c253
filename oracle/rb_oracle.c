/*
 * rb_oracle.c -- CPU restatement of the reference recursive-block SpTRSV
 * pipeline (recblock-sptrsv).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product library (librecblock.so) never links, loads or calls it.
 *
 * Every function below cites the reference file:line it restates.  The
 * reference's own host code cannot be compiled here (its headers include
 * cusparse.h / cuda_runtime.h, which this image lacks), so each function is
 * pinned instead by hand-worked fixtures, closed forms and invariants in
 * tests/test_oracle.py.
 *
 * Conventions (shared with the reference):
 *   - CSC input: colptr[n+1], rowidx[nnz], val[nnz]; 0-based; m == n.
 *   - substitution 0 = forward (lower triangular L, diagonal first in each
 *     column), 1 = backward (upper triangular U, diagonal last).
 *   - block sequence: tri0, sq1, tri1, sq2, ... (2^lv triangular blocks and
 *     2^lv - 1 square blocks, in-order traversal of the recursive bisection).
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>

#define ORC_LONGROW_THRESHOLD 2048 /* R/utils_spmv_cuda.h:9 */

/* ------------------------------------------------------------------ A1 */
/* R/main.cu:245-254: smallest li >= 1 with m / 2^(li+1) < threshold. */
int orc_lv_auto(int m, int device_id)
{
    int li;
    double thr = device_id == 0 ? 92160.0 : 58880.0;
    for (li = 1; li <= 100; li++)
        if ((double)m / pow(2.0, (double)(li + 1)) < thr)
            break;
    return li;
}

/* ------------------------------------------------------------------ A3 */
/* R/utils.h:96-113 (in-place exclusive scan). */
static void orc_exclusive_scan(int *a, int len)
{
    int i, old, nv;
    if (len <= 1) {
        if (len == 1) a[0] = 0;
        return;
    }
    old = a[0];
    a[0] = 0;
    for (i = 1; i < len; i++) {
        nv = a[i];
        a[i] = old + a[i - 1];
        old = nv;
    }
}

/* R/tranpose.h:7-54: histogram -> exclusive scan -> serial insertion in
 * row order.  Input is CSR-like (m "rows" with ptr[m+1]); output is the
 * transpose with n "rows" (out_ptr[n+1]).  Within each output row the
 * entries come out in ascending input-row order.  val/out_val may be NULL. */
void orc_transpose(int m, int n, int nnz, const int *ptr, const int *idx,
                   const double *val, int *out_idx, int *out_ptr,
                   double *out_val)
{
    int i, j;
    int *incr = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    memset(out_ptr, 0, sizeof(int) * (size_t)(n + 1));
    for (i = 0; i < nnz; i++)
        out_ptr[idx[i]]++;
    orc_exclusive_scan(out_ptr, n + 1);
    memcpy(incr, out_ptr, sizeof(int) * (size_t)(n + 1));
    for (i = 0; i < m; i++) {
        for (j = ptr[i]; j < ptr[i + 1]; j++) {
            int c = idx[j];
            out_idx[incr[c]] = i;
            if (val && out_val)
                out_val[incr[c]] = val[j];
            incr[c]++;
        }
    }
    free(incr);
}

/* ------------------------------------------------------------------ A2 */
/* R/findlevel.h:7-83 (host) and R/utils_cuda.h:29-97 (device, identical):
 * Kahn BFS in *discovery order*.  indegree = CSR row count including the
 * diagonal; roots are rows with indegree 1 in ascending id; each level's
 * nodes are visited in order, their CSC columns in storage order, and a row
 * is appended when its indegree drops to 1.
 * Returns 0, or -1 where the reference would spin forever (a round that
 * appends nothing while rows remain: missing diagonal or a cycle). */
int orc_findlevel(const int *colptr, const int *rowidx, const int *rowptr,
                  int m, int *nlevel, int *levelPtr, int *levelItem)
{
    int i, j, lvi, ptr = 0;
    int *indeg = (int *)malloc(sizeof(int) * (size_t)(m > 0 ? m : 1));
    for (i = 0; i < m; i++)
        indeg[i] = rowptr[i + 1] - rowptr[i];
    levelPtr[0] = 0;
    for (i = 0; i < m; i++)
        if (indeg[i] == 1)
            levelItem[ptr++] = i;
    levelPtr[1] = ptr;
    lvi = 1;
    while (levelPtr[lvi] != m) {
        int before = ptr;
        for (i = levelPtr[lvi - 1]; i < levelPtr[lvi]; i++) {
            int node = levelItem[i];
            for (j = colptr[node]; j < colptr[node + 1]; j++) {
                int v = rowidx[j];
                indeg[v]--;
                if (indeg[v] == 1)
                    levelItem[ptr++] = v;
            }
        }
        if (ptr == before) { /* the reference hangs here */
            free(indeg);
            *nlevel = lvi;
            return -1;
        }
        lvi++;
        levelPtr[lvi] = ptr;
    }
    *nlevel = lvi;
    free(indeg);
    return 0;
}

/* R/utils_reordering.h:102-210 (host) / R/utils_cuda.h:706-779 (device):
 * symmetric level-order permutation.  Forward: new column i = old column
 * levelItem[i]; backward: new column i = old column levelItem[n-1-i].
 * Row ids go through the inverse permutation (forward) or
 * perm[levelItem[i]] = m-1-i (backward).  Entry order inside each column is
 * kept.  On return levelItem is the final permutation used for b/x
 * (reversed for backward, R/utils_reordering.h:189-197), and
 * nlevel/levelPtr are the global BFS levels (which the reference computes
 * and then discards). */
int orc_reorder(const int *colptr, const int *rowidx, const double *val,
                int m, int n, int nnz, int subst, int *colptr_new,
                int *rowidx_new, double *val_new, int *levelItem,
                int *nlevel, int *levelPtr)
{
    int i, j, rc;
    int *rowptr = (int *)malloc(sizeof(int) * (size_t)(m + 1));
    int *perm = (int *)malloc(sizeof(int) * (size_t)(m > 0 ? m : 1));
    /* R/tranpose.h:92-108 (matrix_transposition_litelite): row counts */
    memset(rowptr, 0, sizeof(int) * (size_t)(m + 1));
    for (i = 0; i < nnz; i++)
        rowptr[rowidx[i]]++;
    orc_exclusive_scan(rowptr, m + 1);
    rc = orc_findlevel(colptr, rowidx, rowptr, m, nlevel, levelPtr, levelItem);
    if (rc != 0) {
        free(rowptr);
        free(perm);
        return rc;
    }
    /* inverse permutation (the stable sort of R/utils_cuda.h:740-742) */
    for (i = 0; i < m; i++)
        perm[levelItem[i]] = subst == 0 ? i : m - 1 - i;
    colptr_new[0] = 0;
    for (i = 0; i < n; i++) {
        int idx = subst == 0 ? levelItem[i] : levelItem[n - 1 - i];
        int cnt = colptr[idx + 1] - colptr[idx];
        colptr_new[i + 1] = colptr_new[i] + cnt;
        for (j = 0; j < cnt; j++) {
            rowidx_new[colptr_new[i] + j] = perm[rowidx[colptr[idx] + j]];
            if (val && val_new)
                val_new[colptr_new[i] + j] = val[colptr[idx] + j];
        }
    }
    if (subst != 0) {
        for (i = 0; i < m / 2; i++) {
            int t = levelItem[i];
            levelItem[i] = levelItem[m - 1 - i];
            levelItem[m - 1 - i] = t;
        }
    }
    free(rowptr);
    free(perm);
    return 0;
}

/* ------------------------------------------------------------------ A4 */
/* R/recblocking_preprocess.h:15-296 (host) / R/utils_cuda.h:241-482
 * (device): recursive bisection.  The reference's figure[] / judge[] loops
 * are restated as written (search for the power of two, subtracting the
 * largest power of two below temp; judge[i] = sum of the i+1 top powers).
 * For triangular blocks left/right hold the column range (== up/down). */
void orc_partition(int m, int lv, int subst, int *blk_m, int *blk_n,
                   int *loc_off, int *tmp_off, int *up, int *down, int *left,
                   int *right)
{
    int tri_block = 1 << lv;
    int sum_block = 2 * tri_block - 1;
    int step = m / tri_block;
    int i, j;
    int *figure = (int *)malloc(sizeof(int) * (size_t)(lv + 1));
    int *judge = (int *)calloc((size_t)(lv > 0 ? lv : 1), sizeof(int));
    int tri_up, tri_down, tri_count = 0;
    for (i = 0; i <= lv; i++)
        figure[i] = 1 << i;
    for (i = 0; i < lv; i++)
        for (j = lv - 1; j >= lv - 1 - i; j--)
            judge[i] += figure[j];
    for (i = 0; i < sum_block; i++) {
        loc_off[i] = 0;
        tmp_off[i] = 0;
    }
    if (subst == 0) {
        tri_up = 0;
        tri_down = step;
    } else {
        tri_up = m - step;
        tri_down = m;
    }
    for (i = 0; i < sum_block; i++) {
        if (i % 2 == 0) {
            blk_m[i] = tri_down - tri_up;
            blk_n[i] = tri_down - tri_up;
            up[i] = tri_up;
            down[i] = tri_down;
            left[i] = tri_up;
            right[i] = tri_down;
            loc_off[i] = tri_up;
            if (subst == 0) {
                tri_up += step;
                tri_down += step;
            } else {
                tri_up -= step;
                tri_down -= step;
            }
            tri_count++;
        } else {
            int rec_up = 0, rec_down = 0, rec_left, rec_right, temp = tri_count;
            int flag = 0;
            if (subst == 0)
                rec_up = tri_down - step;
            else
                rec_down = tri_up + step;
            while (1) {
                for (j = 0; j <= lv; j++) {
                    if (temp == figure[j]) {
                        if (subst == 0)
                            rec_down = rec_up + temp * step;
                        else
                            rec_up = rec_down - temp * step;
                        flag = 1;
                        break;
                    }
                }
                if (flag)
                    break;
                for (j = 0; j <= lv; j++) {
                    if (temp < figure[j]) {
                        temp -= figure[j - 1];
                        break;
                    }
                }
            }
            if (subst == 0) {
                rec_right = rec_up;
                rec_left = rec_right - (rec_down - rec_up);
            } else {
                rec_left = rec_down;
                rec_right = rec_left + (rec_down - rec_up);
            }
            for (j = 0; j < lv; j++) {
                if (judge[j] == tri_count) {
                    if (subst == 0)
                        rec_down = m;
                    else
                        rec_up = 0;
                    break;
                }
            }
            blk_m[i] = rec_down - rec_up;
            blk_n[i] = rec_right - rec_left;
            tmp_off[i] = rec_up;
            loc_off[i] = rec_left;
            if (i == sum_block - 2) {
                if (subst == 0)
                    tri_down = m;
                else
                    tri_up = 0;
            }
            up[i] = rec_up;
            down[i] = rec_down;
            left[i] = rec_left;
            right[i] = rec_right;
        }
    }
    free(figure);
    free(judge);
}

/* R/recblocking_preprocess.h:83-95,151-161 (and the cal_triblk_nnz /
 * cal_recblk_nnz kernels R/utils_cuda.h:191-239): nnz per block. */
void orc_block_nnz(const int *colptr, const int *rowidx, int lv, int subst,
                   const int *up, const int *down, const int *left,
                   const int *right, int *blk_nnz)
{
    int sum_block = 2 * (1 << lv) - 1;
    int i, j, k;
    for (i = 0; i < sum_block; i++) {
        int cnt = 0;
        for (j = left[i]; j < right[i]; j++) {
            for (k = colptr[j]; k < colptr[j + 1]; k++) {
                int r = rowidx[k];
                if (i % 2 == 0) {
                    if (subst == 0 ? r < down[i] : r >= up[i])
                        cnt++;
                } else if (r >= up[i] && r < down[i]) {
                    cnt++;
                }
            }
        }
        blk_nnz[i] = cnt;
    }
}

/* ------------------------------------------------------------- A5 / A6 */
/* Per-block extraction, local level analysis and the adaptive selector, as
 * the device path runs them (R/recblocking_solver_cuda.h:133-581 forward,
 * :671-1121 backward; host twin R/recblocking_preprocess.h:387-1472).
 *
 * Outputs per block i (arrays of size 2^(lv+1)-1):
 *   method[i]  triangular: 0 diagonal-only, 1 (nlv > 20000), 2 level-set,
 *              3 sync-free; square: 0 CSR-thread, 1 DCSR-thread,
 *              2 CSR-warp, 3 DCSR-warp, -1 when the block is empty.
 *   nlv[i]     triangular: local level count (1 for diagonal-only).
 *   nnzr[i]    triangular: blk_nnz / blk_m; square: (blk_nnz - longlen) /
 *              m_new, -1 when m_new == 0 (the reference divides by zero).
 *   m_new, lenmax, longrow, longlen, empty_ratio: square-block statistics
 *              (cal_longrow, R/utils_cuda.h:798-830).
 * Concatenated per triangular block, in triangular-block order:
 *   lvptr_local  (blk_m + 1 entries per block) local levelPtr
 *   lvitem_local (blk_m entries) local levelItem (BFS discovery order)
 *   nnz_lv       (blk_m + 1 entries; first nlv used) find_level_nnz
 *                (R/utils_cuda.h:781-796), which sums row counts over the
 *                *position* range [lvptr[li], lvptr[li+1]).
 * Returns 0, or -1 when a local level analysis stalls. */
int orc_analyze(const int *colptr, const int *rowidx, int m, int lv,
                int subst, const int *blk_m, const int *blk_n,
                const int *blk_nnz, const int *up, const int *down,
                const int *left, const int *right, int *method, int *nlv,
                int *nnzr, int *m_new, int *lenmax, int *longrow,
                int *longlen, double *empty_ratio, int *lvptr_local,
                int *lvitem_local, int *nnz_lv)
{
    int sum_block = 2 * (1 << lv) - 1;
    int i, j, k;
    int lvp_off = 0, item_off = 0;
    (void)m;
    for (i = 0; i < sum_block; i++) {
        int bm = blk_m[i], bn = blk_n[i], bz = blk_nnz[i];
        int *sptr = (int *)malloc(sizeof(int) * (size_t)(bn + 1));
        int *sidx = (int *)malloc(sizeof(int) * (size_t)(bz > 0 ? bz : 1));
        int *rptr = (int *)malloc(sizeof(int) * (size_t)(bm + 1));
        int *ridx = (int *)malloc(sizeof(int) * (size_t)(bz > 0 ? bz : 1));
        int q = 0;
        /* extraction: store_into_sub{tri,rec}mat_* R/utils_cuda.h:484-620 */
        sptr[0] = 0;
        for (j = left[i]; j < right[i]; j++) {
            for (k = colptr[j]; k < colptr[j + 1]; k++) {
                int r = rowidx[k];
                int in;
                if (i % 2 == 0)
                    in = subst == 0 ? r < down[i] : r >= up[i];
                else
                    in = r >= up[i] && r < down[i];
                if (in)
                    sidx[q++] = r - up[i];
            }
            sptr[j - left[i] + 1] = q;
        }
        /* transposition to CSR: R/utils_cuda.h:644-670 */
        orc_transpose(bn, bm, bz, sptr, sidx, NULL, ridx, rptr, NULL);
        m_new[i] = lenmax[i] = longrow[i] = longlen[i] = 0;
        empty_ratio[i] = 0.0;
        nlv[i] = 0;
        nnzr[i] = 0;
        if (i % 2 == 0) {
            int *lp = lvptr_local + lvp_off;
            int *li = lvitem_local + item_off;
            int *nl = nnz_lv + lvp_off;
            int fasttrack = bm == bz;
            if (fasttrack) {
                /* the reference skips the analysis; one level of bm rows.
                 * nnzr is not formed on this branch by the reference; it is
                 * reported as blk_nnz / blk_m (1), or 0 for an empty block */
                nlv[i] = 1;
                nnzr[i] = bm > 0 ? bz / bm : 0;
                lp[0] = 0;
                lp[1] = bm;
                for (j = 0; j < bm; j++)
                    li[j] = j;
                method[i] = 0;
            } else {
                int nl_ = 0;
                if (orc_findlevel(sptr, sidx, rptr, bm, &nl_, lp, li) != 0) {
                    free(sptr); free(sidx); free(rptr); free(ridx);
                    return -1;
                }
                nlv[i] = nl_;
                nnzr[i] = bz / bm;
                if (nl_ > 20000)
                    method[i] = 1;
                else if ((nnzr[i] <= 15 && nl_ <= 20) ||
                         (nnzr[i] == 1 && nl_ <= 100))
                    method[i] = 2;
                else
                    method[i] = 3;
            }
            for (j = 0; j < nlv[i]; j++) {
                int s = 0;
                for (k = lp[j]; k < lp[j + 1]; k++)
                    s += rptr[k + 1] - rptr[k];
                nl[j] = s;
            }
            lvp_off += bm + 1;
            item_off += bm;
        } else {
            /* cal_longrow R/utils_cuda.h:798-830 */
            int inew = 1;
            int lmax = bm > 0 ? rptr[1] - rptr[0] : 0;
            int lr = 0, ll = 0;
            for (j = 1; j <= bm; j++) {
                int len = rptr[j] - rptr[j - 1];
                lmax = len > lmax ? len : lmax;
                if (rptr[j] != rptr[j - 1]) {
                    if (len > ORC_LONGROW_THRESHOLD) {
                        lr++;
                        ll += len;
                    }
                    inew++;
                }
            }
            m_new[i] = inew - 1;
            lenmax[i] = lmax;
            longrow[i] = lr;
            longlen[i] = ll;
            empty_ratio[i] = bm > 0 ? 100.0 * (double)(bm - m_new[i]) / (double)bm : 0.0;
            if (bz == 0 || m_new[i] == 0) {
                nnzr[i] = -1;
                method[i] = -1;
            } else {
                /* R/recblocking_solver_cuda.h:464-546 */
                nnzr[i] = (bz - ll) / m_new[i];
                if (nnzr[i] <= 12 && empty_ratio[i] <= 50)
                    method[i] = 0;
                else if (nnzr[i] <= 12 && empty_ratio[i] > 50)
                    method[i] = 1;
                else if (nnzr[i] > 12 && empty_ratio[i] <= 15)
                    method[i] = 2;
                else
                    method[i] = 3;
            }
        }
        free(sptr);
        free(sidx);
        free(rptr);
        free(ridx);
    }
    return 0;
}

/* --------------------------------------------------------- A7 - A13 */
/* Block-ordered solve on the CPU, following L_calculate / U_calculate
 * (R/recblocking_calculate.h:15-325): b is permuted into level order
 * (R/utils_cuda.h:858-866), blocks run in sequence -- a triangular block is
 * solved by substitution restricted to its own rows (the arithmetic of
 * R/utils_sptrsv_cuda.h: x = (b - sum) / diag), a square block computes
 * y = A_sq * x[left:right] and subtracts it from b[up:down]
 * (R/utils_spmv_cuda.h, subKernel :205-212) -- and x is permuted back
 * (R/utils_cuda.h:868-876).  Inputs are the *reordered* CSC and the final
 * levelItem from orc_reorder.  Returns -1 on a zero or missing diagonal. */
int orc_solve_blocks(const int *colptr, const int *rowidx, const double *val,
                     int m, int lv, int subst, const int *up,
                     const int *down, const int *left, const int *right,
                     const int *levelItem, const double *b, double *x)
{
    int sum_block = 2 * (1 << lv) - 1;
    int i, j, k;
    double *bp = (double *)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    double *xp = (double *)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    double *y = (double *)calloc((size_t)(m > 0 ? m : 1), sizeof(double));
    for (i = 0; i < m; i++)
        bp[i] = b[levelItem[i]];
    for (i = 0; i < sum_block; i++) {
        if (i % 2 == 0) {
            if (subst == 0) {
                for (j = up[i]; j < down[i]; j++) {
                    int s = colptr[j], e = colptr[j + 1];
                    double d, xv;
                    if (e <= s || rowidx[s] != j || val[s] == 0.0)
                        goto fail;
                    d = val[s];
                    xv = bp[j] / d;
                    xp[j] = xv;
                    for (k = s + 1; k < e; k++)
                        if (rowidx[k] < down[i])
                            bp[rowidx[k]] -= val[k] * xv;
                }
            } else {
                for (j = down[i] - 1; j >= up[i]; j--) {
                    int s = colptr[j], e = colptr[j + 1];
                    double d, xv;
                    if (e <= s || rowidx[e - 1] != j || val[e - 1] == 0.0)
                        goto fail;
                    d = val[e - 1];
                    xv = bp[j] / d;
                    xp[j] = xv;
                    for (k = s; k < e - 1; k++)
                        if (rowidx[k] >= up[i])
                            bp[rowidx[k]] -= val[k] * xv;
                }
            }
        } else {
            for (j = up[i]; j < down[i]; j++)
                y[j] = 0.0;
            for (j = left[i]; j < right[i]; j++)
                for (k = colptr[j]; k < colptr[j + 1]; k++) {
                    int r = rowidx[k];
                    if (r >= up[i] && r < down[i])
                        y[r] += val[k] * xp[j];
                }
            for (j = up[i]; j < down[i]; j++)
                bp[j] -= y[j];
        }
    }
    for (i = 0; i < m; i++)
        x[levelItem[i]] = xp[i];
    free(bp);
    free(xp);
    free(y);
    return 0;
fail:
    free(bp);
    free(xp);
    free(y);
    return -1;
}

/* Textbook serial substitution on CSR (row-oriented pull), independent of
 * the block pipeline: forward for lower, backward for upper.  The diagonal
 * is found by column id, wherever it sits in the row.  Returns -1 on a zero
 * or missing diagonal, -2 on an entry outside the triangle. */
int orc_sptrsv_serial(const int *rowptr, const int *colidx, const double *val,
                      int m, int subst, const double *b, double *x)
{
    int t, j;
    for (t = 0; t < m; t++) {
        int i = subst == 0 ? t : m - 1 - t;
        double s = b[i], d = 0.0;
        int have = 0;
        for (j = rowptr[i]; j < rowptr[i + 1]; j++) {
            int c = colidx[j];
            if (c == i) {
                d = val[j];
                have = 1;
            } else if (subst == 0 ? c < i : c > i) {
                s -= val[j] * x[c];
            } else {
                return -2;
            }
        }
        if (!have || d == 0.0)
            return -1;
        x[i] = s / d;
    }
    return 0;
}

/* Sparse matrix-vector product y = A x on CSR (used for b = L x and the
 * residual b - L x by the tests). */
void orc_spmv(const int *rowptr, const int *colidx, const double *val, int m,
              const double *x, double *y)
{
    int i, j;
    for (i = 0; i < m; i++) {
        double s = 0.0;
        for (j = rowptr[i]; j < rowptr[i + 1]; j++)
            s += val[j] * x[colidx[j]];
        y[i] = s;
    }
}
