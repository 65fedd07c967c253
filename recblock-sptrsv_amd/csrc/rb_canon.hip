// Canonical CSC for general inputs (rb_opts.canonicalize): row ids in any
// order inside a column and repeated (row, column) entries are accepted;
// the plan is built from the matrix with every column's row ids strictly
// ascending and repeated entries summed.  The reference accepts neither
// (its level analysis and block extraction assume sorted, unique row ids,
// R/findlevel.h:7-83, R/utils_cuda.h:484-620); this has no counterpart
// there.
//
// Entries are ordered by (column, row, input position) with two stable
// radix sorts -- by row, then by column -- and each run of equal (column,
// row) is summed left to right in input order by the run's first entry, so
// the result does not depend on thread timing.  colptr must be a valid
// pointer array and row ids must lie in [0, n): those are checked first
// (RB_ERR_BAD_STRUCTURE), since the sorts rely on them.
#include "rb_internal.h"

namespace rb {

namespace {

constexpr int kT = 256;
inline int nblk(long n) { return (int)((n + kT - 1) / kT); }

__global__ void k_canon_check(const int *__restrict__ colptr,
                              const int *__restrict__ rowidx, int n, int nnz,
                              int *err) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int s = colptr[j], e = colptr[j + 1];
    if (s < 0 || e < s || e > nnz) {
        atomicOr(err, 1);
        return;
    }
    for (int q = s; q < e; q++)
        if (rowidx[q] < 0 || rowidx[q] >= n) {
            atomicOr(err, 1);
            return;
        }
}

// column of every entry, and the identity permutation
__global__ void k_canon_colof(const int *__restrict__ colptr, int n, int *ecol,
                              int *iota) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    for (int q = colptr[j]; q < colptr[j + 1]; q++) {
        ecol[q] = j;
        iota[q] = q;
    }
}

__global__ void k_canon_gather(const int *__restrict__ src,
                               const int *__restrict__ idx, int n, int *dst) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[idx[k]];
}

// head[k] = 1 where a new (column, row) starts in sorted order
__global__ void k_canon_heads(const int *__restrict__ col,
                              const int *__restrict__ row, int n, int *head) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    head[k] = k == 0 || col[k] != col[k - 1] || row[k] != row[k - 1];
}

// each run's first entry sums the run in input order
__global__ void k_canon_merge(const int *__restrict__ col,
                              const int *__restrict__ row,
                              const int *__restrict__ idx,
                              const double *__restrict__ val,
                              const int *__restrict__ head,
                              const int *__restrict__ pos, int n, int *out_row,
                              double *out_val, int *colcnt) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n || !head[k]) return;
    double s = val[idx[k]];
    for (int t = k + 1; t < n && !head[t]; t++) s += val[idx[t]];
    out_row[pos[k]] = row[k];
    out_val[pos[k]] = s;
    atomicAdd(&colcnt[col[k]], 1);
}

int bits_for(int n) {
    int b = 1;
    while (b < 31 && (1 << b) < n) b++;
    return b;
}

}  // namespace

rb_status canonicalize_csc(const int *d_colptr, const int *d_rowidx,
                           const double *d_val, int n, int nnz, Arena &out,
                           int **c_colptr, int **c_rowidx, double **c_val,
                           int *c_nnz, hipStream_t s) {
    Arena tmp;
    struct Guard {
        Arena &a;
        ~Guard() { a.release(); }
    } guard{tmp};
    int hp[2] = {0, 0};
    RB_HIP(hipMemcpyAsync(&hp[0], d_colptr, sizeof(int), hipMemcpyDeviceToHost, s));
    RB_HIP(hipMemcpyAsync(&hp[1], d_colptr + n, sizeof(int), hipMemcpyDeviceToHost, s));
    RB_HIP(hipStreamSynchronize(s));
    if (hp[0] != 0 || hp[1] != nnz) {
        set_error(RB_ERR_BAD_STRUCTURE,
                  "colptr[0] must be 0 and colptr[n] must equal nnz");
        return RB_ERR_BAD_STRUCTURE;
    }
    int *err;
    RB_TRY(tmp.alloc(&err, 1));
    RB_HIP(hipMemsetAsync(err, 0, sizeof(int), s));
    if (n > 0) k_canon_check<<<nblk(n), kT, 0, s>>>(d_colptr, d_rowidx, n, nnz, err);
    RB_HIP(hipGetLastError());
    int herr = 0;
    RB_HIP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s));
    RB_HIP(hipStreamSynchronize(s));
    if (herr) {
        set_error(RB_ERR_BAD_STRUCTURE,
                  "CSC structure invalid: colptr not monotone or a row id out "
                  "of range");
        return RB_ERR_BAD_STRUCTURE;
    }
    int *colcnt;
    RB_TRY(out.alloc(c_colptr, (size_t)n + 1));
    RB_TRY(tmp.alloc(&colcnt, (size_t)n + 1));
    RB_HIP(hipMemsetAsync(colcnt, 0, sizeof(int) * ((size_t)n + 1), s));
    if (nnz == 0) {
        RB_TRY(out.alloc(c_rowidx, 1));
        RB_TRY(out.alloc(c_val, 1));
        RB_HIP(hipMemsetAsync(*c_colptr, 0, sizeof(int) * ((size_t)n + 1), s));
        *c_nnz = 0;
        RB_HIP(hipStreamSynchronize(s));
        return RB_OK;
    }
    int *ecol, *iota, *row1, *idx1, *key2, *col2, *idx2, *row2, *head, *pos;
    for (int **a : {&ecol, &iota, &row1, &idx1, &key2, &col2, &idx2, &row2,
                    &head, &pos})
        RB_TRY(tmp.alloc(a, (size_t)nnz));
    k_canon_colof<<<nblk(n), kT, 0, s>>>(d_colptr, n, ecol, iota);
    RB_HIP(hipGetLastError());
    const int bits = bits_for(n);
    // stable by row, then stable by column: (column, row, input position)
    RB_TRY(sort_pairs_i32(d_rowidx, row1, iota, idx1, nnz, bits, s));
    k_canon_gather<<<nblk(nnz), kT, 0, s>>>(ecol, idx1, nnz, key2);
    RB_HIP(hipGetLastError());
    RB_TRY(sort_pairs_i32(key2, col2, idx1, idx2, nnz, bits, s));
    k_canon_gather<<<nblk(nnz), kT, 0, s>>>(d_rowidx, idx2, nnz, row2);
    k_canon_heads<<<nblk(nnz), kT, 0, s>>>(col2, row2, nnz, head);
    RB_HIP(hipGetLastError());
    RB_TRY(exclusive_scan_i32(head, pos, nnz, s));
    int last[2];
    RB_HIP(hipMemcpyAsync(&last[0], pos + nnz - 1, sizeof(int), hipMemcpyDeviceToHost, s));
    RB_HIP(hipMemcpyAsync(&last[1], head + nnz - 1, sizeof(int), hipMemcpyDeviceToHost, s));
    RB_HIP(hipStreamSynchronize(s));
    const int nu = last[0] + last[1];
    RB_TRY(out.alloc(c_rowidx, (size_t)nu));
    RB_TRY(out.alloc(c_val, (size_t)nu));
    k_canon_merge<<<nblk(nnz), kT, 0, s>>>(col2, row2, idx2, d_val, head, pos, nnz,
                                           *c_rowidx, *c_val, colcnt);
    RB_HIP(hipGetLastError());
    RB_TRY(exclusive_scan_i32(colcnt, *c_colptr, n + 1, s));
    *c_nnz = nu;
    return RB_OK;
}

}  // namespace rb
