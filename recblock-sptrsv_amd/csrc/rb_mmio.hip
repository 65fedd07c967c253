// Matrix Market input and the benchmark driver's triangular factor: the
// host side of the reference's `sptrsv-double` program (main.cu +
// mmio_highlevel.h), as C entry points of the library.
//
// Reader (rb_mmio_info / rb_mmio_data), mmio_highlevel.h:7-296: coordinate
// files of real / integer / pattern / complex (real part kept) values,
// 1-based indices.  Symmetric and hermitian files are mirrored: an
// off-diagonal entry (i, j) also stands for (j, i) with the same value.  CSR
// is built by counting entries per row, an exclusive scan, and one pass over
// the entries in file order, each writing at its row's next position (a
// mirrored pair writes (i, j) and then (j, i)): columns inside a row keep
// file order, not sorted order -- the reference's factor depends on it.
// Two differences, both where the reference reads past its input: an index
// outside 1..m / 1..n or a truncated entry list returns -5 instead of
// writing out of bounds, and `array` (dense) files return -2 instead of
// being misparsed.
//
// Triangle (rb_mmio_triangle), main.cu:150-228: forward keeps the entries
// with column < row in row order and appends the diagonal last; backward
// puts the diagonal first and keeps the entries with column > row.  With
// `reference_values` every kept entry becomes rand() % 10 + 1 and the
// diagonal 1.0, drawn from the C library's rand() after srand(seed), in the
// reference's order; x_ref[r * n + i] = rand() % 10 + 1 follows (in both
// modes, after srand(seed)), and
// b[r * m + i] = sum_j T[i, j] * x_ref[r * n + j] over the row's entries
// in order.  Without it the file's values are kept and the diagonal is the
// file's own (RB_ERR_MISSING_DIAG if a row has none; several diagonal
// entries in a row are summed, as duplicate coordinates add up).
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "rb_internal.h"

namespace {

struct MMHeader {
    bool real = false, integer = false, pattern = false, complex = false;
    bool symmetric = false;
    long m = 0, n = 0, nnz = 0;
};

// whole file into memory (a factor's file is read twice, as by the
// reference's two-call interface)
bool slurp(const char *path, std::string &buf) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    char tmp[1 << 16];
    size_t k;
    while ((k = std::fread(tmp, 1, sizeof tmp, f)) > 0) buf.append(tmp, k);
    std::fclose(f);
    return true;
}

bool ieq(const char *a, const char *b) {
    for (; *a && *b; a++, b++)
        if ((*a | 0x20) != (*b | 0x20)) return false;
    return *a == *b;
}

// Banner and size line.  Returns 0, -2 (banner) or -4 (size line); `pos`
// is left at the first entry.
int parse_header(const std::string &s, MMHeader &h, size_t &pos) {
    size_t e = s.find('\n');
    std::string line = s.substr(0, e == std::string::npos ? s.size() : e);
    char t0[64] = {0}, t1[64] = {0}, t2[64] = {0}, t3[64] = {0}, t4[64] = {0};
    if (std::sscanf(line.c_str(), "%63s %63s %63s %63s %63s", t0, t1, t2, t3, t4) != 5)
        return -2;
    if (std::strcmp(t0, "%%MatrixMarket") != 0 || !ieq(t1, "matrix") ||
        !ieq(t2, "coordinate"))
        return -2;
    h.real = ieq(t3, "real") || ieq(t3, "double");
    h.integer = ieq(t3, "integer");
    h.pattern = ieq(t3, "pattern");
    h.complex = ieq(t3, "complex");
    if (!(h.real || h.integer || h.pattern || h.complex)) return -2;
    if (ieq(t4, "symmetric") || ieq(t4, "hermitian")) h.symmetric = true;
    else if (!ieq(t4, "general") && !ieq(t4, "skew-symmetric")) return -2;
    pos = e == std::string::npos ? s.size() : e + 1;
    // comment and blank lines, then "M N nz"
    while (pos < s.size()) {
        size_t le = s.find('\n', pos);
        if (le == std::string::npos) le = s.size();
        const char *p = s.c_str() + pos;
        size_t q = 0;
        while (pos + q < le && (p[q] == ' ' || p[q] == '\t' || p[q] == '\r')) q++;
        if (pos + q == le || p[q] == '%') {
            pos = le + 1;
            continue;
        }
        std::string sz = s.substr(pos, le - pos);
        if (std::sscanf(sz.c_str(), "%ld %ld %ld", &h.m, &h.n, &h.nnz) != 3 ||
            h.m < 0 || h.n < 0 || h.nnz < 0 || h.m > 0x7fffffffL ||
            h.n > 0x7fffffffL || h.nnz > 0x7fffffffL)
            return -4;
        pos = le + 1;
        return 0;
    }
    return -4;
}

// The entries (0-based) in file order.  Returns 0 or -5.
int parse_entries(const std::string &s, size_t pos, const MMHeader &h,
                  std::vector<int> &ri, std::vector<int> &ci,
                  std::vector<double> &v) {
    ri.resize(h.nnz);
    ci.resize(h.nnz);
    v.resize(h.nnz);
    const char *p = s.c_str() + pos;
    const char *end = s.c_str() + s.size();
    for (long k = 0; k < h.nnz; k++) {
        // skip blank / comment lines between entries
        while (p < end) {
            while (p < end && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n')) p++;
            if (p < end && *p == '%') {
                while (p < end && *p != '\n') p++;
                continue;
            }
            break;
        }
        if (p >= end) return -5;
        char *q;
        errno = 0;
        const long i = std::strtol(p, &q, 10);
        if (q == p) return -5;
        p = q;
        const long j = std::strtol(p, &q, 10);
        if (q == p) return -5;
        p = q;
        double val = 1.0;
        if (h.real || h.complex) {
            val = std::strtod(p, &q);
            if (q == p) return -5;
            p = q;
            if (h.complex) {  // imaginary part, dropped
                std::strtod(p, &q);
                if (q == p) return -5;
                p = q;
            }
        } else if (h.integer) {
            const long iv = std::strtol(p, &q, 10);
            if (q == p) return -5;
            p = q;
            val = (double)iv;
        }
        if (i < 1 || i > h.m || j < 1 || j > h.n) return -5;
        if (h.symmetric && j > h.m) return -5;  // its mirror has no row
        ri[k] = (int)(i - 1);
        ci[k] = (int)(j - 1);
        v[k] = val;
        while (p < end && *p != '\n') p++;  // rest of the line
    }
    return 0;
}

int read_all(const char *path, MMHeader &h, std::vector<int> &ri,
             std::vector<int> &ci, std::vector<double> &v) {
    std::string s;
    if (!path || !slurp(path, s)) return -1;
    size_t pos = 0;
    int rc = parse_header(s, h, pos);
    if (rc) return rc;
    return parse_entries(s, pos, h, ri, ci, v);
}

// row counts (mirrored pairs counted twice) -> exclusive scan
void row_pointer(const MMHeader &h, const std::vector<int> &ri,
                 const std::vector<int> &ci, std::vector<long> &ptr) {
    ptr.assign(h.m + 1, 0);
    for (long k = 0; k < h.nnz; k++) {
        ptr[ri[k] + 1]++;
        if (h.symmetric && ri[k] != ci[k]) ptr[ci[k] + 1]++;
    }
    for (long i = 0; i < h.m; i++) ptr[i + 1] += ptr[i];
}

}  // namespace

extern "C" int rb_mmio_info(int *m, int *n, int *nnz, int *is_symmetric,
                            const char *path) {
    MMHeader h;
    std::vector<int> ri, ci;
    std::vector<double> v;
    const int rc = read_all(path, h, ri, ci, v);
    if (rc) return rc;
    std::vector<long> ptr;
    row_pointer(h, ri, ci, ptr);
    if (ptr[h.m] > 0x7fffffffL) return -5;
    if (m) *m = (int)h.m;
    if (n) *n = (int)h.n;
    if (nnz) *nnz = (int)ptr[h.m];
    if (is_symmetric) *is_symmetric = h.symmetric ? 1 : 0;
    return 0;
}

extern "C" int rb_mmio_data(int *csr_rowptr, int *csr_colidx, double *csr_val,
                            const char *path) {
    if (!csr_rowptr || !csr_colidx || !csr_val) return -1;
    MMHeader h;
    std::vector<int> ri, ci;
    std::vector<double> v;
    const int rc = read_all(path, h, ri, ci, v);
    if (rc) return rc;
    std::vector<long> ptr;
    row_pointer(h, ri, ci, ptr);
    if (ptr[h.m] > 0x7fffffffL) return -5;
    std::vector<long> next(ptr.begin(), ptr.end() - 1);
    for (long k = 0; k < h.nnz; k++) {
        long o = next[ri[k]]++;
        csr_colidx[o] = ci[k];
        csr_val[o] = v[k];
        if (h.symmetric && ri[k] != ci[k]) {
            o = next[ci[k]]++;
            csr_colidx[o] = ri[k];
            csr_val[o] = v[k];
        }
    }
    for (long i = 0; i <= h.m; i++) csr_rowptr[i] = (int)ptr[i];
    return 0;
}

extern "C" rb_status rb_mmio_triangle(int m, const int *rowptr,
                                      const int *colidx, const double *val,
                                      int substitution, int reference_values,
                                      unsigned seed, int nrhs, int *tri_rowptr,
                                      int *tri_colidx, double *tri_val,
                                      int *tri_nnz, double *x_ref, double *b) {
    using namespace rb;
    if (m < 0 || !rowptr || !tri_rowptr || !tri_nnz ||
        (substitution != 0 && substitution != 1) || nrhs < 0 ||
        (nrhs > 0 && (!x_ref || !b))) {
        set_error(RB_ERR_INVALID_ARG, "rb_mmio_triangle: bad argument");
        return RB_ERR_INVALID_ARG;
    }
    const bool fwd = substitution == 0;
    std::srand(seed);
    long nz = 0;
    tri_rowptr[0] = 0;
    for (int i = 0; i < m; i++) {
        double dsum = 0.0;
        bool hasd = false;
        if (!reference_values)
            for (int k = rowptr[i]; k < rowptr[i + 1]; k++)
                if (colidx[k] == i) {
                    dsum += val[k];
                    hasd = true;
                }
        if (!reference_values && !hasd) {
            set_error(RB_ERR_MISSING_DIAG,
                      "row " + std::to_string(i) + " has no diagonal entry");
            return RB_ERR_MISSING_DIAG;
        }
        const double dv = reference_values ? 1.0 : dsum;
        if (!fwd) {
            if (tri_colidx) {
                tri_colidx[nz] = i;
                tri_val[nz] = dv;
            }
            nz++;
        }
        for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
            const int j = colidx[k];
            if (fwd ? j < i : j > i) {
                if (tri_colidx) {
                    tri_colidx[nz] = j;
                    tri_val[nz] = reference_values ? (double)(std::rand() % 10 + 1)
                                                   : val[k];
                }
                nz++;
            }
        }
        if (fwd) {
            if (tri_colidx) {
                tri_colidx[nz] = i;
                tri_val[nz] = dv;
            }
            nz++;
        }
        if (nz > 0x7fffffffL) {
            set_error(RB_ERR_INVALID_ARG, "triangle has more than 2^31-1 entries");
            return RB_ERR_INVALID_ARG;
        }
        tri_rowptr[i + 1] = (int)nz;
    }
    *tri_nnz = (int)nz;
    if (!tri_colidx || nrhs == 0) return RB_OK;
    for (long i = 0; i < (long)m * nrhs; i++) x_ref[i] = std::rand() % 10 + 1;
    for (int r = 0; r < nrhs; r++)
        for (int i = 0; i < m; i++) {
            double s = 0.0;
            for (int k = tri_rowptr[i]; k < tri_rowptr[i + 1]; k++)
                s += tri_val[k] * x_ref[(long)r * m + tri_colidx[k]];
            b[(long)r * m + i] = s;
        }
    return RB_OK;
}
