"""Matrix Market input and the benchmark driver's factor, through the
library's host entry points (include/recblock.h: rb_mmio_info, rb_mmio_data,
rb_mmio_triangle), which restate recblock-sptrsv/mmio_highlevel.h and
main.cu:150-228.  No device is touched.
"""
import ctypes

import numpy as np

from . import gen
from ._lib import RbError, check, lib

_MM_ERRORS = {-1: "cannot open", -2: "bad banner (not a coordinate matrix "
              "of a known type)", -4: "bad size line",
              -5: "missing, malformed or out-of-range entry"}


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def read_csr(path):
    """CSR of the file's matrix, built as the reference builds it (columns
    of a row in file order; symmetric / hermitian files mirrored).  Returns
    (m, n, rowptr, colidx, val, is_symmetric); raises ValueError on a
    reader error."""
    L = lib()
    m, n, nnz, sym = (ctypes.c_int() for _ in range(4))
    bp = path.encode()
    rc = L.rb_mmio_info(ctypes.byref(m), ctypes.byref(n), ctypes.byref(nnz),
                        ctypes.byref(sym), bp)
    if rc != 0:
        raise ValueError("%s: %s (%d)" % (path, _MM_ERRORS.get(rc, "error"), rc))
    rowptr = np.zeros(m.value + 1, np.int32)
    colidx = np.zeros(max(1, nnz.value), np.int32)
    val = np.zeros(max(1, nnz.value), np.float64)
    rc = L.rb_mmio_data(_p(rowptr), _p(colidx), _p(val), bp)
    if rc != 0:
        raise ValueError("%s: %s (%d)" % (path, _MM_ERRORS.get(rc, "error"), rc))
    return (m.value, n.value, rowptr, colidx[:nnz.value], val[:nnz.value],
            bool(sym.value))


def triangle(m, rowptr, colidx, val, substitution, reference_values=False,
             seed=0, nrhs=1):
    """The driver's triangular factor (CSR), x_ref (m * nrhs, column-major)
    and b = T x_ref.  Returns (rowptr, colidx, val, x_ref, b)."""
    L = lib()
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    colidx = np.ascontiguousarray(colidx, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    tp = np.zeros(m + 1, np.int32)
    nz = ctypes.c_int()
    check(L.rb_mmio_triangle(m, _p(rowptr), _p(colidx), _p(val), substitution,
                             int(reference_values), seed, 0, _p(tp), None, None,
                             ctypes.byref(nz), None, None))
    tc = np.zeros(max(1, nz.value), np.int32)
    tv = np.zeros(max(1, nz.value), np.float64)
    xref = np.zeros(max(1, m * nrhs))
    b = np.zeros(max(1, m * nrhs))
    check(L.rb_mmio_triangle(m, _p(rowptr), _p(colidx), _p(val), substitution,
                             int(reference_values), seed, nrhs, _p(tp), _p(tc),
                             _p(tv), ctypes.byref(nz), _p(xref), _p(b)))
    return tp, tc[:nz.value], tv[:nz.value], xref[:m * nrhs], b[:m * nrhs]


def load_factor(path, substitution=0, reference_values=False, seed=0, nrhs=1):
    """A file's triangular factor as a ``gen.Factor`` (CSC, ready for
    ``Plan.from_factor``) plus x_ref and b, as the driver builds them."""
    m, n, rp, ci, v, _ = read_csr(path)
    if m != n:
        raise RbError(2, "we need square matrix")
    tp, tc, tv, xref, b = triangle(m, rp, ci, v, substitution,
                                   reference_values, seed, nrhs)
    cp, ri, cv = gen.csc_transpose(m, tp, tc, tv)
    f = gen.Factor(m, cp, ri, cv, substitution == 0, path)
    return f, xref.reshape(nrhs, m), b.reshape(nrhs, m)


def write_mtx(path, rows, cols, vals, m, n, symmetry="general", field="real"):
    """Write a coordinate Matrix Market file (1-based), for tests and
    examples; entries in the given order."""
    with open(path, "w") as fh:
        fh.write("%%%%MatrixMarket matrix coordinate %s %s\n" % (field, symmetry))
        fh.write("%% written by rbsptrsv.mmio\n")
        fh.write("%d %d %d\n" % (m, n, len(rows)))
        for i, j, x in zip(rows, cols, vals):
            if field == "pattern":
                fh.write("%d %d\n" % (i + 1, j + 1))
            elif field == "integer":
                fh.write("%d %d %d\n" % (i + 1, j + 1, int(x)))
            elif field == "complex":
                fh.write("%d %d %.17g 0.5\n" % (i + 1, j + 1, x))
            else:
                fh.write("%d %d %.17g\n" % (i + 1, j + 1, x))
