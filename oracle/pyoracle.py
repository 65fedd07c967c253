"""ctypes front end of the C oracle (``rb_oracle.c``).

Test infrastructure only (the checker).  Each wrapper names the C function
and, through it, the reference file:line it restates.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_I = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build(force=False):
    """Compile liboracle.so with the committed Makefile (gcc only)."""
    src = os.path.join(_HERE, "rb_oracle.c")
    if force or not os.path.exists(_LIB_PATH) or \
            os.path.getmtime(_LIB_PATH) < os.path.getmtime(src):
        subprocess.check_call(["make", "-s", "-B", "-C", _HERE, "liboracle.so"])
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIB_PATH)
        c_int, c_dp = ctypes.c_int, ctypes.POINTER(ctypes.c_int)
        L.orc_lv_auto.argtypes = [c_int, c_int]
        L.orc_lv_auto.restype = c_int
        L.orc_transpose.argtypes = [c_int, c_int, c_int, _I, _I, ctypes.c_void_p,
                                    _I, _I, ctypes.c_void_p]
        L.orc_transpose.restype = None
        L.orc_findlevel.argtypes = [_I, _I, _I, c_int, c_dp, _I, _I]
        L.orc_findlevel.restype = c_int
        L.orc_reorder.argtypes = [_I, _I, _D, c_int, c_int, c_int, c_int,
                                  _I, _I, _D, _I, c_dp, _I]
        L.orc_reorder.restype = c_int
        L.orc_partition.argtypes = [c_int, c_int, c_int] + [_I] * 8
        L.orc_partition.restype = None
        L.orc_block_nnz.argtypes = [_I, _I, c_int, c_int, _I, _I, _I, _I, _I]
        L.orc_block_nnz.restype = None
        L.orc_analyze.argtypes = ([_I, _I, c_int, c_int, c_int] + [_I] * 7 +
                                  [_I] * 7 + [_D] + [_I] * 3)
        L.orc_analyze.restype = c_int
        L.orc_solve_blocks.argtypes = [_I, _I, _D, c_int, c_int, c_int,
                                       _I, _I, _I, _I, _I, _D, _D]
        L.orc_solve_blocks.restype = c_int
        L.orc_sptrsv_serial.argtypes = [_I, _I, _D, c_int, c_int, _D, _D]
        L.orc_sptrsv_serial.restype = c_int
        L.orc_spmv.argtypes = [_I, _I, _D, c_int, _D, _D]
        L.orc_spmv.restype = None
        _lib = L
    return _lib


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def lv_auto(m, device_id=0):
    """R/main.cu:245-254."""
    return lib().orc_lv_auto(int(m), int(device_id))


def transpose(nrows, ncols, ptr, idx, val=None):
    """R/tranpose.h:7-54: returns (out_ptr[ncols+1], out_idx, out_val)."""
    ptr, idx = _i32(ptr), _i32(idx)
    nnz = int(ptr[nrows])
    out_ptr = np.zeros(ncols + 1, np.int32)
    out_idx = np.zeros(max(nnz, 1), np.int32)
    if val is not None:
        val = _f64(val)
        out_val = np.zeros(max(nnz, 1), np.float64)
        lib().orc_transpose(nrows, ncols, nnz, ptr, idx, val.ctypes.data,
                            out_idx, out_ptr, out_val.ctypes.data)
        return out_ptr, out_idx[:nnz], out_val[:nnz]
    lib().orc_transpose(nrows, ncols, nnz, ptr, idx, None, out_idx, out_ptr, None)
    return out_ptr, out_idx[:nnz], None


def findlevel(colptr, rowidx, rowptr, m):
    """R/findlevel.h:7-83 -> (rc, nlevel, levelPtr[:nlevel+1], levelItem)."""
    nlev = ctypes.c_int(0)
    lp = np.zeros(m + 2, np.int32)
    li = np.zeros(max(m, 1), np.int32)
    rc = lib().orc_findlevel(_i32(colptr), _i32(rowidx), _i32(rowptr), m,
                             ctypes.byref(nlev), lp, li)
    return rc, nlev.value, lp[:nlev.value + 1].copy(), li[:m].copy()


def reorder(colptr, rowidx, val, m, subst):
    """R/utils_reordering.h:102-210.  Returns a dict with the permuted CSC,
    the final levelItem and the global levels."""
    colptr, rowidx, val = _i32(colptr), _i32(rowidx), _f64(val)
    nnz = int(colptr[m])
    cp = np.zeros(m + 1, np.int32)
    ri = np.zeros(max(nnz, 1), np.int32)
    vv = np.zeros(max(nnz, 1), np.float64)
    li = np.zeros(max(m, 1), np.int32)
    lp = np.zeros(m + 2, np.int32)
    nlev = ctypes.c_int(0)
    rc = lib().orc_reorder(colptr, rowidx, val, m, m, nnz, subst, cp, ri, vv,
                           li, ctypes.byref(nlev), lp)
    if rc != 0:
        raise ValueError("level analysis stalled (missing diagonal or cycle)")
    return dict(colptr=cp, rowidx=ri[:nnz], val=vv[:nnz], levelItem=li[:m],
                nlevel=nlev.value, levelPtr=lp[:nlev.value + 1])


def partition(m, lv, subst):
    """R/recblocking_preprocess.h:15-296 (bounds only)."""
    nb = 2 * (1 << lv) - 1
    arrs = [np.zeros(nb, np.int32) for _ in range(8)]
    lib().orc_partition(int(m), int(lv), int(subst), *arrs)
    names = ["blk_m", "blk_n", "loc_off", "tmp_off", "up", "down", "left", "right"]
    return dict(zip(names, arrs))


def block_nnz(colptr, rowidx, lv, subst, part):
    nb = 2 * (1 << lv) - 1
    out = np.zeros(nb, np.int32)
    lib().orc_block_nnz(_i32(colptr), _i32(rowidx), lv, subst, part["up"],
                        part["down"], part["left"], part["right"], out)
    return out


def analyze(colptr, rowidx, m, lv, subst, part, blk_nnz):
    """Per-block selector inputs and decisions
    (R/recblocking_solver_cuda.h:133-581 and :671-1121)."""
    nb = 2 * (1 << lv) - 1
    ints = {k: np.zeros(nb, np.int32) for k in
            ["method", "nlv", "nnzr", "m_new", "lenmax", "longrow", "longlen"]}
    er = np.zeros(nb, np.float64)
    ntri = 1 << lv
    lvptr = np.zeros(m + ntri + 1, np.int32)
    lvitem = np.zeros(max(m, 1), np.int32)
    nnzlv = np.zeros(m + ntri + 1, np.int32)
    rc = lib().orc_analyze(_i32(colptr), _i32(rowidx), m, lv, subst,
                           part["blk_m"], part["blk_n"], _i32(blk_nnz),
                           part["up"], part["down"], part["left"], part["right"],
                           ints["method"], ints["nlv"], ints["nnzr"], ints["m_new"],
                           ints["lenmax"], ints["longrow"], ints["longlen"], er,
                           lvptr, lvitem, nnzlv)
    if rc != 0:
        raise ValueError("local level analysis stalled")
    out = dict(ints)
    out["empty_ratio"] = er
    # split the concatenated per-triangular-block arrays
    lp_list, li_list, nl_list = [], [], []
    po = io = 0
    for t in range(ntri):
        bm = int(part["blk_m"][2 * t])
        nlv = int(ints["nlv"][2 * t])
        lp_list.append(lvptr[po:po + nlv + 1].copy())
        li_list.append(lvitem[io:io + bm].copy())
        nl_list.append(nnzlv[po:po + nlv].copy())
        po += bm + 1
        io += bm
    out["lvptr_local"] = lp_list
    out["lvitem_local"] = li_list
    out["nnz_lv"] = nl_list
    return out


def solve_blocks(re, m, lv, subst, part, b):
    """Block-ordered CPU solve (R/recblocking_calculate.h:15-325)."""
    x = np.zeros(m, np.float64)
    rc = lib().orc_solve_blocks(re["colptr"], _i32(re["rowidx"]), _f64(re["val"]), m,
                                lv, subst, part["up"], part["down"], part["left"],
                                part["right"], _i32(re["levelItem"]), _f64(b), x)
    if rc != 0:
        raise ValueError("zero or missing diagonal")
    return x


def sptrsv_serial(rowptr, colidx, val, m, subst, b):
    """Textbook row-oriented substitution on CSR."""
    x = np.zeros(m, np.float64)
    rc = lib().orc_sptrsv_serial(_i32(rowptr), _i32(colidx), _f64(val), m, subst,
                                 _f64(b), x)
    if rc != 0:
        raise ValueError("serial substitution failed (rc=%d)" % rc)
    return x


def spmv(rowptr, colidx, val, m, x):
    y = np.zeros(m, np.float64)
    lib().orc_spmv(_i32(rowptr), _i32(colidx), _f64(val), m, _f64(x), y)
    return y


def csc_to_csr(colptr, rowidx, val, m):
    """Transpose a square CSC to CSR with the oracle's transposition."""
    return transpose(m, m, colptr, rowidx, val)


def full_pipeline(colptr, rowidx, val, m, subst, lv):
    """Reorder + partition + block nnz + selector, as one dict."""
    re = reorder(colptr, rowidx, val, m, subst)
    part = partition(m, lv, subst)
    bnz = block_nnz(re["colptr"], re["rowidx"], lv, subst, part)
    an = analyze(re["colptr"], re["rowidx"], m, lv, subst, part, bnz)
    return re, part, bnz, an
