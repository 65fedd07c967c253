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


_CPU_PATH = os.path.join(_HERE, "libcpubase.so")
_cpu = None


def build(force=False):
    """Compile liboracle.so and libcpubase.so with the committed Makefile
    (gcc only)."""
    for lib_path, src in ((_LIB_PATH, "rb_oracle.c"), (_CPU_PATH, "rb_cpu.c")):
        src = os.path.join(_HERE, src)
        if force or not os.path.exists(lib_path) or \
                os.path.getmtime(lib_path) < os.path.getmtime(src):
            subprocess.check_call(["make", "-s", "-B", "-C", _HERE,
                                   os.path.basename(lib_path)])
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


def cpu_lib():
    """The CPU baseline (rb_cpu.c, OpenMP)."""
    global _cpu
    if _cpu is None:
        build()
        L = ctypes.CDLL(_CPU_PATH)
        c_int, vp = ctypes.c_int, ctypes.c_void_p
        L.cpu_plan_create.argtypes = [_I, _I, _D, c_int, c_int, c_int] + \
            [_I] * 5 + [_I, _I, _I]
        L.cpu_plan_create.restype = vp
        L.cpu_plan_solve.argtypes = [vp, _D, _D, c_int]
        L.cpu_plan_solve.restype = None
        L.cpu_plan_destroy.argtypes = [vp]
        L.cpu_plan_destroy.restype = None
        L.cpu_max_threads.argtypes = []
        L.cpu_max_threads.restype = c_int
        _cpu = L
    return _cpu


class CpuSolver:
    """The reference's block-ordered solve on the host cores (rb_cpu.c):
    built from the oracle's level-ordered matrix, partition and local
    levels.  ``solve(b, threads)`` returns x in the original order."""

    def __init__(self, f, lv):
        s = f.substitution
        re, part, bnz, an = full_pipeline(f.colptr, f.rowidx, f.val, f.n,
                                          s, lv)
        self.m = f.n
        lvp = np.concatenate([np.concatenate([lp, np.zeros(
            int(part["blk_m"][2 * t]) + 1 - len(lp), np.int32)])
            for t, lp in enumerate(an["lvptr_local"])]).astype(np.int32)
        lvi = np.concatenate(an["lvitem_local"]).astype(np.int32) \
            if self.m else np.zeros(1, np.int32)
        self._h = cpu_lib().cpu_plan_create(
            _i32(re["colptr"]), _i32(re["rowidx"]), _f64(re["val"]), f.n, lv,
            s, part["up"], part["down"], part["left"], part["right"],
            _i32(an["nlv"]), lvp, lvi, _i32(re["levelItem"]))
        if not self._h:
            raise ValueError("zero or missing diagonal")

    def solve(self, b, threads=0):
        x = np.zeros(self.m, np.float64)
        cpu_lib().cpu_plan_solve(self._h, _f64(b), x, int(threads))
        return x

    def close(self):
        if self._h:
            cpu_lib().cpu_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def cpu_threads():
    return cpu_lib().cpu_max_threads()


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


# ---------------------------------------------------------------- Matrix
# Market input and the driver's factor (pure Python: small files only)

def mm_read_csr(path):
    """The reference's mmio_info + mmio_data
    (recblock-sptrsv/mmio_highlevel.h:7-137, 139-296): banner, comment lines,
    "M N nz", then nz entries "i j [value [imag]]" (1-based); real / integer
    / pattern (value 1.0) / complex (real part).  Symmetric and hermitian
    files are mirrored.  CSR: per-row counts (an off-diagonal entry of a
    symmetric file also counts for its column's row), exclusive scan, then
    one pass in file order writing at each row's next position, a mirrored
    pair as (i, j) then (j, i).  Returns (m, n, rowptr, colidx, val,
    is_symmetric)."""
    with open(path) as fh:
        lines = fh.read().split("\n")
    banner = lines[0].split()
    kind, sym = banner[3].lower(), banner[4].lower()
    symmetric = sym in ("symmetric", "hermitian")
    k = 1
    while lines[k].strip() == "" or lines[k].lstrip().startswith("%"):
        k += 1
    m, n, nz = (int(t) for t in lines[k].split()[:3])
    k += 1
    ent = []
    while len(ent) < nz:
        t = lines[k].split()
        k += 1
        if not t or t[0].startswith("%"):
            continue
        i, j = int(t[0]) - 1, int(t[1]) - 1
        if kind == "pattern":
            v = 1.0
        elif kind == "integer":
            v = float(int(t[2]))
        else:
            v = float(t[2])
        ent.append((i, j, v))
    cnt = [0] * (m + 1)
    for i, j, _ in ent:
        cnt[i + 1] += 1
        if symmetric and i != j:
            cnt[j + 1] += 1
    for r in range(m):
        cnt[r + 1] += cnt[r]
    rowptr = np.array(cnt, np.int32)
    nxt = list(cnt[:m])
    colidx = np.zeros(cnt[m], np.int32)
    val = np.zeros(cnt[m], np.float64)
    for i, j, v in ent:
        colidx[nxt[i]], val[nxt[i]] = j, v
        nxt[i] += 1
        if symmetric and i != j:
            colidx[nxt[j]], val[nxt[j]] = i, v
            nxt[j] += 1
    return m, n, rowptr, colidx, val, symmetric


def c_rand(seed, count):
    """count values of the C library's rand() after srand(seed) (the
    reference draws its values with it, main.cu:141,166,187,213)."""
    libc = ctypes.CDLL("libc.so.6")
    libc.srand.argtypes = [ctypes.c_uint]
    libc.rand.restype = ctypes.c_int
    libc.srand(seed)
    return [libc.rand() for _ in range(count)]


def glibc_random_model(seed, count):
    """glibc's rand() from its published algorithm (additive feedback,
    r[i] = r[i-3] + r[i-31] mod 2^32 after a Lehmer seeding of 31 words
    (seed 0 taken as 1), 344 outputs discarded, each output r >> 1): an
    independent model to check c_rand against."""
    r = [0] * (344 + count)
    r[0] = (seed & 0xffffffff) or 1  # a zero seed is replaced by 1
    if r[0] >= 2 ** 31:  # the seed is used as a signed 32-bit word
        r[0] -= 2 ** 32
    for i in range(1, 31):
        # C division and remainder (truncating: the first word may be
        # negative)
        hi = r[i - 1] // 127773 if r[i - 1] >= 0 else -((-r[i - 1]) // 127773)
        lo = r[i - 1] - hi * 127773
        w = 16807 * lo - 2836 * hi
        r[i] = w + 2147483647 if w < 0 else w
    for i in range(31, 34):
        r[i] = r[i - 31]
    for i in range(34, 344 + count):
        r[i] = (r[i - 31] + r[i - 3]) & 0xffffffff
    return [(r[i] & 0xffffffff) >> 1 for i in range(344, 344 + count)]


def mm_triangle(m, rowptr, colidx, val, subst, reference_values, seed, nrhs):
    """The driver's factor (recblock-sptrsv/main.cu:150-228): forward keeps
    columns < row (row order) and appends the diagonal; backward puts the
    diagonal first, then columns > row.  reference_values: kept entries
    rand() % 10 + 1 and diagonal 1.0 (rand after srand(seed)); else the
    file's values with the row's diagonal entries summed (ValueError when a
    row has none).  Then x_ref = rand() % 10 + 1 (m * nrhs, column-major)
    and b = T * x_ref per column in T's row order.  Returns (rowptr, colidx,
    val, x_ref, b) of T in CSR."""
    rows = []
    nkept = 0
    for i in range(m):
        cols = colidx[rowptr[i]:rowptr[i + 1]]
        kept = [k for k in range(rowptr[i], rowptr[i + 1])
                if (colidx[k] < i if subst == 0 else colidx[k] > i)]
        nkept += len(kept)
        rows.append(kept)
        if not reference_values and not (cols == i).any():
            raise ValueError("row %d has no diagonal" % i)
    draws = iter(c_rand(seed, nkept + m * nrhs))
    tp, tc, tv = [0], [], []
    for i in range(m):
        d = 1.0 if reference_values else float(
            sum(val[k] for k in range(rowptr[i], rowptr[i + 1]) if colidx[k] == i))
        if subst == 1:
            tc.append(i)
            tv.append(d)
        for k in rows[i]:
            tc.append(int(colidx[k]))
            tv.append(float(next(draws) % 10 + 1) if reference_values else float(val[k]))
        if subst == 0:
            tc.append(i)
            tv.append(d)
        tp.append(len(tc))
    xref = np.array([float(next(draws) % 10 + 1) for _ in range(m * nrhs)])
    b = np.zeros(m * nrhs)
    for r in range(nrhs):
        for i in range(m):
            s = 0.0
            for k in range(tp[i], tp[i + 1]):
                s += tv[k] * xref[r * m + tc[k]]
            b[r * m + i] = s
    return (np.array(tp, np.int32), np.array(tc, np.int32), np.array(tv),
            xref, b)
