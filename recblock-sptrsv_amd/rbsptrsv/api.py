"""Host-side mirror of the reference interface over the C ABI.

``recblocking_solver_cuda`` / ``recblocking_solver`` keep the reference's
names and argument meaning (R/recblocking_solver_cuda.h:21-32,
R/recblocking_solver.h:16-28); ``Plan`` is the split preprocessing / solve
interface.  Device memory and streams come from torch (ROCm build); the
compute is librecblock.so.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no GPU visible: librecblock needs a gfx950 device")
    return torch


def _dev_i32(a):
    torch = _torch()
    return torch.as_tensor(np.ascontiguousarray(a, np.int32), device="cuda")


def _dev_f64(a):
    torch = _torch()
    return torch.as_tensor(np.ascontiguousarray(a, np.float64), device="cuda")


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream_ptr(stream):
    if stream is None:
        return None
    return ctypes.c_void_p(stream.cuda_stream)


class Plan:
    """Preprocessed factor resident in HBM.

    ``colptr``/``rowidx``/``val`` are host arrays or torch device tensors of
    a square CSC (lower for substitution 0, upper for 1).
    """

    def __init__(self, colptr, rowidx, val, substitution=0, lv=-1,
                 schedule=_lib.SCHED_AUTO, stream=None, device_id=0,
                 fused_waves=0, graph=True, mixed=False, tune_rhs=1,
                 canonicalize=False):
        torch = _torch()
        self._keep = []
        dcp = colptr if torch.is_tensor(colptr) else _dev_i32(colptr)
        dri = rowidx if torch.is_tensor(rowidx) else _dev_i32(rowidx)
        dv = val if torch.is_tensor(val) else _dev_f64(val)
        self.m = int(dcp.numel()) - 1
        self.nnz = int(dri.numel())
        opts = _lib.rb_opts()
        lib().rb_opts_default(ctypes.byref(opts))
        opts.schedule = schedule
        opts.device_id = device_id
        opts.fused_waves = fused_waves
        opts.no_graph = 0 if graph else 1
        opts.mixed = 1 if mixed else 0
        opts.tune_rhs = tune_rhs
        opts.canonicalize = 1 if canonicalize else 0
        h = ctypes.c_void_p()
        check(lib().rb_plan_create(_ptr(dcp), _ptr(dri), _ptr(dv), self.m,
                                   self.m, self.nnz, substitution, lv,
                                   ctypes.byref(opts), _stream_ptr(stream),
                                   ctypes.byref(h)))
        self._h = h
        self.substitution = substitution
        info = self.info()
        self.nnz = info.nnz  # the merged count when canonicalized
        self.lv = info.lv
        self.nblocks = info.nblocks
        self.nlevel = info.nlevel

    @classmethod
    def from_factor(cls, f, lv=-1, **kw):
        return cls(f.colptr, f.rowidx, f.val, f.substitution, lv, **kw)

    def close(self):
        if getattr(self, "_h", None):
            lib().rb_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        inf = _lib.rb_plan_info()
        check(lib().rb_plan_info_get(self._h, ctypes.byref(inf)))
        return inf

    # ------------------------------------------------------------ solve
    def solve(self, b, x=None, stream=None):
        """Solve for b (torch device tensor, shape (m,) or (m, k) with
        column-major storage, i.e. b.t() contiguous).  Returns x."""
        torch = _torch()
        if not torch.is_tensor(b):
            b = _dev_f64(b)
        if b.dim() == 1:
            nrhs, ldb = 1, self.m
            if x is None:
                x = torch.empty_like(b)
            ldx = self.m
        else:
            nrhs = b.shape[1]
            if b.stride(0) != 1:
                raise ValueError("b must be column-major (b.t() contiguous)")
            ldb = b.stride(1)
            if x is None:
                x = torch.empty((nrhs, self.m), dtype=b.dtype,
                                device=b.device).t()
            ldx = x.stride(1)
        check(lib().rb_solve(self._h, _ptr(b), ldb, _ptr(x), ldx, nrhs,
                             _stream_ptr(stream)))
        return x

    def solve_f32ir(self, b, x=None, tol=1e-12, max_iter=10, stream=None):
        """FP32 solves with FP64 refinement (plan built with mixed=True).
        b as in solve().  Returns (x, iterations, relative residual)."""
        torch = _torch()
        if not torch.is_tensor(b):
            b = _dev_f64(b)
        if b.dim() == 1:
            nrhs, ldb = 1, self.m
            if x is None:
                x = torch.empty_like(b)
            ldx = self.m
        else:
            nrhs = b.shape[1]
            if b.stride(0) != 1:
                raise ValueError("b must be column-major (b.t() contiguous)")
            ldb = b.stride(1)
            if x is None:
                x = torch.empty((nrhs, self.m), dtype=b.dtype,
                                device=b.device).t()
            ldx = x.stride(1)
        it = ctypes.c_int(0)
        rr = ctypes.c_double(0.0)
        check(lib().rb_solve_f32ir(self._h, _ptr(b), ldb, _ptr(x), ldx, nrhs,
                                   ctypes.c_double(tol), max_iter,
                                   _stream_ptr(stream), ctypes.byref(it),
                                   ctypes.byref(rr)))
        return x, it.value, rr.value

    def set_timing(self, enable=True):
        """Record events around each solve's main launch (see
        rb_plan_set_timing)."""
        check(lib().rb_plan_set_timing(self._h, 1 if enable else 0))

    def kernel_ms(self):
        """Duration of the last timed main launch, in ms."""
        ms = ctypes.c_double(0.0)
        check(lib().rb_plan_kernel_ms(self._h, ctypes.byref(ms)))
        return ms.value

    def check(self, stream=None):
        """Synchronise and raise if the device watchdog fired."""
        check(lib().rb_plan_check(self._h, _stream_ptr(stream)))

    # ------------------------------------------------------- inspection
    def blocks(self):
        nb = self.nblocks
        names = ["blk_m", "blk_n", "blk_nnz", "up", "down", "left", "right",
                 "method", "nlv", "nnzr", "m_new", "lenmax", "longrow",
                 "longlen"]
        arrs = {k: np.zeros(nb, np.int32) for k in names}
        er = np.zeros(nb, np.float64)
        check(lib().rb_plan_blocks(self._h, *[arrs[k].ctypes.data for k in names],
                                   er.ctypes.data))
        arrs["empty_ratio"] = er
        return arrs

    def levels(self):
        li = np.zeros(max(self.m, 1), np.int32)
        lp = np.zeros(self.nlevel + 1, np.int32)
        check(lib().rb_plan_levels(self._h, li.ctypes.data, lp.ctypes.data))
        return li[:self.m], lp

    def tri_levels(self, t):
        bl = self.blocks()
        bm = int(bl["blk_m"][2 * t])
        nlv = int(bl["nlv"][2 * t])
        lp = np.zeros(nlv + 1, np.int32)
        li = np.zeros(max(bm, 1), np.int32)
        nl = np.zeros(max(nlv, 1), np.int32)
        check(lib().rb_plan_tri_levels(self._h, t, lp.ctypes.data,
                                       li.ctypes.data, nl.ctypes.data))
        return lp, li[:bm], nl[:nlv]

    def block_csr(self, i):
        bl = self.blocks()
        bm, nz = int(bl["blk_m"][i]), int(bl["blk_nnz"][i])
        rp = np.zeros(bm + 1, np.int32)
        ci = np.zeros(max(nz, 1), np.int32)
        vv = np.zeros(max(nz, 1), np.float64)
        check(lib().rb_plan_block_csr(self._h, i, rp.ctypes.data,
                                      ci.ctypes.data, vv.ctypes.data))
        return rp, ci[:nz], vv[:nz]

    def trace(self):
        """Fused-schedule counters of the last solve, shape (waves, 16), or
        None (see rb_plan_trace)."""
        nw = ctypes.c_int(0)
        check(lib().rb_plan_trace(self._h, None, ctypes.byref(nw)))
        if nw.value == 0:
            return None
        out = np.zeros((nw.value, 16), np.uint64)
        check(lib().rb_plan_trace(self._h, out.ctypes.data, ctypes.byref(nw)))
        return out

    def permuted_csc(self):
        cp = np.zeros(self.m + 1, np.int32)
        ri = np.zeros(max(self.nnz, 1), np.int32)
        vv = np.zeros(max(self.nnz, 1), np.float64)
        check(lib().rb_plan_permuted_csc(self._h, cp.ctypes.data,
                                         ri.ctypes.data, vv.ctypes.data))
        return cp, ri[:self.nnz], vv[:self.nnz]


def _colmajor(torch, b, m):
    """(pointer, ld, k) of a (m,) or column-major (m, k) device tensor."""
    if b.dim() == 1:
        return _ptr(b), m, 1
    if b.stride(0) != 1:
        raise ValueError("right-hand sides must be column-major (b.t() contiguous)")
    return _ptr(b), b.stride(1), b.shape[1]


class Batch:
    """Many independent factors preprocessed as one block-diagonal plan and
    solved together (rb_batch_*).  ``factors``: objects with ``colptr``,
    ``rowidx``, ``val`` (host arrays or device tensors) and
    ``substitution`` (all equal)."""

    def __init__(self, factors, lv=-1, schedule=_lib.SCHED_AUTO, stream=None,
                 device_id=0, fused_waves=0, mixed=False, tune_rhs=1):
        torch = _torch()
        if not factors:
            raise ValueError("empty batch")
        subst = factors[0].substitution
        if any(f.substitution != subst for f in factors):
            raise ValueError("all factors of a batch share one substitution")
        self._keep = []
        cps, ris, vs, ms, zs = [], [], [], [], []
        for f in factors:
            cp = f.colptr if torch.is_tensor(f.colptr) else _dev_i32(f.colptr)
            ri = f.rowidx if torch.is_tensor(f.rowidx) else _dev_i32(f.rowidx)
            v = f.val if torch.is_tensor(f.val) else _dev_f64(f.val)
            self._keep += [cp, ri, v]
            cps.append(cp.data_ptr())
            ris.append(ri.data_ptr())
            vs.append(v.data_ptr())
            ms.append(int(cp.numel()) - 1)
            zs.append(int(ri.numel()))
        nf = len(factors)
        P = ctypes.c_void_p * nf
        I = ctypes.c_int * nf
        opts = _lib.rb_opts()
        lib().rb_opts_default(ctypes.byref(opts))
        opts.schedule = schedule
        opts.device_id = device_id
        opts.fused_waves = fused_waves
        opts.mixed = 1 if mixed else 0
        opts.tune_rhs = tune_rhs
        h = ctypes.c_void_p()
        check(lib().rb_batch_create(nf, P(*cps), P(*ris), P(*vs), I(*ms),
                                    I(*zs), subst, lv, ctypes.byref(opts),
                                    _stream_ptr(stream), ctypes.byref(h)))
        self._h = h
        self.m = ms
        self.nf = nf
        self.substitution = subst

    def close(self):
        if getattr(self, "_h", None):
            lib().rb_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        inf = _lib.rb_plan_info()
        check(lib().rb_plan_info_get(lib().rb_batch_plan(self._h),
                                     ctypes.byref(inf)))
        return inf

    def _tables(self, bs, xs):
        torch = _torch()
        nf = self.nf
        if xs is None:
            xs = [None if b is None else
                  (torch.empty_like(b) if b.dim() == 1 else
                   torch.empty((b.shape[1], b.shape[0]), dtype=b.dtype,
                               device=b.device).t()) for b in bs]
        bp, xp, lb, lx, nr = [], [], [], [], []
        for f in range(nf):
            if bs[f] is None:
                bp.append(None), xp.append(None)
                lb.append(self.m[f]), lx.append(self.m[f]), nr.append(0)
                continue
            pb, ldb, k = _colmajor(torch, bs[f], self.m[f])
            px, ldx, _ = _colmajor(torch, xs[f], self.m[f])
            bp.append(pb.value), xp.append(px.value)
            lb.append(ldb), lx.append(ldx), nr.append(k)
        P = ctypes.c_void_p * nf
        I = ctypes.c_int * nf
        return xs, (P(*bp), I(*lb), P(*xp), I(*lx), I(*nr))

    def solve(self, bs, xs=None, stream=None):
        """bs[f]: factor f's right-hand sides, (m_f,) or column-major
        (m_f, k_f) device tensor, or None.  Returns the list of solutions."""
        xs, t = self._tables(bs, xs)
        check(lib().rb_batch_solve(self._h, *t, _stream_ptr(stream)))
        return xs

    def solve_f32ir(self, bs, xs=None, tol=1e-12, max_iter=10, stream=None):
        """FP32 solves with FP64 refinement (batch built with mixed=True).
        Returns (solutions, iterations, relative residual)."""
        xs, t = self._tables(bs, xs)
        it = ctypes.c_int(0)
        rr = ctypes.c_double(0.0)
        check(lib().rb_batch_solve_f32ir(self._h, *t, ctypes.c_double(tol),
                                         max_iter, _stream_ptr(stream),
                                         ctypes.byref(it), ctypes.byref(rr)))
        return xs, it.value, rr.value


def recblocking_solver_cuda(d_cscColPtrTR, d_cscRowIdxTR, d_cscValTR, m, n,
                            nnzTR, d_x, d_b, substitution, lv):
    """Reference entry point (R/recblocking_solver_cuda.h:21-32) on torch
    device tensors.  Returns (cal_time_ms, preprocess_time_ms); raises on a
    failed call instead of hanging."""
    ct, pt = ctypes.c_double(0.0), ctypes.c_double(0.0)
    lib().recblocking_solver_cuda(_ptr(d_cscColPtrTR), _ptr(d_cscRowIdxTR),
                                  _ptr(d_cscValTR), m, n, nnzTR, _ptr(d_x),
                                  _ptr(d_b), substitution, lv,
                                  ctypes.byref(ct), ctypes.byref(pt))
    check(lib().rb_last_status())
    return ct.value, pt.value


def recblocking_solver(cscColPtr, cscRowIdx, cscVal, m, n, nnz, b, x, x_ref,
                       rhs, lv, substitution):
    """Reference host-pointer entry point (R/recblocking_solver.h:16-28) on
    numpy arrays; x is written in place.  Returns cal_time_ms."""
    cp = np.ascontiguousarray(cscColPtr, np.int32)
    ri = np.ascontiguousarray(cscRowIdx, np.int32)
    vv = np.ascontiguousarray(cscVal, np.float64)
    bb = np.ascontiguousarray(b, np.float64)
    assert x.dtype == np.float64 and x.flags["C_CONTIGUOUS"]
    ct = ctypes.c_double(0.0)
    lib().recblocking_solver(cp.ctypes.data, ri.ctypes.data, vv.ctypes.data,
                             m, n, nnz, bb.ctypes.data, x.ctypes.data, None,
                             rhs, lv, substitution, ctypes.byref(ct))
    check(lib().rb_last_status())
    return ct.value


def set_bench_repeat(n):
    return lib().rb_set_bench_repeat(int(n))


def lv_auto(m, device_id=0):
    return lib().rb_lv_auto(int(m), int(device_id))
