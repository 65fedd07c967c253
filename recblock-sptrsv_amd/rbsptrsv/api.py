"""Host-side mirror of the reference interface over the C ABI.

``recblocking_solver_cuda`` / ``recblocking_solver`` keep the reference's
names and argument meaning (R/recblocking_solver_cuda.h:21-32,
R/recblocking_solver.h:16-28); ``Plan`` is the split preprocessing / solve
interface.  Device memory and streams come from torch (ROCm build); the
compute is librecblock.so.
"""
import ctypes
import sys

import numpy as np

from . import _lib
from ._lib import check, lib


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no GPU visible: librecblock needs a gfx950 device")
    return torch


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _dev_vec(t, dtype, n=None, what="array"):
    """A 1-D, contiguous device tensor of ``dtype`` holding ``t`` (a host
    array or a torch tensor on any device).  Tensors already in that form
    pass through; others are copied.  The C API reads raw device pointers,
    so a host pointer or a wrong element width must never reach it."""
    torch = _torch()
    if not torch.is_tensor(t):
        np_t = np.int32 if dtype == torch.int32 else np.float64
        t = torch.as_tensor(np.ascontiguousarray(t, np_t), device="cuda")
    if t.dim() != 1:
        raise ValueError("%s must be one-dimensional" % what)
    if not t.is_cuda or t.dtype != dtype or t.stride(0) != 1:
        t = t.to(device="cuda", dtype=dtype).contiguous()
    if n is not None and t.numel() != n:
        raise ValueError("%s has %d entries, expected %d" % (what, t.numel(), n))
    return t


def _rhs(torch, b, m, what="b"):
    """(tensor, ld, k) of right-hand sides b: shape (m,) or column-major
    (m, k), float64, on the device.  b is copied into that form when it is
    not (a host array, another dtype or device, a row-major block)."""
    if not torch.is_tensor(b):
        b = torch.as_tensor(np.asarray(b, np.float64), device="cuda")
    if b.dim() not in (1, 2) or b.shape[0] != m:
        raise ValueError("%s must have shape (%d,) or (%d, k)" % (what, m, m))
    if b.dim() == 1:
        if not b.is_cuda or b.dtype != torch.float64 or b.stride(0) != 1:
            b = b.to(device="cuda", dtype=torch.float64).contiguous()
        return b, m, 1
    k = b.shape[1]
    if (not b.is_cuda or b.dtype != torch.float64 or b.stride(0) != 1 or
            (k > 1 and b.stride(1) < m)):
        b = b.to(device="cuda", dtype=torch.float64).t().contiguous().t()
    return b, (b.stride(1) if k > 1 else m), k


def _out(torch, x, m, k, like_2d):
    """The output tensor for k solutions of length m: a new column-major
    block, or the caller's x after checking it can be written in place."""
    if x is None:
        if not like_2d:
            return torch.empty(m, dtype=torch.float64, device="cuda"), m
        return torch.empty((k, m), dtype=torch.float64, device="cuda").t(), m
    if not torch.is_tensor(x) or not x.is_cuda or x.dtype != torch.float64:
        raise ValueError("x must be a float64 CUDA tensor")
    if x.shape[0] != m or x.stride(0) != 1 or (x.dim() == 2 and x.shape[1] != k) \
            or x.dim() not in (1, 2) or (x.dim() == 1 and k != 1):
        raise ValueError("x must be (m,) or column-major (m, k) matching b")
    ld = x.stride(1) if x.dim() == 2 and k > 1 else m
    if x.dim() == 2 and k > 1 and ld < m:
        raise ValueError("x columns overlap")
    return x, ld


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
                 canonicalize=False, watchdog_spins=0, block_kernels=0):
        torch = _torch()
        self._keep = []
        dcp = _dev_vec(colptr, torch.int32, what="colptr")
        self.m = int(dcp.numel()) - 1
        if self.m < 0:
            raise ValueError("colptr needs m + 1 >= 1 entries")
        dri = _dev_vec(rowidx, torch.int32, what="rowidx")
        self.nnz = int(dri.numel())
        dv = _dev_vec(val, torch.float64, self.nnz, what="val")
        opts = _lib.rb_opts()
        lib().rb_opts_default(ctypes.byref(opts))
        opts.schedule = schedule
        opts.device_id = device_id
        opts.fused_waves = fused_waves
        opts.no_graph = 0 if graph else 1
        opts.mixed = 1 if mixed else 0
        opts.tune_rhs = tune_rhs
        opts.canonicalize = 1 if canonicalize else 0
        opts.watchdog_spins = int(watchdog_spins)
        opts.block_kernels = int(block_kernels)
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
        # at interpreter exit the HIP runtime (and a profiler attached to
        # it) may already be finalised: leave the memory to the process end
        if sys.is_finalizing():
            return
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
        """Solve for b: shape (m,) or (m, k), column-major on the device
        (b.t() contiguous); other forms are copied first.  Returns x (the
        caller's x when given: float64, on the device, same shape)."""
        torch = _torch()
        b, ldb, nrhs = _rhs(torch, b, self.m)
        x, ldx = _out(torch, x, self.m, nrhs, b.dim() == 2)
        check(lib().rb_solve(self._h, _ptr(b), ldb, _ptr(x), ldx, nrhs,
                             _stream_ptr(stream)))
        return x

    def solve_f32ir(self, b, x=None, tol=1e-12, max_iter=10, stream=None,
                    report=True):
        """FP32 solves with FP64 refinement (plan built with mixed=True).
        b and x as in solve().  Returns (x, iterations, relative
        residual).  report=False: no host synchronisation; returns (x, None,
        None) and a column that missed tol shows in check()."""
        torch = _torch()
        b, ldb, nrhs = _rhs(torch, b, self.m)
        x, ldx = _out(torch, x, self.m, nrhs, b.dim() == 2)
        it = ctypes.c_int(0)
        rr = ctypes.c_double(0.0)
        check(lib().rb_solve_f32ir(self._h, _ptr(b), ldb, _ptr(x), ldx, nrhs,
                                   ctypes.c_double(tol), max_iter,
                                   _stream_ptr(stream),
                                   ctypes.byref(it) if report else None,
                                   ctypes.byref(rr) if report else None))
        return (x, it.value, rr.value) if report else (x, None, None)

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

    def block_kernels(self):
        """Kernels the blocked schedule runs per block: dict of kern
        (triangular: 0 reference selector's, 1 sync-free grid, 2 one
        workgroup), lanes (square: lanes per row) and tuned_ms."""
        nb = self.nblocks
        kern = np.zeros(nb, np.int32)
        lanes = np.zeros(nb, np.int32)
        ms = np.zeros(nb, np.float64)
        check(lib().rb_plan_block_kernels(self._h, kern.ctypes.data,
                                          lanes.ctypes.data, ms.ctypes.data))
        return {"kern": kern, "lanes": lanes, "tuned_ms": ms}

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
            cp = _dev_vec(f.colptr, torch.int32, what="colptr")
            ri = _dev_vec(f.rowidx, torch.int32, what="rowidx")
            v = _dev_vec(f.val, torch.float64, ri.numel(), what="val")
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
        # at interpreter exit the HIP runtime (and a profiler attached to
        # it) may already be finalised: leave the memory to the process end
        if sys.is_finalizing():
            return
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
        if len(bs) != nf or (xs is not None and len(xs) != nf):
            raise ValueError("one entry of bs (and xs) per factor")
        bs = list(bs)
        xs = [None] * nf if xs is None else list(xs)
        self._call_keep = []  # converted right-hand sides live until return
        bp, xp, lb, lx, nr = [], [], [], [], []
        for f in range(nf):
            if bs[f] is None:
                bp.append(None), xp.append(None)
                lb.append(self.m[f]), lx.append(self.m[f]), nr.append(0)
                continue
            b, ldb, k = _rhs(torch, bs[f], self.m[f])
            bs[f] = b
            xs[f], ldx = _out(torch, xs[f], self.m[f], k, b.dim() == 2)
            self._call_keep.append(b)
            bp.append(b.data_ptr()), xp.append(xs[f].data_ptr())
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
