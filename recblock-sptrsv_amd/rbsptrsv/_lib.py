"""ctypes binding of librecblock.so (the C ABI declared in include/recblock.h).

The library is built in-tree (``recblock-sptrsv_amd/librecblock.so``) by
``build()``; loading fails loudly when it is missing -- there is no CPU
fallback.
"""
import ctypes
import os
import subprocess

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# RECBLOCK_LIB names another build of the library to load (a variant built
# with different compile-time settings); the in-tree build by default
LIB_PATH = os.environ.get("RECBLOCK_LIB") or os.path.join(PKG_ROOT,
                                                          "librecblock.so")
CSRC = os.path.join(PKG_ROOT, "csrc")

# Every exported symbol of include/recblock.h.
EXPORTS = [
    "recblocking_solver_cuda", "recblocking_solver", "rb_last_status",
    "rb_last_error", "rb_set_bench_repeat", "rb_lv_auto", "rb_opts_default",
    "rb_plan_create", "rb_solve", "rb_plan_destroy", "rb_plan_check",
    "rb_plan_info_get",
    "rb_plan_blocks", "rb_plan_levels", "rb_plan_tri_levels",
    "rb_plan_block_csr", "rb_plan_permuted_csc", "rb_plan_trace",
    "rb_plan_set_timing", "rb_plan_kernel_ms", "rb_solve_f32ir",
    "rb_batch_create", "rb_batch_solve", "rb_batch_solve_f32ir",
    "rb_batch_plan", "rb_batch_destroy",
    "rb_mmio_info", "rb_mmio_data", "rb_mmio_triangle",
    "rb_solve_noalloc", "rb_noalloc_limit", "rb_plan_block_kernels",
]

RB_OK = 0
STATUS_NAMES = {0: "RB_OK", 1: "RB_ERR_INVALID_ARG", 2: "RB_ERR_NOT_SQUARE",
                3: "RB_ERR_BAD_STRUCTURE", 4: "RB_ERR_MISSING_DIAG",
                5: "RB_ERR_ZERO_DIAG", 6: "RB_ERR_HIP", 7: "RB_ERR_NO_MEMORY",
                8: "RB_ERR_INTERNAL", 9: "RB_ERR_NOT_CONVERGED"}

SCHED_AUTO, SCHED_BLOCKED, SCHED_FUSED = 0, 1, 2


class RbError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (STATUS_NAMES.get(status, status), msg))
        self.status = status


class rb_opts(ctypes.Structure):
    _fields_ = [("schedule", ctypes.c_int), ("device_id", ctypes.c_int),
                ("fused_waves", ctypes.c_int), ("no_graph", ctypes.c_int),
                ("mixed", ctypes.c_int), ("tune_rhs", ctypes.c_int),
                ("canonicalize", ctypes.c_int),
                ("watchdog_spins", ctypes.c_int),
                ("block_kernels", ctypes.c_int)]


class rb_plan_info(ctypes.Structure):
    _fields_ = [("m", ctypes.c_int), ("nnz", ctypes.c_int),
                ("substitution", ctypes.c_int), ("lv", ctypes.c_int),
                ("nblocks", ctypes.c_int), ("nlevel", ctypes.c_int),
                ("schedule", ctypes.c_int), ("preprocess_ms", ctypes.c_double),
                ("device_bytes", ctypes.c_size_t),
                ("phase_ms", ctypes.c_double * 4),
                ("tune_ms", ctypes.c_double * 2)]


def build(jobs=8):
    """Compile librecblock.so for gfx950 (hipcc cross-compiles without a
    GPU).  Always recompiles."""
    subprocess.check_call(["make", "-s", "-B", "-j%d" % jobs, "-C", CSRC])
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("librecblock.so is not built (%s); run "
                           "__graft_entry__.build()" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp, ip, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_void_p
    i = ctypes.c_int
    L.rb_last_status.restype = i
    L.rb_last_error.restype = ctypes.c_char_p
    L.rb_set_bench_repeat.argtypes = [i]
    L.rb_set_bench_repeat.restype = i
    L.rb_lv_auto.argtypes = [i, i]
    L.rb_lv_auto.restype = i
    L.rb_opts_default.argtypes = [ctypes.POINTER(rb_opts)]
    L.rb_plan_create.argtypes = [vp, vp, vp, i, i, i, i, i,
                                 ctypes.POINTER(rb_opts), vp,
                                 ctypes.POINTER(vp)]
    L.rb_plan_create.restype = i
    L.rb_solve.argtypes = [vp, vp, i, vp, i, i, vp]
    L.rb_solve.restype = i
    L.rb_solve_noalloc.argtypes = [vp, vp, i, vp, i, i, vp]
    L.rb_solve_noalloc.restype = i
    L.rb_noalloc_limit.argtypes = []
    L.rb_noalloc_limit.restype = ctypes.c_size_t
    L.rb_plan_check.argtypes = [vp, vp]
    L.rb_plan_check.restype = i
    L.rb_plan_destroy.argtypes = [vp]
    L.rb_plan_destroy.restype = i
    L.rb_plan_info_get.argtypes = [vp, ctypes.POINTER(rb_plan_info)]
    L.rb_plan_info_get.restype = i
    L.rb_plan_blocks.argtypes = [vp] + [vp] * 15
    L.rb_plan_blocks.restype = i
    L.rb_plan_block_kernels.argtypes = [vp, vp, vp, vp]
    L.rb_plan_block_kernels.restype = i
    L.rb_plan_levels.argtypes = [vp, vp, vp]
    L.rb_plan_levels.restype = i
    L.rb_plan_tri_levels.argtypes = [vp, i, vp, vp, vp]
    L.rb_plan_tri_levels.restype = i
    L.rb_plan_block_csr.argtypes = [vp, i, vp, vp, vp]
    L.rb_plan_block_csr.restype = i
    L.rb_plan_permuted_csc.argtypes = [vp, vp, vp, vp]
    L.rb_plan_permuted_csc.restype = i
    L.rb_plan_trace.argtypes = [vp, vp, ctypes.POINTER(i)]
    L.rb_plan_trace.restype = i
    L.rb_plan_set_timing.argtypes = [vp, i]
    L.rb_plan_set_timing.restype = i
    L.rb_plan_kernel_ms.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    L.rb_plan_kernel_ms.restype = i
    L.rb_solve_f32ir.argtypes = [vp, vp, i, vp, i, i, ctypes.c_double, i, vp,
                                 ctypes.POINTER(i),
                                 ctypes.POINTER(ctypes.c_double)]
    L.rb_solve_f32ir.restype = i
    L.rb_batch_create.argtypes = [i, vp, vp, vp, vp, vp, i, i,
                                  ctypes.POINTER(rb_opts), vp,
                                  ctypes.POINTER(vp)]
    L.rb_batch_create.restype = i
    L.rb_batch_solve.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.rb_batch_solve.restype = i
    L.rb_batch_solve_f32ir.argtypes = [vp, vp, vp, vp, vp, vp,
                                       ctypes.c_double, i, vp,
                                       ctypes.POINTER(i),
                                       ctypes.POINTER(ctypes.c_double)]
    L.rb_batch_solve_f32ir.restype = i
    L.rb_batch_plan.argtypes = [vp]
    L.rb_batch_plan.restype = vp
    L.rb_batch_destroy.argtypes = [vp]
    L.rb_batch_destroy.restype = i
    L.recblocking_solver_cuda.argtypes = [vp, vp, vp, i, i, i, vp, vp, i, i,
                                          ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_double)]
    L.recblocking_solver_cuda.restype = None
    L.recblocking_solver.argtypes = [vp, vp, vp, i, i, i, vp, vp, vp, i, i, i,
                                     ctypes.POINTER(ctypes.c_double)]
    L.recblocking_solver.restype = None
    cp = ctypes.c_char_p
    L.rb_mmio_info.argtypes = [ctypes.POINTER(i)] * 4 + [cp]
    L.rb_mmio_info.restype = i
    L.rb_mmio_data.argtypes = [vp, vp, vp, cp]
    L.rb_mmio_data.restype = i
    L.rb_mmio_triangle.argtypes = [i, vp, vp, vp, i, i, ctypes.c_uint, i,
                                   vp, vp, vp, ctypes.POINTER(i), vp, vp]
    L.rb_mmio_triangle.restype = i
    _lib = L
    return L


def check(status):
    if status != RB_OK:
        raise RbError(status, lib().rb_last_error().decode())
    return status
