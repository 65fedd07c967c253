"""MI355X-native recursive-block sparse triangular solve.

The compute path is librecblock.so (hand-written HIP for gfx950, C ABI in
include/recblock.h); this package is its Python front end plus seeded
workload generators.
"""
from . import gen  # noqa: F401
from ._lib import (EXPORTS, LIB_PATH, RbError, SCHED_AUTO, SCHED_BLOCKED,  # noqa: F401
                   SCHED_FUSED, build, lib)
from .api import (Batch, Plan, lv_auto, recblocking_solver,  # noqa: F401
                  recblocking_solver_cuda, set_bench_repeat)
