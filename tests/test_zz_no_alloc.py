"""rb_solve_noalloc through the C ABI against rb_solve on the same plan.

The entry point takes its per-solve scratch from a fixed device array
instead of allocating, so these tests check that (a) its x is bitwise equal
to rb_solve's x for both schedules and several right-hand sides, (b) two
calls in a row on different plans, the second smaller, both come out right
(the shared array keeps the first call's data), and (c) a plan whose one
column needs more scratch than the array holds is rejected with
RB_ERR_INVALID_ARG instead of running.
"""
import ctypes

import numpy as np
import pytest

from rbsptrsv import Plan, gen
from rbsptrsv._lib import RB_OK, SCHED_BLOCKED, SCHED_FUSED, lib

pytestmark = pytest.mark.gpu

RB_ERR_INVALID_ARG = 1


@pytest.fixture(scope="module")
def torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a device"
    return torch


def noalloc(torch, plan, b):
    """x from rb_solve_noalloc for b (m,) or (m, k) column-major."""
    k = 1 if b.dim() == 1 else b.shape[1]
    bc = b.reshape(plan.m, k).t().contiguous()      # k x m: column-major
    xc = torch.full_like(bc, float("nan"))
    st = lib().rb_solve_noalloc(plan._h, ctypes.c_void_p(bc.data_ptr()),
                                plan.m, ctypes.c_void_p(xc.data_ptr()),
                                plan.m, k, None)
    assert st == RB_OK, lib().rb_last_error().decode()
    plan.check()
    return xc.t().reshape(b.shape)


def reference(torch, plan, b):
    x = plan.solve(b)
    plan.check()
    return x


def rhs(torch, m, k, seed):
    g = np.random.default_rng(seed)
    return torch.as_tensor(g.uniform(-1, 1, (m, k)) if k > 1 else
                           g.uniform(-1, 1, m), device="cuda")


@pytest.mark.parametrize("schedule", [SCHED_BLOCKED, SCHED_FUSED])
def test_noalloc_matches_solve_two_calls(torch, schedule):
    print("noalloc two calls, schedule", schedule, flush=True)
    big = Plan.from_factor(gen.laplacian(300), lv=-1, schedule=schedule)
    small = Plan.from_factor(gen.banded(5000, seed=3), lv=-1,
                             schedule=schedule)
    try:
        for plan, k, seed in ((big, 3, 1), (small, 1, 2)):
            b = rhs(torch, plan.m, k, seed)
            x0 = reference(torch, plan, b).cpu().numpy()
            x1 = noalloc(torch, plan, b).cpu().numpy()
            assert np.isfinite(x1).all()
            assert np.array_equal(x0, x1), (plan.m, k)
    finally:
        big.close()
        small.close()


def test_noalloc_repeated_and_interleaved(torch):
    """The shared array survives interleaving with ordinary solves and
    repeated calls on one plan."""
    f = gen.laplacian(200)
    p = Plan.from_factor(f, lv=-1, schedule=SCHED_FUSED)
    try:
        for seed in range(3):
            b = rhs(torch, p.m, 1, 10 + seed)
            x0 = reference(torch, p, b).cpu().numpy()
            x1 = noalloc(torch, p, b).cpu().numpy()
            x2 = noalloc(torch, p, b).cpu().numpy()
            assert np.array_equal(x0, x1) and np.array_equal(x0, x2)
    finally:
        p.close()


def test_noalloc_rejects_oversized_plan(torch):
    limit = lib().rb_noalloc_limit()
    assert limit == 64 << 20
    # 16 bytes per row (permuted b and x) exceed the array: checked on the
    # host before anything is launched
    m = limit // 16 + 1024
    f = gen.Factor(m, np.arange(m + 1, dtype=np.int32),
                   np.arange(m, dtype=np.int32), np.full(m, 2.0), True,
                   "diagonal")
    p = Plan.from_factor(f, lv=1, schedule=SCHED_BLOCKED)
    try:
        b = torch.ones(p.m, dtype=torch.float64, device="cuda")
        x = torch.zeros_like(b)
        st = lib().rb_solve_noalloc(p._h, ctypes.c_void_p(b.data_ptr()), p.m,
                                    ctypes.c_void_p(x.data_ptr()), p.m, 1,
                                    None)
        assert st == RB_ERR_INVALID_ARG
        assert "scratch" in lib().rb_last_error().decode()
        assert float(x.abs().max()) == 0.0   # nothing ran
    finally:
        p.close()
