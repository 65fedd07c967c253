"""Kernel choice per block of the blocked schedule (rb_opts.block_kernels).

Every triangular-block kernel (the reference selector's, the sync-free grid
kernel, the one-workgroup kernel with its LDS ring) and every lanes-per-row
choice of the square blocks must give x within relative 1e-10 of the
oracle's block-ordered solve (R/recblocking_calculate.h:15-325) and a
relative residual below 1e-12, forward and backward.  The three triangular
kernels sum each row's entries in entry order, so on factors whose level
kernel uses a thread per row they agree bit for bit.  The one-workgroup
kernel applies only when every in-block dependency lies within its ring;
blocks whose dependencies reach further keep the other kernels.
"""
import numpy as np
import pytest

from helpers import rel_err, residual
from rbsptrsv import gen

pytestmark = pytest.mark.gpu

X_TOL = 1e-10
RES_TOL = 1e-12
REF, SYNCFREE, GROUP = 0, 1, 2

CASES = [
    ("banded-20000", lambda: gen.banded(20000, seed=4)),
    ("banded-20000-upper", lambda: gen.banded(20000, seed=4).transpose()),
    ("laplacian-120", lambda: gen.laplacian(120)),
    ("laplacian-120-upper", lambda: gen.laplacian(120, lower=False)),
    ("powerlaw-60000", lambda: gen.powerlaw(60000, seed=5)),
    ("random-5000-long", lambda: gen.random_lower(5000, 8, seed=6)),
]


@pytest.fixture(scope="module")
def torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a device"
    return torch


def _solve(torch, f, b, **kw):
    from rbsptrsv import Plan
    p = Plan.from_factor(f, lv=-1, schedule=1, **kw)
    try:
        x = p.solve(torch.as_tensor(b, device="cuda"))
        p.check()
        return x.cpu().numpy(), p.block_kernels(), p.blocks()
    finally:
        p.close()


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_block_kernels_against_oracle(oracle, torch, name, make):
    print("block kernels", name, flush=True)
    f = make()
    lv = oracle.lv_auto(f.n)
    re, part, bnz, an = oracle.full_pipeline(f.colptr, f.rowidx, f.val, f.n,
                                             f.substitution, lv)
    b = gen.rhs(f.n, seed=21)
    xo = oracle.solve_blocks(re, f.n, lv, f.substitution, part, b)
    xs = {}
    for mode in (1, 2, 3, 0):
        x, bk, bl = _solve(torch, f, b, block_kernels=mode)
        tri = np.arange(0, len(bl["method"]), 2)
        nz = tri[(bl["method"][tri] != 0) & (bl["blk_m"][tri] > 0)]
        if mode == 1:
            assert (bk["kern"] == REF).all()
        elif mode == 2:
            assert (bk["kern"][nz] == SYNCFREE).all()
        elif mode == 3:
            # the one-workgroup kernel wherever its ring reaches
            assert set(bk["kern"][nz]) <= {REF, GROUP}
        sq = np.arange(1, len(bl["method"]), 2)
        live = sq[bl["method"][sq] >= 0]
        assert set(bk["lanes"][live]) <= {1, 2, 4, 8, 16, 32, 64}
        e = rel_err(x, xo)
        print("  mode", mode, "kern", bk["kern"][tri].tolist(), "rel", e,
              flush=True)
        assert e < X_TOL, (mode, e)
        assert residual(f, x, b) < RES_TOL
        xs[mode] = x
    # entry-order sums: with the same lanes-per-row rule for the square
    # blocks, the sync-free and one-workgroup kernels agree bit for bit, and
    # so does the reference's kernel where its level kernel keeps a thread
    # per row (not on the power-law factor's long rows)
    if "powerlaw" not in name:
        assert np.array_equal(xs[2], xs[3])
        assert np.array_equal(xs[1], xs[2])


def test_group_kernel_runs_on_deep_blocks(torch):
    """The banded factor's triangular blocks are deep and narrow (hundreds
    of levels): every one of them fits the ring, mode 3 runs them all with
    the one-workgroup kernel and the tuner may pick it."""
    f = gen.banded(30000, seed=7)
    b = gen.rhs(f.n, seed=3)
    x3, bk3, bl = _solve(torch, f, b, block_kernels=3)
    tri = np.arange(0, len(bl["method"]), 2)
    nz = tri[(bl["method"][tri] != 0) & (bl["blk_m"][tri] > 0)]
    assert len(nz) > 0 and (bk3["kern"][nz] == GROUP).all()
    x1, _, _ = _solve(torch, f, b, block_kernels=1)
    assert np.array_equal(x1, x3)
    assert residual(f, x3, b) < RES_TOL


def test_group_kernel_refused_beyond_ring(torch):
    """Dependencies farther back than the ring reaches keep a triangular
    block off the one-workgroup kernel (a chain whose row i depends on row
    i - 1 and on row i - 9000)."""
    from helpers import csc_from_deps
    n = 30000
    deps = {i: [i - 1] + ([i - 9000] if i >= 9000 else []) for i in range(1, n)}
    f = csc_from_deps(n, deps, seed=8)
    b = gen.rhs(n, seed=4)
    x, bk, bl = _solve(torch, f, b, block_kernels=3)
    tri = np.arange(0, len(bl["method"]), 2)
    nz = tri[(bl["method"][tri] != 0) & (bl["blk_m"][tri] > 0)]
    assert (bk["kern"][nz] != GROUP).any()
    assert residual(f, x, b) < RES_TOL


def test_group_kernel_repeated_solves_stay_clear(torch):
    """A wave of the one-workgroup kernel may run ahead of the others only
    so far that the ring slots it overwrites have no reader left: twenty
    solves of the 1341^2 Laplacian lower factor (1.8M rows, triangular
    blocks of ~1300-row levels, every one on the kernel) leave the
    watchdog clear and give the same x each time.  (Before the round gate,
    waves ran ahead, overwrote slots other rows still waited for, and those
    rows waited out the watchdog: the same factor's plan build took
    111 s.)"""
    from rbsptrsv import Plan
    f = gen.laplacian(1341)
    b = gen.rhs(f.n, seed=13)
    p = Plan.from_factor(f, lv=-1, schedule=1, block_kernels=3)
    try:
        bk, bl = p.block_kernels(), p.blocks()
        tri = np.arange(0, len(bl["method"]), 2)
        nz = tri[(bl["method"][tri] != 0) & (bl["blk_m"][tri] > 0)]
        assert len(nz) > 0 and (bk["kern"][nz] == GROUP).all()
        db = torch.as_tensor(b, device="cuda")
        x0 = None
        for k in range(20):
            x = p.solve(db)
            p.check()
            if x0 is None:
                x0 = x.clone()
            else:
                assert torch.equal(x, x0), k
        assert residual(f, x0.cpu().numpy(), b) < RES_TOL
    finally:
        p.close()
