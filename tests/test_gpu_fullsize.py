"""GPU parity at the benchmark configurations' real sizes, against the
oracle on the same seeded inputs.

* banded n = 1e5 (8 draws in a 2000-column window), forward;
* Laplacian 1500 x 1500 lower (forward) and upper = L^T (backward);
* power-law n = 4e6, Zipf(2) row counts, forward;
each with the lv auto rule.  Level sets, permutation, partition, block nnz
and the selector's inputs and decisions must be bit-exact; x must agree
with the oracle's block-ordered solve (R/recblocking_calculate.h:15-325)
within relative 1e-10 and leave a relative residual below 1e-12, for the
schedule the tuner picks and for the other one.

The served stream: 16 resident factors of the three recipes (n log-spaced in
[1e4, 4e5], ten times smaller than the benchmark's so the oracle keeps up),
Poisson arrivals, Zipf(1.1) factor choice; every answer, FP64 and FP32 with
FP64 refinement, against the oracle's solve of the same right-hand side.
"""
import numpy as np
import pytest

from helpers import rel_err, residual
from rbsptrsv import gen

pytestmark = pytest.mark.gpu

X_TOL = 1e-10
RES_TOL = 1e-12

CONFIGS = [
    ("banded-1e5", lambda: gen.banded(100_000, seed=0)),
    ("laplacian-1500-forward", lambda: gen.laplacian(1500)),
    ("laplacian-1500-backward", lambda: gen.laplacian(1500, lower=False)),
    ("powerlaw-4e6", lambda: gen.powerlaw(4_000_000, seed=0)),
]


@pytest.fixture(scope="module")
def torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a device"
    return torch


@pytest.mark.parametrize("name,make", CONFIGS, ids=[c[0] for c in CONFIGS])
def test_config_full_size_against_oracle(oracle, torch, name, make):
    from rbsptrsv import Plan
    print("full-size", name, flush=True)
    f = make()
    s = f.substitution
    lv = oracle.lv_auto(f.n)
    re, part, bnz, an = oracle.full_pipeline(f.colptr, f.rowidx, f.val, f.n,
                                             s, lv)
    b = gen.rhs(f.n, seed=11)
    xo = oracle.solve_blocks(re, f.n, lv, s, part, b)
    db = torch.as_tensor(b, device="cuda")
    p = Plan.from_factor(f, lv=-1)
    try:
        assert p.lv == lv and p.nlevel == re["nlevel"]
        li, lp = p.levels()
        assert np.array_equal(lp, re["levelPtr"])
        assert np.array_equal(li, re["levelItem"])
        bl = p.blocks()
        for k in ("blk_m", "blk_n", "up", "down", "left", "right"):
            assert np.array_equal(bl[k], part[k]), k
        assert np.array_equal(bl["blk_nnz"], bnz)
        for k in ("method", "nnzr"):
            assert np.array_equal(bl[k], an[k]), k
        tri = np.arange(0, p.nblocks, 2)
        sq = np.arange(1, p.nblocks, 2)
        assert np.array_equal(bl["nlv"][tri], an["nlv"][tri])
        for k in ("m_new", "lenmax", "longrow", "longlen", "empty_ratio"):
            assert np.array_equal(bl[k][sq], an[k][sq]), k
        chosen = p.info().schedule
        x = p.solve(db)
        p.check()
        x = x.cpu().numpy()
        print("  schedule", chosen, "rel", rel_err(x, xo), flush=True)
        assert rel_err(x, xo) < X_TOL
        assert residual(f, x, b) < RES_TOL
    finally:
        p.close()
    # the schedule the tuner did not pick
    other = 1 if chosen == 2 else 2
    q = Plan.from_factor(f, lv=-1, schedule=other)
    try:
        y = q.solve(db)
        q.check()
        y = y.cpu().numpy()
        print("  schedule", other, "rel", rel_err(y, xo), flush=True)
        assert rel_err(y, xo) < X_TOL
        assert residual(f, y, b) < RES_TOL
    finally:
        q.close()


@pytest.mark.parametrize("mode", ["fp64", "f32ir"])
def test_served_stream_against_oracle(oracle, torch, mode):
    from rbsptrsv.serve import Server, poisson_requests
    facs = gen.serve_factors(16, 1e4, 4e5, seed=100)
    sizes = [f.n for f in facs]
    reqs = poisson_requests(16, sizes, 160, rate=4000.0, seed=9)
    print("served stream", mode, sizes, flush=True)
    pipes = {}
    for fi in sorted(set(r[0] for r in reqs)):
        f = facs[fi]
        lv = oracle.lv_auto(f.n)
        re, part, bnz, an = oracle.full_pipeline(f.colptr, f.rowidx, f.val,
                                                 f.n, f.substitution, lv)
        pipes[fi] = (re, part, lv)
    srv = Server(facs, mode=mode, max_batch=32, streams=4)
    try:
        stats, out = srv.run(reqs)
        print("  stats", stats, flush=True)
        assert stats["requests"] == len(reqs)
        assert stats["mean_batch"] > 1.0
        worst = 0.0
        for r in out:
            f = facs[r.factor]
            re, part, lv = pipes[r.factor]
            b = np.asarray(reqs[r.rid][1])
            xo = oracle.solve_blocks(re, f.n, lv, f.substitution, part, b)
            x = r.x.cpu().numpy()
            e = rel_err(x, xo)
            worst = max(worst, e)
            assert e < X_TOL, (r.rid, r.factor, e)
            assert residual(f, x, b) < RES_TOL
        print("  worst rel err", worst, flush=True)
    finally:
        srv.close()
