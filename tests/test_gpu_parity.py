"""GPU parity: the device pipeline against the CPU oracle on the same seeded
inputs.  Integer outputs (level sets, permutation, partition, selector
inputs and decisions, block contents) must be bit-exact; x must agree
within relative 1e-10 (FP64), and the residual ||b - T x|| / ||b|| is
checked against 1e-12."""
import numpy as np
import pytest

from helpers import block_sets, csc_from_deps, rel_err, residual
from rbsptrsv import gen

pytestmark = pytest.mark.gpu

X_TOL = 1e-10      # relative max error of x against the oracle (FP64)
RES_TOL = 1e-12    # relative residual


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a device"
    return torch


def _plan(f, lv, **kw):
    from rbsptrsv import Plan
    return Plan.from_factor(f, lv=lv, **kw)


def _check_pipeline(oracle, f, lv):
    """Levels, permutation, partition, selector and block contents."""
    s = f.substitution
    re, part, bnz, an = oracle.full_pipeline(f.colptr, f.rowidx, f.val, f.n, s, lv)
    p = _plan(f, lv)
    li, lp = p.levels()
    assert p.nlevel == re["nlevel"]
    assert np.array_equal(lp, re["levelPtr"])
    assert np.array_equal(li, re["levelItem"])
    cp, ri, vv = p.permuted_csc()
    assert np.array_equal(cp, re["colptr"])
    assert np.array_equal(ri, re["rowidx"])
    assert np.array_equal(vv, re["val"])
    bl = p.blocks()
    for k in ("blk_m", "blk_n", "up", "down", "left", "right"):
        assert np.array_equal(bl[k], part[k]), k
    assert np.array_equal(bl["blk_nnz"], bnz)
    tri = np.arange(0, p.nblocks, 2)
    sq = np.arange(1, p.nblocks, 2)
    assert np.array_equal(bl["method"], an["method"])
    assert np.array_equal(bl["nlv"][tri], an["nlv"][tri])
    assert np.array_equal(bl["nnzr"], an["nnzr"])
    for k in ("m_new", "lenmax", "longrow", "longlen"):
        assert np.array_equal(bl[k][sq], an[k][sq]), k
    assert np.array_equal(bl["empty_ratio"][sq], an["empty_ratio"][sq])
    for t in range(len(tri)):
        lvp, lvi, nlv = p.tri_levels(t)
        assert np.array_equal(lvp, an["lvptr_local"][t]), t
        assert np.array_equal(lvi, an["lvitem_local"][t]), t
        assert np.array_equal(nlv, an["nnz_lv"][t]), t
    # block contents as sets, against the oracle's permuted matrix
    for i in range(p.nblocks):
        want = block_sets(re["colptr"], re["rowidx"], re["val"], part, i, s)
        rp, ci, v = p.block_csr(i)
        got = {}
        for r in range(len(rp) - 1):
            for q in range(rp[r], rp[r + 1]):
                got[(r, int(ci[q]))] = float(v[q])
        assert got.keys() == want.keys(), i
        for k in want:
            assert got[k] == pytest.approx(want[k], rel=1e-15), (i, k)
    return p, re, part


SMALL = [
    ("random300", lambda: gen.random_lower(300, 3, seed=1, window=40)),
    ("random2000", lambda: gen.random_lower(2000, 6, seed=2)),
    ("laplacian40x30", lambda: gen.laplacian(40, 30)),
    ("banded20000", lambda: gen.banded(20000, seed=5)),
    ("powerlaw30000", lambda: gen.powerlaw(30000, seed=6)),
]


@pytest.mark.parametrize("name,make", SMALL)
@pytest.mark.parametrize("lower", [True, False])
@pytest.mark.parametrize("lv", [1, 2, 4])
def test_pipeline_bit_exact(oracle, torch, name, make, lower, lv):
    f = make()
    if not lower:
        f = f.transpose()
    _say("pipeline", name, "lower" if lower else "upper", "lv", lv)
    _check_pipeline(oracle, f, lv)


@pytest.mark.parametrize("name,make", SMALL)
@pytest.mark.parametrize("lower", [True, False])
@pytest.mark.parametrize("lv", [1, 3, -1])
@pytest.mark.parametrize("schedule", [1, 2])
def test_solve_matches_oracle(oracle, torch, name, make, lower, lv, schedule):
    f = make()
    if not lower:
        f = f.transpose()
    _say("solve", name, "lower" if lower else "upper", "lv", lv,
         "schedule", schedule)
    b = gen.rhs(f.n, seed=3)
    lvv = oracle.lv_auto(f.n) if lv == -1 else lv
    re, part, bnz, an = oracle.full_pipeline(f.colptr, f.rowidx, f.val, f.n,
                                             f.substitution, lvv)
    xo = oracle.solve_blocks(re, f.n, lvv, f.substitution, part, b)
    p = _plan(f, lv, schedule=schedule)
    assert p.info().schedule == schedule
    x = p.solve(torch.as_tensor(b, device="cuda")).cpu().numpy()
    p.check()
    assert rel_err(x, xo) < X_TOL
    assert residual(f, x, b) < RES_TOL


@pytest.mark.parametrize("lower", [True, False])
def test_wide_and_narrow_levels(oracle, torch, lower):
    """400k-row power-law factor: 47 of its 91 levels are wider than one
    workgroup's frontier, so both paths of the level analysis run."""
    f = gen.powerlaw(400_000, seed=3)
    if not lower:
        f = f.transpose()
    _say("wide levels", "lower" if lower else "upper")
    lv = 3
    re, part, bnz, an = oracle.full_pipeline(f.colptr, f.rowidx, f.val, f.n,
                                             f.substitution, lv)
    p = _plan(f, lv)
    li, lp = p.levels()
    assert np.array_equal(lp, re["levelPtr"])
    assert np.array_equal(li, re["levelItem"])
    bl = p.blocks()
    assert np.array_equal(bl["blk_nnz"], bnz)
    for k in ("method", "nlv", "nnzr", "m_new", "lenmax", "longrow", "longlen"):
        assert np.array_equal(bl[k], an[k]), k
    for t in range(1 << lv):
        lvp, lvi, nlv = p.tri_levels(t)
        assert np.array_equal(lvp, an["lvptr_local"][t])
        assert np.array_equal(lvi, an["lvitem_local"][t])
    b = gen.rhs(f.n, seed=5)
    xo = oracle.solve_blocks(re, f.n, lv, f.substitution, part, b)
    for sched in (1, 2):
        q = _plan(f, lv, schedule=sched)
        x = q.solve(torch.as_tensor(b, device="cuda")).cpu().numpy()
        q.check()
        assert rel_err(x, xo) < X_TOL, sched


def test_handworked_fixture_on_device(oracle, torch):
    import json
    import os
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                    "handworked.json")))["bfs_forward"]
    from rbsptrsv import Plan
    val = np.arange(1, len(g["rowidx"]) + 1, dtype=np.float64)
    p = Plan(np.array(g["colptr"]), np.array(g["rowidx"]), val, 0, 1)
    li, lp = p.levels()
    assert li.tolist() == g["levelItem_final"]
    assert lp.tolist() == g["levelPtr"]
    cp, ri, _ = p.permuted_csc()
    assert cp.tolist() == g["colptr_new"] and ri.tolist() == g["rowidx_new"]


@pytest.mark.parametrize("schedule,k", [(0, 3), (1, 3), (2, 3), (2, 40)])
def test_multi_rhs_and_leading_dimension(oracle, torch, schedule, k):
    """Columns at a leading dimension past n, through the tuned, blocked
    and fused schedules.  With at most 4 fused waves, 3 columns are an
    ordinary launch (at most 12 workgroups) and 40 a cooperative one.  Rows
    n..ld-1 of x stay untouched."""
    f = gen.banded(5000, seed=8)
    p = _plan(f, 2, schedule=schedule, fused_waves=4)
    ld = f.n + 17
    B = np.random.default_rng(1).uniform(-1, 1, (k, ld))
    Bd = torch.as_tensor(B, device="cuda").t()          # (ld, k) column-major
    X = torch.zeros((k, ld), dtype=torch.float64, device="cuda").t()
    from rbsptrsv._lib import check, lib
    import ctypes
    check(lib().rb_solve(p._h, ctypes.c_void_p(Bd.data_ptr()), ld,
                         ctypes.c_void_p(X.data_ptr()), ld, k, None))
    torch.cuda.synchronize()
    Xh = X.t().cpu().numpy()
    re, part, bnz, an = oracle.full_pipeline(f.colptr, f.rowidx, f.val, f.n, 0, 2)
    for j in range(k):
        xo = oracle.solve_blocks(re, f.n, 2, 0, part, B[j, :f.n].copy())
        assert rel_err(Xh[j, :f.n], xo) < X_TOL
        assert np.all(Xh[j, f.n:] == 0.0)   # padding rows untouched


@pytest.mark.parametrize("kernels", [1, 2, 3])
@pytest.mark.parametrize("name,make", [
    ("powerlaw60000", lambda: gen.powerlaw(60000, seed=3)),
    ("banded20000-upper", lambda: gen.banded(20000, seed=6).transpose()),
])
def test_blocked_columns_match_single(torch, kernels, name, make):
    """The blocked walk solves k columns at once on a grid of the next
    power of two (graphs kept per width, the column count read from device
    memory): every column equals the single-column solve bit for bit, for
    widths inside and across the buckets, in rising and falling order, with
    each triangular kernel (reference selector's, sync-free with its
    interleaved columns, one-workgroup)."""
    f = make()
    _say("blocked columns", name, kernels)
    p = _plan(f, 3, schedule=1, block_kernels=kernels)
    rng = np.random.default_rng(4)
    B = rng.uniform(-1, 1, (33, f.n))
    single = [p.solve(torch.as_tensor(B[j], device="cuda")).cpu().numpy()
              for j in range(33)]
    for k in (3, 17, 33, 5, 2, 32, 1):
        X = p.solve(torch.as_tensor(B[:k], device="cuda").t()).t().cpu().numpy()
        for j in range(k):
            assert np.array_equal(X[j].view(np.int64), single[j].view(np.int64)), (k, j)
    p.check()
    assert residual(f, single[0], B[0]) < RES_TOL
    p.close()


@pytest.mark.parametrize("schedule", [0, 1, 2])
def test_plan_build_phases(torch, schedule):
    """The plan reports its build phases: non-negative, the first three
    adding up to preprocess_ms; no timed tuning when the fused schedule is
    asked for or the blocked schedule keeps the reference selector's
    kernels, and no fused build when the blocked schedule is asked for
    (those phases are then two clock reads apart)."""
    f = gen.laplacian(60)
    inf = _plan(f, 2, schedule=schedule, block_kernels=1).info()
    ph = list(inf.phase_ms)
    assert all(t >= 0.0 for t in ph)
    assert abs(sum(ph[:3]) - inf.preprocess_ms) <= 1e-9 * max(1.0, inf.preprocess_ms)
    assert ph[0] > 0.0
    if schedule != 0:
        assert ph[3] < 0.05
    if schedule == 1:
        assert ph[2] < 0.05


def test_deterministic_bitwise(torch):
    f = gen.powerlaw(50000, seed=2)
    p = _plan(f, 3)
    b = torch.as_tensor(gen.rhs(f.n, 4), device="cuda")
    x1 = p.solve(b).cpu().numpy()
    x2 = p.solve(b).cpu().numpy()
    assert np.array_equal(x1.view(np.int64), x2.view(np.int64))


# ------------------------------------------------------------- edge cases
@pytest.mark.parametrize("n,lv", [(1, 1), (3, 2), (5, 3), (7, 1), (100, 6)])
def test_tiny_and_more_blocks_than_rows(oracle, torch, n, lv):
    f = gen.random_lower(n, 2, seed=n)
    _say("tiny", n, lv)
    p, re, part = _check_pipeline(oracle, f, lv)
    b = gen.rhs(n, 1)
    x = p.solve(torch.as_tensor(b, device="cuda")).cpu().numpy()
    assert residual(f, x, b) < RES_TOL


def test_diagonal_only(oracle, torch):
    n = 1000
    f = csc_from_deps(n, {}, diag=2.0)
    p, re, part = _check_pipeline(oracle, f, 3)
    b = gen.rhs(n, 2)
    x = p.solve(torch.as_tensor(b, device="cuda")).cpu().numpy()
    assert np.array_equal(x, b * 0.5)


def test_long_rows_and_empty_square_blocks(oracle, torch):
    n = 8000
    deps = {7000: list(range(0, 3000)), 7001: list(range(100, 2600)),
            7002: [5]}
    f = csc_from_deps(n, deps, diag=float(3100))
    p, re, part = _check_pipeline(oracle, f, 2)
    bl = p.blocks()
    assert int(bl["longrow"].sum()) == 2
    b = gen.rhs(n, 3)
    x = p.solve(torch.as_tensor(b, device="cuda")).cpu().numpy()
    assert residual(f, x, b) < RES_TOL


def _expect_status(torch, colptr, rowidx, val, subst, code):
    from rbsptrsv import Plan, RbError
    with pytest.raises(RbError) as ei:
        Plan(np.asarray(colptr), np.asarray(rowidx), np.asarray(val, float),
             subst, 1)
    assert ei.value.status == code


def test_error_codes_not_hangs(torch):
    # missing diagonal (the reference's level analysis never terminates)
    _expect_status(torch, [0, 1, 1, 2], [0, 2], [1.0, 1.0], 0, 4)
    # zero diagonal
    _expect_status(torch, [0, 2, 3], [0, 1, 1], [0.0, 1.0, 1.0], 0, 5)
    # entry above the diagonal in a lower solve
    _expect_status(torch, [0, 1, 3], [0, 0, 1], [1.0, 1.0, 1.0], 0, 3)
    # unsorted column
    _expect_status(torch, [0, 3, 4, 5], [0, 2, 1, 1, 2], [1.0] * 5, 0, 3)
    # diagonal not last in a backward solve column
    _expect_status(torch, [0, 1, 3], [0, 1, 0], [1.0] * 3, 1, 3)


# ---------------------------------------- general inputs (canonicalize)
def _scramble(f, seed, dup_frac=0.3):
    """The same matrix with row ids shuffled inside every column and a
    fraction of the entries split into two or three repeated entries whose
    values add up to the original (exact halves / thirds are not needed:
    the split parts are a and v - a)."""
    rng = np.random.default_rng(seed)
    cp, ri, v = [0], [], []
    for j in range(f.n):
        rows = list(f.rowidx[f.colptr[j]:f.colptr[j + 1]])
        vals = list(f.val[f.colptr[j]:f.colptr[j + 1]])
        er, ev = [], []
        for r, x in zip(rows, vals):
            k = 1 + (rng.random() < dup_frac) + (rng.random() < dup_frac / 3)
            parts = [float(rng.uniform(-1, 1)) for _ in range(k - 1)]
            parts.append(x - sum(parts))
            er += [r] * k
            ev += parts
        perm = rng.permutation(len(er))
        ri += [er[q] for q in perm]
        v += [ev[q] for q in perm]
        cp.append(len(ri))
    return (np.array(cp, np.int32), np.array(ri, np.int32), np.array(v))


@pytest.mark.parametrize("name,make,lower", [
    ("random3000", lambda: gen.random_lower(3000, 4, seed=21), True),
    ("laplacian40", lambda: gen.laplacian(40), True),
    ("banded5000-upper", lambda: gen.banded(5000, seed=2).transpose(), False),
])
@pytest.mark.parametrize("schedule", [1, 2])
def test_canonicalize_general_input(oracle, torch, name, make, lower, schedule):
    """Unsorted row ids and repeated entries are rejected by default and,
    with canonicalize, solved as the merged matrix: x within 1e-10 of the
    oracle's solve of the original factor, plan nnz = the original nnz."""
    from rbsptrsv import Plan, RbError
    f = make()
    subst = 0 if lower else 1
    cp, ri, v = _scramble(f, 5)
    print("canonicalize", name, "schedule", schedule, "entries", len(ri),
          "->", f.nnz, flush=True)
    with pytest.raises(RbError) as ei:
        Plan(cp, ri, v, subst, -1, schedule=schedule)
    assert ei.value.status == 3
    p = Plan(cp, ri, v, subst, -1, schedule=schedule, canonicalize=True)
    assert p.nnz == f.nnz
    b = gen.rhs(f.n, seed=8)
    x = p.solve(torch.as_tensor(b, device="cuda")).cpu().numpy()
    p.check()
    lvv = oracle.lv_auto(f.n)
    re, part, bnz, an = oracle.full_pipeline(f.colptr, f.rowidx, f.val, f.n,
                                             subst, lvv)
    xo = oracle.solve_blocks(re, f.n, lvv, subst, part, b)
    assert rel_err(x, xo) < X_TOL
    # the merged matrix is the original one: same permuted CSC as a plan of
    # the original factor
    q = Plan.from_factor(f, -1, schedule=schedule)
    assert np.array_equal(p.permuted_csc()[0], q.permuted_csc()[0])
    assert np.array_equal(p.permuted_csc()[1], q.permuted_csc()[1])
    assert np.allclose(p.permuted_csc()[2], q.permuted_csc()[2], rtol=1e-14,
                       atol=1e-15)


def test_canonicalize_still_checks_structure(torch):
    """Canonicalizing does not hide real errors: a row id out of range, an
    entry outside the triangle, a missing diagonal."""
    from rbsptrsv import Plan, RbError
    cases = [([0, 2, 3], [0, 5, 1], [1.0, 1.0, 1.0], 0, 3),
             ([0, 1, 3], [0, 1, 0], [1.0, 1.0, 1.0], 0, 3),
             ([0, 1, 1, 2], [0, 2], [1.0, 1.0], 0, 4),
             ([0, 2, 3], [0, 0, 1], [1.0, -1.0, 1.0], 0, 5)]  # 1 + -1 = 0
    for cp, ri, v, subst, code in cases:
        with pytest.raises(RbError) as ei:
            Plan(np.asarray(cp), np.asarray(ri), np.asarray(v), subst, 1,
                 canonicalize=True)
        assert ei.value.status == code, (cp, ri, v)


# ------------------------------------------------ legacy entry points
def test_legacy_recblocking_solver_cuda(oracle, torch):
    from rbsptrsv import recblocking_solver_cuda, set_bench_repeat
    f = gen.banded(30000, seed=4)
    b = f.matvec(np.ones(f.n))
    d = lambda a, t: torch.as_tensor(np.asarray(a, t), device="cuda")
    dcp, dri, dv = d(f.colptr, np.int32), d(f.rowidx, np.int32), d(f.val, np.float64)
    db = d(b, np.float64)
    dx = torch.zeros(f.n, dtype=torch.float64, device="cuda")
    old = set_bench_repeat(3)
    try:
        ct, pt = recblocking_solver_cuda(dcp, dri, dv, f.n, f.n, f.nnz, dx, db,
                                         0, -1)
    finally:
        set_bench_repeat(old)
    assert ct > 0 and pt > 0
    x = dx.cpu().numpy()
    assert np.abs(x - 1.0).max() < 1e-12          # b = L * 1  =>  x = 1
    assert np.array_equal(db.cpu().numpy(), b)     # b left unchanged


def test_legacy_recblocking_solver_host_multi_rhs(oracle, torch):
    from rbsptrsv import recblocking_solver, set_bench_repeat
    f = gen.laplacian(50, 40).transpose()          # backward solve
    rhs = 2
    B = np.random.default_rng(0).uniform(-1, 1, (rhs, f.n))
    X = np.zeros((rhs, f.n))
    old = set_bench_repeat(2)
    try:
        recblocking_solver(f.colptr, f.rowidx, f.val, f.n, f.n, f.nnz,
                           B.ravel(), X.reshape(-1), None, rhs, 2, 1)
    finally:
        set_bench_repeat(old)
    for j in range(rhs):
        assert residual(f, X[j], B[j]) < RES_TOL


@pytest.mark.parametrize("kernels", [1, 3])
def test_blocked_graph_replays_match_eager(oracle, torch, kernels):
    """The blocked schedule replays a captured graph of its block walk; every
    replay (with eager work in between) must equal the kernel-by-kernel
    launch sequence bitwise.  Both plans take their block kernels by a fixed
    rule (the reference selector's, or the one-workgroup kernel wherever it
    applies): kernels chosen by timing may differ between two plans, and
    with them the order of the sums."""
    f = gen.powerlaw(60000, seed=9)
    p = _plan(f, 3, schedule=1, block_kernels=kernels)
    q = _plan(f, 3, schedule=1, graph=False, block_kernels=kernels)
    bs = [gen.rhs(f.n, seed=s) for s in (11, 12, 13)]
    eager = [q.solve(torch.as_tensor(b, device="cuda")).cpu().numpy()
             for b in bs]
    q.check()
    re, part, bnz, an = oracle.full_pipeline(f.colptr, f.rowidx, f.val, f.n, 0, 3)
    for rep in range(3):
        for i, b in enumerate(bs):
            x = p.solve(torch.as_tensor(b, device="cuda")).cpu().numpy()
            torch.cuda.synchronize()
            _ = torch.ones(1000, device="cuda").sum().item()  # eager work
            assert np.array_equal(x.view(np.int64), eager[i].view(np.int64))
            if rep == 0:
                xo = oracle.solve_blocks(re, f.n, 3, 0, part, b)
                assert rel_err(x, xo) < X_TOL
    p.check()


@pytest.mark.parametrize("schedule", [1, 2])
def test_kernel_timing_events(torch, schedule):
    f = gen.laplacian(60, 50)
    p = _plan(f, 2, schedule=schedule)
    b = torch.as_tensor(gen.rhs(f.n, 2), device="cuda")
    from rbsptrsv import RbError
    with pytest.raises(RbError):
        p.kernel_ms()                       # never enabled
    p.set_timing(True)
    x = p.solve(b)
    ms = p.kernel_ms()
    assert 0.0 < ms < 1000.0
    p.set_timing(False)
    x2 = p.solve(b)
    assert np.array_equal(x.cpu().numpy(), x2.cpu().numpy())


def _solve_many(torch, p, B, ld=None):
    """Solve the columns of B (k, n) in one call; returns X (k, n)."""
    import ctypes
    from rbsptrsv._lib import check, lib
    k, n = B.shape
    ld = ld or n
    Bp = np.zeros((k, ld))
    Bp[:, :n] = B
    Bd = torch.as_tensor(Bp, device="cuda")
    Xd = torch.zeros((k, ld), dtype=torch.float64, device="cuda")
    check(lib().rb_solve(p._h, ctypes.c_void_p(Bd.data_ptr()), ld,
                         ctypes.c_void_p(Xd.data_ptr()), ld, k, None))
    torch.cuda.synchronize()
    return Xd.cpu().numpy()[:, :n]


@pytest.mark.parametrize("name,make", [
    ("laplacian60x50", lambda: gen.laplacian(60, 50)),
    ("random2000", lambda: gen.random_lower(2000, 6, seed=2)),
    ("banded20000-upper", lambda: gen.banded(20000, seed=5).transpose()),
])
def test_fused_many_rhs_match_single(oracle, torch, name, make):
    """Several right-hand sides per fused launch (waves replicated per
    column) give bitwise the single-column results; counts that need more
    than one launch, and changing counts between calls, included."""
    f = make()
    p = _plan(f, 2, schedule=2)
    rng = np.random.default_rng(7)
    re, part, bnz, an = oracle.full_pipeline(f.colptr, f.rowidx, f.val, f.n,
                                             f.substitution, 2)
    for k in (3, 1, 70, 2, 131, 5):
        _say("fused rhs", name, k)
        B = rng.uniform(-1, 1, (k, f.n))
        X = _solve_many(torch, p, B, ld=f.n + 3)
        p.check()
        for j in sorted({0, k // 2, k - 1}):
            xs = p.solve(torch.as_tensor(B[j], device="cuda")).cpu().numpy()
            assert np.array_equal(X[j].view(np.int64), xs.view(np.int64)), (k, j)
            xo = oracle.solve_blocks(re, f.n, 2, f.substitution, part, B[j])
            assert rel_err(X[j], xo) < X_TOL
        assert max(residual(f, X[j], B[j]) for j in range(k)) < RES_TOL


# ----------------------------------------------- FP32 + FP64 refinement
@pytest.mark.parametrize("name,make", [
    ("random2000", lambda: gen.random_lower(2000, 6, seed=2)),
    ("laplacian40x30-upper", lambda: gen.laplacian(40, 30, lower=False)),
    ("banded20000", lambda: gen.banded(20000, seed=5)),
    ("powerlaw30000", lambda: gen.powerlaw(30000, seed=6)),
])
@pytest.mark.parametrize("schedule", [0, 1, 2])
def test_f32_refinement_reaches_fp64_accuracy(oracle, torch, name, make, schedule):
    """schedule 2: the FP32 corrections run on the fused dataflow kernel
    (FP32 copies of its records); 1: on the blocked walk."""
    f = make()
    _say("f32ir", name, schedule)
    p = _plan(f, 2, mixed=True, schedule=schedule)
    b = gen.rhs(f.n, seed=4)
    x, it, rel = p.solve_f32ir(torch.as_tensor(b, device="cuda"), tol=1e-12)
    x = x.cpu().numpy()
    p.check()
    assert rel <= 1e-12 and 1 <= it <= 6, (it, rel)
    # the reported max-norm residual is the true one, up to the rounding of
    # the residual itself (another summation order on the host): a few ulps
    # of max_i (|A| |x|)_i
    import dataclasses
    r = b - f.matvec(x)
    absax = dataclasses.replace(f, val=np.abs(f.val)).matvec(np.abs(x))
    slack = 64 * np.finfo(float).eps * absax.max() / np.abs(b).max()
    assert abs(np.abs(r).max() / np.abs(b).max() - rel) <= slack
    assert residual(f, x, b) < RES_TOL
    re, part, bnz, an = oracle.full_pipeline(f.colptr, f.rowidx, f.val, f.n,
                                             f.substitution, 2)
    xo = oracle.solve_blocks(re, f.n, 2, f.substitution, part, b)
    assert rel_err(x, xo) < X_TOL
    # one FP32 correction alone is FP32-accurate only
    from rbsptrsv import RbError
    with pytest.raises(RbError) as ei:
        p.solve_f32ir(torch.as_tensor(b, device="cuda"), tol=1e-14, max_iter=1)
    assert ei.value.status == 9                      # not converged
    # without a report: the same x, no synchronisation, and a miss shows in
    # the plan's check (then cleared)
    xa, it_a, rel_a = p.solve_f32ir(torch.as_tensor(b, device="cuda"),
                                    tol=1e-12, report=False)
    assert it_a is None and rel_a is None
    p.check()
    assert np.array_equal(xa.cpu().numpy(), x)
    p.solve_f32ir(torch.as_tensor(b, device="cuda"), tol=1e-14, max_iter=1,
                  report=False)
    with pytest.raises(RbError) as ei:
        p.check()
    assert ei.value.status == 9
    p.check()                                        # cleared


@pytest.mark.parametrize("schedule,k", [(0, 3), (1, 3), (2, 3), (2, 40)])
def test_f32_refinement_many_rhs_and_errors(torch, schedule, k):
    f = gen.laplacian(50, 40)
    _say("f32ir many", schedule, k)
    p = _plan(f, 2, mixed=True, schedule=schedule)
    B = np.random.default_rng(3).uniform(-1, 1, (k, f.n))
    B[1] = 0.0                                       # zero column: x = 0
    Bd = torch.as_tensor(B, device="cuda").t()
    X, it, rel = p.solve_f32ir(Bd)
    X = X.t().cpu().numpy()
    assert rel <= 1e-12
    assert np.all(X[1] == 0.0)
    for j in [0] + list(range(2, k)):
        assert residual(f, X[j], B[j]) < RES_TOL
    # the same columns again: the FP32 slot buffers start over each solve
    X2 = p.solve_f32ir(Bd)[0].t().cpu().numpy()
    assert np.array_equal(X2, X)
    q = _plan(f, 2)                                  # no FP32 copies
    from rbsptrsv import RbError
    with pytest.raises(RbError) as ei:
        q.solve_f32ir(torch.as_tensor(B[0], device="cuda"))
    assert ei.value.status == 1


# ------------------------------------------------------------ batches
def _small_batch():
    sizes = [300, 1200, 50, 2500, 800, 1]
    facs = [gen.banded(n, seed=10 + j, window=min(200, max(n - 1, 1)))
            if n > 1 else csc_from_deps(1, {}, diag=3.0)
            for j, n in enumerate(sizes)]
    facs[2] = gen.random_lower(50, 4, seed=3)
    return facs, [2, 0, 1, 5, 3, 1]


@pytest.mark.parametrize("schedule", [0, 1, 2])
def test_batch_matches_single_factor_solves(oracle, torch, schedule):
    from rbsptrsv import Batch
    facs, nrhs = _small_batch()
    _say("batch", schedule)
    bt = Batch(facs, lv=2, schedule=schedule)
    rng = np.random.default_rng(11)
    Bs = [rng.uniform(-1, 1, (k, f.n)) for f, k in zip(facs, nrhs)]
    bs = [None if k == 0 else
          (torch.as_tensor(B[0], device="cuda") if k == 1 else
           torch.as_tensor(B, device="cuda").t()) for B, k in zip(Bs, nrhs)]
    xs = bt.solve(bs)
    torch.cuda.synchronize()
    for f, B, k, x in zip(facs, Bs, nrhs, xs):
        if k == 0:
            assert x is None
            continue
        X = x.cpu().numpy()[None, :] if k == 1 else x.t().cpu().numpy()
        lv = 2 if f.n >= 4 else 0
        re, part, bnz, an = oracle.full_pipeline(f.colptr, f.rowidx, f.val,
                                                 f.n, 0, lv)
        for j in range(k):
            xo = oracle.solve_blocks(re, f.n, lv, 0, part, B[j])
            assert rel_err(X[j], xo) < X_TOL
            assert residual(f, X[j], B[j]) < RES_TOL
    # the same solve again is bitwise the same
    xs2 = bt.solve(bs)
    for x, x2 in zip(xs, xs2):
        if x is not None:
            assert torch.equal(x, x2)


@pytest.mark.parametrize("schedule", [1, 2])
def test_batch_plan_solved_directly_after_batches(torch, schedule):
    """A batch's blocked walk skips, in each column, the rows of factors
    with fewer columns; a solve called on the union plan itself (not
    through the batch) afterwards solves every row in every column."""
    import ctypes
    from rbsptrsv import Batch
    from rbsptrsv._lib import check, lib
    facs, nrhs = _small_batch()
    bt = Batch(facs, lv=2, schedule=schedule)
    rng = np.random.default_rng(13)
    bs = [None if k == 0 else torch.as_tensor(rng.uniform(-1, 1, (k, f.n)),
                                              device="cuda").t()
          for f, k in zip(facs, nrhs)]
    bt.solve(bs)
    n = sum(f.n for f in facs)
    k = max(nrhs)
    Bu = rng.uniform(-1, 1, (k, n))
    Bd = torch.as_tensor(Bu, device="cuda")
    Xd = torch.zeros_like(Bd)
    check(lib().rb_solve(lib().rb_batch_plan(bt._h),
                         ctypes.c_void_p(Bd.data_ptr()), n,
                         ctypes.c_void_p(Xd.data_ptr()), n, k, None))
    X = Xd.cpu().numpy()
    off = 0
    for f in facs:
        for j in range(k):
            assert residual(f, X[j, off:off + f.n], Bu[j, off:off + f.n]) < RES_TOL
        off += f.n
    bt.close()


def test_batch_f32_refinement(torch):
    from rbsptrsv import Batch
    facs, nrhs = _small_batch()
    bt = Batch(facs, lv=2, mixed=True)
    rng = np.random.default_rng(12)
    bs = [None if k == 0 else torch.as_tensor(rng.uniform(-1, 1, (k, f.n)),
                                              device="cuda").t()
          for f, k in zip(facs, nrhs)]
    xs, it, rel = bt.solve_f32ir(bs, tol=1e-12)
    assert rel <= 1e-12 and it >= 1
    for f, b, x in zip(facs, bs, xs):
        if b is None:
            continue
        B, X = b.t().cpu().numpy(), x.t().cpu().numpy()
        for j in range(B.shape[0]):
            assert residual(f, X[j], B[j]) < RES_TOL


# ------------------------------------------------------------ serving
@pytest.mark.parametrize("mode", ["fp64", "f32ir"])
def test_serving_results_same_alone_and_in_the_mix(torch, mode):
    """Requests served with continuous cross-request batching give the
    result of the same request solved alone: bitwise for fp64, and within
    the refinement tolerance for f32ir; a memory budget of one request at a
    time still completes everything."""
    from rbsptrsv.serve import Server, poisson_requests
    facs = [gen.banded(3000, seed=1, window=300), gen.laplacian(40, 40),
            gen.random_lower(2500, 5, seed=4)]
    sizes = [f.n for f in facs]
    reqs = poisson_requests(len(facs), sizes, 60, rate=2e4, seed=2)
    srv = Server(facs, mode=mode, max_batch=16, streams=2)
    stats, out = srv.run(reqs)
    _say("serve", mode, stats)
    assert stats["requests"] == 60 and stats["p99_ms"] >= stats["p50_ms"] > 0
    assert stats["mean_batch"] > 1.0          # requests were batched
    for r in out:
        f = facs[r.factor]
        x = r.x.cpu().numpy()
        assert residual(f, x, r.b.cpu().numpy()) < RES_TOL
        alone = srv.plans[r.factor]
        bd = torch.as_tensor(np.asarray(reqs[r.rid][1]), device="cuda")
        if mode == "fp64":
            xa = alone.solve(bd).cpu().numpy()
            assert np.array_equal(x.view(np.int64), xa.view(np.int64))
        else:
            xa, it, rel = alone.solve_f32ir(bd)
            assert rel_err(x, xa.cpu().numpy()) < 1e-10
    tight = Server(facs, mode="fp64", max_batch=16, streams=2,
                   mem_budget=16 * 3000)
    stats2, out2 = tight.run(reqs[:20])
    assert stats2["requests"] == 20
    for r in out2:
        assert residual(facs[r.factor], r.x.cpu().numpy(),
                        r.b.cpu().numpy()) < RES_TOL


@pytest.mark.parametrize("schedule", [1, 2])
def test_watchdog_trip_gives_all_nan_and_check_raises(torch, schedule):
    """Force the failure path from the host: a plan whose bounded waits give
    up after one poll.  Every kernel still exits normally; the solve must
    then hand out x with every entry NaN (never a finite wrong answer),
    rb_plan_check must report RB_ERR_INTERNAL and clear the flag, and a
    healthy plan on the same factor is unaffected."""
    from rbsptrsv import Plan
    from rbsptrsv._lib import RbError
    if schedule == 1:   # lane k of a wave waits on lane k-1: always waits
        f = csc_from_deps(3000, {i: [i - 1] for i in range(1, 3000)})
    else:               # cross-wave dependencies on every level
        f = gen.banded(20000, seed=4)
    print("watchdog trip, schedule", schedule, flush=True)
    b = torch.as_tensor(gen.rhs(f.n, seed=2), device="cuda")
    p = Plan.from_factor(f, lv=1, schedule=schedule, watchdog_spins=1)
    try:
        x = p.solve(b)
        with pytest.raises(RbError) as e:
            p.check()
        assert e.value.status == 8
        assert bool(torch.isnan(x).all())
        p.check()   # the flag was cleared by the report
    finally:
        p.close()
    q = Plan.from_factor(f, lv=1, schedule=schedule)
    try:
        y = q.solve(b)
        q.check()
        assert residual(f, y.cpu().numpy(), b.cpu().numpy()) < 1e-12
    finally:
        q.close()


def test_watchdog_trip_in_fp32_refinement_on_fused_schedule(torch):
    """The FP32 refinement on the fused schedule under a forced trip (bounded
    waits that give up after one poll; every kernel still exits normally):
    the call reports a miss, x is all NaN, the plan's check reports the
    watchdog, and the next call on the same plan, whose slot buffers the
    epilogues re-armed, trips the same way instead of reading stale
    slots."""
    from rbsptrsv import Plan
    from rbsptrsv._lib import RbError
    f = gen.banded(20000, seed=4)
    print("watchdog trip, fp32 refinement, fused", flush=True)
    b = torch.as_tensor(gen.rhs(f.n, seed=2), device="cuda")
    p = Plan.from_factor(f, lv=1, schedule=2, mixed=True, watchdog_spins=1)
    try:
        for _ in range(2):
            with pytest.raises(RbError) as e:
                p.solve_f32ir(b, tol=1e-12, max_iter=3)
            assert e.value.status == 9                 # not converged
            with pytest.raises(RbError) as e:
                p.check()
            assert e.value.status == 8                 # the watchdog
            x = p.solve_f32ir(b, tol=1e-12, max_iter=1, report=False)[0]
            with pytest.raises(RbError):
                p.check()
            assert bool(torch.isnan(x).all())
    finally:
        p.close()
