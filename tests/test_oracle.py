"""Pins the CPU oracle (oracle/rb_oracle.c) before it is trusted as the
checker: hand-worked fixtures, closed forms, invariants and independent
library routines.  Each test names the reference behaviour it pins and the
mistake it would catch."""
import json
import os

import numpy as np
import pytest
import scipy.linalg
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from helpers import check_bfs_order, csc_from_deps, rel_err
from rbsptrsv import gen

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "handworked.json")))


def _rowptr(colptr, rowidx, m):
    rp = np.zeros(m + 1, np.int32)
    np.cumsum(np.bincount(np.asarray(rowidx), minlength=m), out=rp[1:])
    return rp


# ------------------------------------------------------------------ A2/A3
def test_findlevel_handworked_forward(oracle):
    g = GOLD["bfs_forward"]
    n = g["n"]
    rp = _rowptr(g["colptr"], g["rowidx"], n)
    assert np.diff(rp).tolist() == g["rowcnt"]
    rc, nl, lp, li = oracle.findlevel(g["colptr"], g["rowidx"], rp, n)
    assert rc == 0
    assert nl == g["nlevel"]
    assert lp.tolist() == g["levelPtr"]
    # discovery order [3, 2] inside level 1 -- a row-id sort would fail here
    assert li.tolist() == g["levelItem_bfs"]


def test_reorder_handworked_forward(oracle):
    g = GOLD["bfs_forward"]
    n = g["n"]
    val = np.arange(1, len(g["rowidx"]) + 1, dtype=np.float64)
    re = oracle.reorder(g["colptr"], g["rowidx"], val, n, 0)
    assert re["colptr"].tolist() == g["colptr_new"]
    assert re["rowidx"].tolist() == g["rowidx_new"]
    assert re["levelItem"].tolist() == g["levelItem_final"]
    # values travel with their entries: new column 2 is old column 3
    assert re["val"][5:8].tolist() == val[7:10].tolist()


def test_reorder_handworked_backward(oracle):
    g = GOLD["bfs_backward"]
    n = g["n"]
    val = np.arange(1, len(g["rowidx"]) + 1, dtype=np.float64)
    rp = _rowptr(g["colptr"], g["rowidx"], n)
    rc, nl, lp, li = oracle.findlevel(g["colptr"], g["rowidx"], rp, n)
    assert li.tolist() == g["levelItem_bfs"] and lp.tolist() == g["levelPtr"]
    re = oracle.reorder(g["colptr"], g["rowidx"], val, n, 1)
    assert re["colptr"].tolist() == g["colptr_new"]
    assert re["rowidx"].tolist() == g["rowidx_new"]
    assert re["levelItem"].tolist() == g["levelItem_final"]


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("lower", [True, False])
def test_findlevel_order_invariants(oracle, seed, lower):
    f = gen.random_lower(300, 3, seed=seed, window=40)
    if not lower:
        f = f.transpose()
    rp = _rowptr(f.colptr, f.rowidx, f.n)
    rc, nl, lp, li = oracle.findlevel(f.colptr, f.rowidx, rp, f.n)
    assert rc == 0
    check_bfs_order(f.colptr, f.rowidx, f.n, li, lp, lower=lower)


def test_findlevel_stall_reports_instead_of_hanging(oracle):
    # row 1 has no entry at all (its diagonal is missing): its in-degree
    # is 0, never 1, so the reference's while loop never ends
    colptr = np.array([0, 1, 1, 2], np.int32)
    rowidx = np.array([0, 2], np.int32)
    rp = _rowptr(colptr, rowidx, 3)
    rc, *_ = oracle.findlevel(colptr, rowidx, rp, 3)
    assert rc == -1


@pytest.mark.parametrize("subst", [0, 1])
def test_reorder_is_symmetric_permutation(oracle, subst):
    f = gen.random_lower(200, 4, seed=3)
    if subst:
        f = f.transpose()
    re = oracle.reorder(f.colptr, f.rowidx, f.val, f.n, subst)
    A = sp.csc_matrix((f.val, f.rowidx, f.colptr), shape=(f.n, f.n)).toarray()
    P = sp.csc_matrix((re["val"], re["rowidx"], re["colptr"]),
                      shape=(f.n, f.n)).toarray()
    li = re["levelItem"]
    # P[i, j] = A[li[i], li[j]]
    assert np.array_equal(P, A[np.ix_(li, li)])
    assert np.array_equal(P, np.tril(P) if subst == 0 else np.triu(P))


# ------------------------------------------------------------------- A1/A4
def test_lv_auto_closed_form(oracle):
    for m, lv in GOLD["lv_auto"]:
        assert oracle.lv_auto(m) == lv


@pytest.mark.parametrize("key", ["partition_forward_m10_lv2",
                                 "partition_backward_m10_lv2"])
def test_partition_handworked(oracle, key):
    g = GOLD[key]
    part = oracle.partition(g["m"], g["lv"], g["subst"])
    for k in ("up", "down", "left", "right", "blk_m", "blk_n"):
        assert part[k].tolist() == g[k], k


def _closed_form(m, lv, subst):
    T = 1 << lv
    step = m // T
    out = []
    for i in range(2 * T - 1):
        if i % 2 == 0:
            t = i // 2
            lo, hi = t * step, (m if t == T - 1 else (t + 1) * step)
            if subst:
                lo, hi = m - hi, m - lo
            out.append((lo, hi, lo, hi))
        else:
            t = (i + 1) // 2
            lb = t & -t
            r0, r1 = t * step, (m if t + lb == T else (t + lb) * step)
            c0, c1 = (t - lb) * step, t * step
            if subst:
                r0, r1, c0, c1 = m - r1, m - r0, m - c1, m - c0
            out.append((r0, r1, c0, c1))
    return out


@pytest.mark.parametrize("m", [1, 3, 7, 10, 37, 64, 101, 1000])
@pytest.mark.parametrize("lv", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("subst", [0, 1])
def test_partition_closed_form_and_tiling(oracle, m, lv, subst):
    part = oracle.partition(m, lv, subst)
    cf = _closed_form(m, lv, subst)
    for i, (u, d, l, r) in enumerate(cf):
        assert (part["up"][i], part["down"][i], part["left"][i],
                part["right"][i]) == (u, d, l, r), i
    if m > 200:
        return
    # every position of the triangle lies in exactly one block
    cover = np.zeros((m, m), np.int32)
    for i in range(len(cf)):
        u, d, l, r = cf[i]
        for row in range(u, d):
            for col in range(l, r):
                if i % 2 == 1 or (row >= col if subst == 0 else row <= col):
                    cover[row, col] += 1
    tri = np.tril(np.ones((m, m), np.int32)) if subst == 0 else \
        np.triu(np.ones((m, m), np.int32))
    assert np.array_equal(cover, tri)


@pytest.mark.parametrize("subst", [0, 1])
def test_block_nnz_partitions_all_entries(oracle, subst):
    f = gen.random_lower(500, 5, seed=11)
    if subst:
        f = f.transpose()
    for lv in (1, 2, 3, 5):
        re, part, bnz, an = oracle.full_pipeline(f.colptr, f.rowidx, f.val,
                                                 f.n, subst, lv)
        assert int(bnz.sum()) == f.nnz
        # triangular blocks hold exactly their own diagonals
        assert all(bnz[2 * t] >= part["blk_m"][2 * t] for t in range(1 << lv))


# ---------------------------------------------------------------------- A6
def test_selector_diagonal_only_blocks(oracle):
    n = 64
    f = csc_from_deps(n, {i: [i - 1] for i in range(33, 64)})
    re, part, bnz, an = oracle.full_pipeline(f.colptr, f.rowidx, f.val, n, 0, 1)
    # rows 0..32 have no dependency: level 0 holds them, tri0 = diagonal only
    assert an["method"][0] == 0 and an["nlv"][0] == 1
    assert bnz[0] == part["blk_m"][0]


def test_selector_square_statistics(oracle):
    # rows 64..127 each depend on 20 columns of [0, 64): nnzr 20 > 12,
    # no empty row -> CSR warp (method 2); add a long row > 2048 via a
    # bigger matrix below
    n = 128
    deps = {i: list(range(0, 20)) for i in range(64, 128)}
    f = csc_from_deps(n, deps)
    re, part, bnz, an = oracle.full_pipeline(f.colptr, f.rowidx, f.val, n, 0, 1)
    assert an["m_new"][1] == part["blk_m"][1]
    # the square block is whatever lies in rows [step,m) x cols [0,step)
    assert an["method"][1] in (0, 2)
    # empty square block: no crash, method -1
    g = csc_from_deps(n, {})
    re, part, bnz, an = oracle.full_pipeline(g.colptr, g.rowidx, g.val, n, 0, 2)
    assert all(an["method"][i] == -1 for i in (1, 3, 5))
    assert all(an["nnzr"][i] == -1 for i in (1, 3, 5))


def test_selector_long_rows(oracle):
    n = 6000
    deps = {5000: list(range(0, 2500)), 5001: list(range(0, 100))}
    f = csc_from_deps(n, deps)
    re, part, bnz, an = oracle.full_pipeline(f.colptr, f.rowidx, f.val, n, 0, 1)
    # after the level reorder rows 5000 and 5001 sit somewhere in [0, n);
    # count long rows over all square blocks
    assert int(an["longrow"].sum()) == 1
    assert int(an["longlen"].sum()) == 2500
    sq = [i for i in range(1, len(bnz), 2) if an["longrow"][i] > 0][0]
    assert an["lenmax"][sq] == 2500
    assert an["nnzr"][sq] == (bnz[sq] - 2500) // an["m_new"][sq]


def test_block_diagonal_bfs_equals_per_block_bfs(oracle):
    """The device builds every triangular block's local levels from ONE BFS
    of the block-diagonal matrix D; check that premise on the CPU."""
    f = gen.random_lower(700, 3, seed=5, window=60)
    lv = 2
    re, part, bnz, an = oracle.full_pipeline(f.colptr, f.rowidx, f.val, f.n, 0, lv)
    cp, ri = re["colptr"], re["rowidx"]
    # D: keep in-triangle entries only
    keep_c, keep_r = [], []
    for t in range(1 << lv):
        u, d = part["up"][2 * t], part["down"][2 * t]
        for j in range(u, d):
            for q in range(cp[j], cp[j + 1]):
                if ri[q] < d:
                    keep_c.append(j)
                    keep_r.append(ri[q])
    keep_c = np.array(keep_c)
    dptr = np.zeros(f.n + 1, np.int32)
    np.cumsum(np.bincount(keep_c, minlength=f.n), out=dptr[1:])
    didx = np.array(keep_r, np.int32)
    rp = _rowptr(dptr, didx, f.n)
    rc, nl, lp, li = oracle.findlevel(dptr, didx, rp, f.n)
    assert rc == 0
    lev = np.empty(f.n, np.int64)
    for L in range(nl):
        lev[li[lp[L]:lp[L + 1]]] = L
    for t in range(1 << lv):
        u, d = part["up"][2 * t], part["down"][2 * t]
        sel = li[(li >= u) & (li < d)] - u
        assert sel.tolist() == an["lvitem_local"][t].tolist()
        cnt = np.bincount(lev[u:d], minlength=an["nlv"][2 * t])
        assert np.concatenate([[0], np.cumsum(cnt)]).tolist() == \
            an["lvptr_local"][t].tolist()


# ------------------------------------------------------------------ solve
CASES = [
    ("random", lambda: gen.random_lower(400, 4, seed=2)),
    ("banded", lambda: gen.banded(5000, seed=1)),
    ("laplacian", lambda: gen.laplacian(30, 20)),
    ("powerlaw", lambda: gen.powerlaw(20000, seed=4)),
]


@pytest.mark.parametrize("name,make", CASES)
@pytest.mark.parametrize("lower", [True, False])
@pytest.mark.parametrize("lv", [1, 3])
def test_solve_blocks_vs_scipy(oracle, name, make, lower, lv):
    f = make()
    if not lower:
        f = f.transpose()
    b = gen.rhs(f.n, seed=1)
    re, part, bnz, an = oracle.full_pipeline(f.colptr, f.rowidx, f.val, f.n,
                                             f.substitution, lv)
    x = oracle.solve_blocks(re, f.n, lv, f.substitution, part, b)
    rp, ci, v = f.csr()
    A = sp.csr_matrix((v, ci, rp), shape=(f.n, f.n))
    xs = sla.spsolve_triangular(A, b, lower=lower)
    assert rel_err(x, xs) < 1e-10
    xr = oracle.sptrsv_serial(rp, ci, v, f.n, f.substitution, b)
    assert rel_err(xr, xs) < 1e-12


@pytest.mark.parametrize("n", [1, 2, 5, 17, 64])
def test_solve_dense_brute_force(oracle, n):
    rng = np.random.default_rng(n)
    L = np.tril(rng.uniform(-1, 1, (n, n)))
    L[np.abs(L) < 0.5] = 0.0
    np.fill_diagonal(L, 2.0 + rng.random(n))
    for lower in (True, False):
        M = L if lower else L.T.copy()
        S = sp.csc_matrix(M)
        S.sort_indices()
        b = rng.uniform(-1, 1, n)
        lv = 1 if n >= 2 else 0
        re, part, bnz, an = oracle.full_pipeline(S.indptr, S.indices, S.data,
                                                 n, 0 if lower else 1, lv)
        x = oracle.solve_blocks(re, n, lv, 0 if lower else 1, part, b)
        xd = scipy.linalg.solve_triangular(M, b, lower=lower)
        assert rel_err(x, xd) < 1e-12


def test_solve_ones_closed_form(oracle):
    # banded workload: b = L * 1  =>  x = 1
    f = gen.banded(20000, seed=3)
    b = f.matvec(np.ones(f.n))
    lv = oracle.lv_auto(f.n)
    re, part, bnz, an = oracle.full_pipeline(f.colptr, f.rowidx, f.val, f.n, 0, lv)
    x = oracle.solve_blocks(re, f.n, lv, 0, part, b)
    assert np.abs(x - 1.0).max() < 1e-12


def test_solve_diagonal_closed_form(oracle):
    n = 50
    f = csc_from_deps(n, {}, diag=3.0)
    b = np.arange(n, dtype=np.float64)
    re, part, bnz, an = oracle.full_pipeline(f.colptr, f.rowidx, f.val, n, 0, 2)
    assert all(an["method"][2 * t] == 0 for t in range(4))
    x = oracle.solve_blocks(re, n, 2, 0, part, b)
    assert np.array_equal(x, b / 3.0)


# ---------------------------------------------------------- generators
def test_laplacian_generator_shape(oracle):
    f = gen.laplacian(1500)
    assert f.n == 2_250_000 and f.nnz == 6_747_000
    re = oracle.reorder(f.colptr, f.rowidx, f.val, f.n, 0)
    assert re["nlevel"] == 2999
    # level k of the grid holds the anti-diagonal r + c = k
    assert np.diff(re["levelPtr"])[:5].tolist() == [1, 2, 3, 4, 5]


def test_generators_deterministic():
    a, b = gen.banded(3000, seed=9), gen.banded(3000, seed=9)
    assert np.array_equal(a.rowidx, b.rowidx) and np.array_equal(a.val, b.val)
    p = gen.powerlaw(5000, seed=1)
    # columns sorted, diagonal first, lower triangular
    for j in range(0, p.n, 97):
        seg = p.rowidx[p.colptr[j]:p.colptr[j + 1]]
        assert seg[0] == j and np.all(np.diff(seg) > 0)
