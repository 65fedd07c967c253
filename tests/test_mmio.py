"""Matrix Market input and the driver's factor (CPU only).

The oracle's restatements (oracle/pyoracle.py: mm_read_csr, mm_triangle,
c_rand) are pinned first -- by a hand-worked file whose CSR is written out
below, by scipy.io.mmread, and by an independent model of the C library's
rand() -- and then the library's host entry points (rb_mmio_info,
rb_mmio_data, rb_mmio_triangle) are compared with them bit for bit.
"""
import os

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

from rbsptrsv import gen, mmio


def _write(tmp_path, name, text):
    p = os.path.join(str(tmp_path), name)
    with open(p, "w") as fh:
        fh.write(text)
    return p


# A 4x4 symmetric file, entries deliberately out of row order.  Worked by
# hand from mmio_highlevel.h: counts per row (an off-diagonal entry counts
# for both rows) 1:(3,1),(2,1)->row2,row3... see the arrays below.
HAND = """%%MatrixMarket matrix coordinate real symmetric
% comment line
4 4 5
3 1 2.5
1 1 4
4 3 -1
2 2 5

4 2 7e-1
"""
# entries (0-based, file order): (2,0,2.5) (0,0,4) (3,2,-1) (1,1,5) (3,1,0.7)
# rows: 0 <- (0,2 mirror) then (0,0): [2, 0] values [2.5, 4]
#       1 <- (1,1) then (1,3 mirror): [1, 3] values [5, 0.7]
#       2 <- (2,0) then (2,3 mirror): [0, 3] values [2.5, -1]
#       3 <- (3,2) then (3,1): [2, 1] values [-1, 0.7]
HAND_ROWPTR = [0, 2, 4, 6, 8]
HAND_COL = [2, 0, 1, 3, 0, 3, 2, 1]
HAND_VAL = [2.5, 4.0, 5.0, 0.7, 2.5, -1.0, -1.0, 0.7]


def test_oracle_reader_handworked(oracle, tmp_path):
    """Pins the oracle's order rule: file order inside a row, a mirrored
    pair written as (i, j) then (j, i).  A reader that sorts columns or
    mirrors in a second pass fails."""
    p = _write(tmp_path, "hand.mtx", HAND)
    m, n, rp, ci, v, sym = oracle.mm_read_csr(p)
    assert (m, n, sym) == (4, 4, True)
    assert rp.tolist() == HAND_ROWPTR
    assert ci.tolist() == HAND_COL
    assert v.tolist() == HAND_VAL


def _random_file(tmp_path, name, seed, m, n, nz, symmetry, field):
    rng = np.random.default_rng(seed)
    rows = rng.integers(0, m, nz)
    cols = rng.integers(0, n, nz)
    if symmetry != "general":  # one triangle, as symmetric files store
        rows, cols = np.maximum(rows, cols), np.minimum(rows, cols)
        key = np.unique(rows * n + cols)
        rows, cols = key // n, key % n
        perm = rng.permutation(len(rows))
        rows, cols = rows[perm], cols[perm]
    else:
        key = np.unique(rows * n + cols)
        perm = rng.permutation(len(key))
        rows, cols = key[perm] // n, key[perm] % n
    vals = rng.integers(-9, 10, len(rows)) if field == "integer" else \
        rng.uniform(-1, 1, len(rows))
    p = os.path.join(str(tmp_path), name)
    mmio.write_mtx(p, rows, cols, vals, m, n, symmetry, field)
    return p


CASES = [("general", "real", 40, 40), ("symmetric", "real", 50, 50),
         ("general", "integer", 30, 45), ("symmetric", "pattern", 35, 35),
         ("hermitian", "complex", 25, 25), ("general", "pattern", 20, 20)]


@pytest.mark.parametrize("symmetry,field,m,n", CASES)
def test_oracle_reader_matches_scipy(oracle, tmp_path, symmetry, field, m, n):
    """The oracle's CSR holds the same matrix as scipy.io.mmread (mirrored
    for symmetric/hermitian; real part of complex); pins values, mirroring
    and 1-based indexing."""
    p = _random_file(tmp_path, "a.mtx", 7, m, n, 3 * m, symmetry, field)
    mm, nn, rp, ci, v, sym = oracle.mm_read_csr(p)
    assert (mm, nn, sym) == (m, n, symmetry != "general")
    A = sp.csr_matrix((v, ci, rp), shape=(m, n))
    R = scipy.io.mmread(p)
    R = sp.csr_matrix(R.real if np.iscomplexobj(R) else R, dtype=np.float64)
    assert abs(A - R).max() == 0.0
    assert A.nnz == R.nnz  # no duplicates were merged away


@pytest.mark.parametrize("symmetry,field,m,n", CASES)
def test_reader_matches_oracle(oracle, tmp_path, symmetry, field, m, n):
    """The library's rb_mmio_info / rb_mmio_data build the oracle's CSR bit
    for bit (order inside rows included)."""
    p = _random_file(tmp_path, "b.mtx", 11, m, n, 4 * m, symmetry, field)
    want = oracle.mm_read_csr(p)
    got = mmio.read_csr(p)
    assert got[:2] == want[:2] and got[5] == want[5]
    for g, w in zip(got[2:5], want[2:5]):
        assert np.array_equal(g, w)


def test_reader_handworked(tmp_path):
    p = _write(tmp_path, "hand.mtx", HAND)
    m, n, rp, ci, v, sym = mmio.read_csr(p)
    assert (m, n, sym) == (4, 4, True)
    assert rp.tolist() == HAND_ROWPTR and ci.tolist() == HAND_COL
    assert v.tolist() == HAND_VAL


@pytest.mark.parametrize("text,code", [
    ("not a banner\n1 1 1\n1 1 1\n", -2),
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", -2),
    ("%%MatrixMarket matrix coordinate real general\n2 2\n", -4),
    ("%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1.0\n", -5),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n", -5),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 0 1.0\n", -5),
])
def test_reader_errors(tmp_path, text, code):
    """Malformed files give the reference's codes (-2 banner, -4 size line)
    or -5 where the reference would read or write out of bounds."""
    p = _write(tmp_path, "bad.mtx", text)
    with pytest.raises(ValueError, match=r"\(%d\)" % code):
        mmio.read_csr(p)


def test_reader_missing_file(tmp_path):
    with pytest.raises(ValueError, match=r"\(-1\)"):
        mmio.read_csr(os.path.join(str(tmp_path), "nope.mtx"))


def test_c_rand_matches_glibc_model(oracle):
    """The oracle draws the driver's values with the C library's rand();
    an independent model of glibc's generator pins it (a different
    generator or seeding fails)."""
    for seed in (0, 1, 12345, 2 ** 31 + 7):
        assert oracle.c_rand(seed, 400) == oracle.glibc_random_model(seed, 400)
    assert oracle.c_rand(1, 1) == [1804289383]  # glibc's first value


@pytest.mark.parametrize("subst", [0, 1])
def test_oracle_triangle_closed_form(oracle, tmp_path, subst):
    """The oracle's factor: strictly lower (upper) entries of each row in
    row order with the diagonal appended (prepended); reference values are
    integers 1..10 with a unit diagonal drawn in that order; b = T x_ref."""
    p = _random_file(tmp_path, "t.mtx", 3, 30, 30, 120, "general", "real")
    m, _, rp, ci, v, _ = oracle.mm_read_csr(p)
    tp, tc, tv, xref, b = oracle.mm_triangle(m, rp, ci, v, subst, True, 99, 2)
    draws = oracle.c_rand(99, int(tp[-1]) - m + 2 * m)
    k = 0
    for i in range(m):
        row = tc[tp[i]:tp[i + 1]].tolist()
        vals = tv[tp[i]:tp[i + 1]].tolist()
        if subst == 0:
            assert row[-1] == i and vals[-1] == 1.0
            kept = [c for c in ci[rp[i]:rp[i + 1]].tolist() if c < i]
            assert row[:-1] == kept
            off = vals[:-1]
        else:
            assert row[0] == i and vals[0] == 1.0
            kept = [c for c in ci[rp[i]:rp[i + 1]].tolist() if c > i]
            assert row[1:] == kept
            off = vals[1:]
        assert off == [float(d % 10 + 1) for d in draws[k:k + len(off)]]
        k += len(off)
    assert xref.tolist() == [float(d % 10 + 1) for d in draws[k:k + 2 * m]]
    T = sp.csr_matrix((tv, tc, tp), shape=(m, m)).toarray()
    for r in range(2):
        assert np.allclose(T @ xref[r * m:(r + 1) * m], b[r * m:(r + 1) * m],
                           rtol=0, atol=1e-12)


@pytest.mark.parametrize("subst,refv", [(0, True), (1, True), (0, False),
                                        (1, False)])
def test_triangle_matches_oracle(oracle, tmp_path, subst, refv):
    """rb_mmio_triangle builds the oracle's factor, x_ref and b bit for bit
    (file values: a diagonal is added to every row first)."""
    p = _random_file(tmp_path, "u.mtx", 5, 60, 60, 300, "symmetric", "real")
    m, _, rp, ci, v, _ = oracle.mm_read_csr(p)
    if not refv:  # give every row a diagonal (appended: file order kept)
        rows = np.repeat(np.arange(m), np.diff(rp))
        rows = np.concatenate([rows, np.arange(m)])
        cols = np.concatenate([ci, np.arange(m)])
        vals = np.concatenate([v, np.full(m, 8.0)])
        order = np.argsort(rows, kind="stable")
        ci, v = cols[order].astype(np.int32), vals[order]
        rp = np.zeros(m + 1, np.int32)
        np.cumsum(np.bincount(rows, minlength=m), out=rp[1:])
    want = oracle.mm_triangle(m, rp, ci, v, subst, refv, 4242, 3)
    got = mmio.triangle(m, rp, ci, v, subst, refv, 4242, 3)
    for g, w in zip(got, want):
        assert np.array_equal(g, w)


def test_triangle_missing_diagonal(tmp_path):
    """File values without a diagonal in some row: an error code, not a
    factor the solver would reject later (or the reference would hang on)."""
    from rbsptrsv import RbError
    rp = np.array([0, 1, 2], np.int32)
    ci = np.array([0, 0], np.int32)
    v = np.array([2.0, 1.0])
    with pytest.raises(RbError, match="RB_ERR_MISSING_DIAG"):
        mmio.triangle(2, rp, ci, v, 0, False)
    mmio.triangle(2, rp, ci, v, 0, True)  # the rewrite adds a unit diagonal


def test_load_factor_is_solver_input(tmp_path):
    """load_factor gives the solver's CSC convention: row ids ascending in
    every column, the diagonal first (L) / last (U), and b = T x_ref."""
    f0 = gen.laplacian(12)
    rp, ci, v = f0.csr()
    sym = sp.csr_matrix((v, ci, rp), shape=(f0.n, f0.n))
    full = (sym + sym.T - sp.diags(sym.diagonal())).tocoo()
    p = os.path.join(str(tmp_path), "lap.mtx")
    mmio.write_mtx(p, full.row, full.col, full.data, f0.n, f0.n)
    for subst in (0, 1):
        f, xref, b = mmio.load_factor(p, subst, reference_values=False, seed=1)
        for j in range(f.n):
            rows = f.rowidx[f.colptr[j]:f.colptr[j + 1]]
            assert np.all(np.diff(rows) > 0)
            assert rows[0 if subst == 0 else -1] == j
        assert np.allclose(f.matvec(xref[0]), b[0], rtol=1e-14, atol=1e-12)
