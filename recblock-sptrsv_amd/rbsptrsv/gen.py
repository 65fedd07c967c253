"""Seeded synthetic triangular factors.

All generators return a ``Factor``: a square CSC matrix (int32 column
pointers / row ids, float64 values) with sorted row ids in every column, so
the diagonal is the first entry of a lower-triangular column and the last
entry of an upper-triangular one -- the input convention of
``recblocking_solver_cuda`` (R/recblocking_solver_cuda.h:21-32).

Workloads (sizes as named in the headline benchmark list):
  * ``banded``     n=100,000; diagonal 2+U(0,1); 8 draws (with replacement,
                   duplicates dropped) of columns uniform in
                   [max(0,i-2000), i-1]; values U(-0.1,0.1)/8; b = L*1.
  * ``laplacian``  lower part of the 5-point Laplacian on an nx-by-ny grid in
                   natural order: diagonal 4, west -1, south -1.  1500x1500 is
                   n=2,250,000, nnz=6,747,000, 2999 level sets.
  * ``powerlaw``   n=4,000,000; off-diagonal draw count per row from Zipf(2)
                   clipped to [1,512] (and to i, so row 0 has none); columns
                   uniform in [0,i), duplicates dropped; values U(-1,1);
                   diagonal 1 + row abs-sum.
Every generator is deterministic in its seed (numpy PCG64).
"""
from dataclasses import dataclass

import numpy as np


@dataclass
class Factor:
    n: int
    colptr: np.ndarray
    rowidx: np.ndarray
    val: np.ndarray
    lower: bool
    name: str = ""

    @property
    def nnz(self):
        return int(self.colptr[-1])

    @property
    def substitution(self):
        return 0 if self.lower else 1

    def csr(self):
        """(rowptr, colidx, val) with sorted columns per row."""
        return csc_transpose(self.n, self.colptr, self.rowidx, self.val)

    def transpose(self):
        """The transposed factor (L -> U = L^T and back), still CSC."""
        rp, ci, v = self.csr()
        return Factor(self.n, rp, ci, v, not self.lower, self.name + "^T")

    def matvec(self, x):
        rp, ci, v = self.csr()
        return csr_matvec(rp, ci, v, x)


def csc_transpose(n, ptr, idx, val):
    """Stable transposition of an n-by-n compressed matrix (the input's
    column order is kept inside each output row)."""
    ptr = np.asarray(ptr, np.int64)
    idx = np.asarray(idx, np.int64)
    nnz = int(ptr[-1])
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(ptr))
    order = np.argsort(idx, kind="stable")
    out_ptr = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(idx, minlength=n), out=out_ptr[1:])
    return (out_ptr.astype(np.int32), src[order].astype(np.int32),
            np.asarray(val, np.float64)[order] if nnz else np.zeros(0))


def csr_matvec(rowptr, colidx, val, x):
    rowptr = np.asarray(rowptr, np.int64)
    prod = np.asarray(val) * np.asarray(x)[np.asarray(colidx)]
    out = np.zeros(len(rowptr) - 1)
    nz = np.diff(rowptr) > 0
    if prod.size:
        sums = np.add.reduceat(prod, rowptr[:-1][nz])
        out[nz] = sums
    return out


def _from_rows(n, rows, cols, vals, lower, name):
    """Assemble CSC from COO (row, col) pairs; duplicates must already be
    removed.  Sorted by (col, row)."""
    rows = np.asarray(rows, np.int64)
    cols = np.asarray(cols, np.int64)
    key = cols * n + rows
    order = np.argsort(key, kind="stable")
    cols, rows, vals = cols[order], rows[order], np.asarray(vals)[order]
    colptr = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(cols, minlength=n), out=colptr[1:])
    return Factor(n, colptr.astype(np.int32), rows.astype(np.int32),
                  vals.astype(np.float64), lower, name)


def _dedupe(rows, cols, n):
    key = rows.astype(np.int64) * n + cols.astype(np.int64)
    key = np.unique(key)
    return key // n, key % n


def banded(n=100_000, seed=0, draws=8, window=2000):
    """Random banded lower factor (the n=100,000 workload)."""
    rng = np.random.default_rng(seed)
    i = np.repeat(np.arange(1, n, dtype=np.int64), draws)
    lo = np.maximum(0, i - window)
    c = lo + np.floor(rng.random(i.size) * (i - lo)).astype(np.int64)
    r, c = _dedupe(i, c, n)
    offv = rng.uniform(-0.1, 0.1, r.size) / 8.0
    d = 2.0 + rng.random(n)
    rows = np.concatenate([r, np.arange(n)])
    cols = np.concatenate([c, np.arange(n)])
    vals = np.concatenate([offv, d])
    return _from_rows(n, rows, cols, vals, True, "banded%d" % n)


def laplacian(nx=1500, ny=None, lower=True):
    """Lower (or, with lower=False, upper = transposed) part of the 5-point
    Laplacian on an nx-by-ny grid in natural ordering."""
    ny = nx if ny is None else ny
    n = nx * ny
    i = np.arange(n, dtype=np.int64)
    west = i[(i % nx) != 0]
    south = i[i >= nx]
    rows = np.concatenate([i, west, south])
    cols = np.concatenate([i, west - 1, south - nx])
    vals = np.concatenate([np.full(n, 4.0), np.full(west.size, -1.0),
                           np.full(south.size, -1.0)])
    f = _from_rows(n, rows, cols, vals, True, "laplacian%dx%d" % (nx, ny))
    return f if lower else f.transpose()


def powerlaw(n=4_000_000, seed=0, alpha=2.0, kmax=512):
    """Wide-level power-law lower factor (the n=4,000,000 workload)."""
    rng = np.random.default_rng(seed)
    k = rng.zipf(alpha, n).astype(np.int64)
    k = np.clip(k, 1, kmax)
    k = np.minimum(k, np.arange(n, dtype=np.int64))
    i = np.repeat(np.arange(n, dtype=np.int64), k)
    c = np.floor(rng.random(i.size) * i).astype(np.int64)
    r, c = _dedupe(i, c, n)
    offv = rng.uniform(-1.0, 1.0, r.size)
    absum = np.bincount(r, weights=np.abs(offv), minlength=n)
    d = 1.0 + absum
    rows = np.concatenate([r, np.arange(n)])
    cols = np.concatenate([c, np.arange(n)])
    vals = np.concatenate([offv, d])
    return _from_rows(n, rows, cols, vals, True, "powerlaw%d" % n)


def random_lower(n, density_per_row=4, seed=0, window=None, unit_diag=False):
    """Small random lower factor for tests (diagonally dominant)."""
    rng = np.random.default_rng(seed)
    i = np.repeat(np.arange(1, n, dtype=np.int64), density_per_row) if n > 1 \
        else np.zeros(0, np.int64)
    if window is None:
        c = np.floor(rng.random(i.size) * i).astype(np.int64)
    else:
        lo = np.maximum(0, i - window)
        c = lo + np.floor(rng.random(i.size) * (i - lo)).astype(np.int64)
    r, c = _dedupe(i, c, n) if i.size else (i, i)
    offv = rng.uniform(-1.0, 1.0, r.size)
    absum = np.bincount(r, weights=np.abs(offv), minlength=n)
    d = np.ones(n) if unit_diag else 1.0 + absum
    rows = np.concatenate([r, np.arange(n)])
    cols = np.concatenate([c, np.arange(n)])
    vals = np.concatenate([offv, d])
    return _from_rows(n, rows, cols, vals, True, "random%d" % n)


def batch_factors(count=64, seed=0, nmin=1_000, nmax=500_000):
    """Uneven batch: sizes log-uniform in [nmin, nmax], each a ``banded``
    factor with its own seed; RHS counts uniform in 1..32."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(nmin), np.log(nmax), count)).astype(int)
    nrhs = rng.integers(1, 33, count)
    facs = [banded(int(s), seed=seed * 1000 + j + 1) for j, s in enumerate(sizes)]
    return facs, [int(r) for r in nrhs]


def serve_factors(count=16, nmin=1e4, nmax=4e6, seed=100):
    """Resident factors of the serving workload: n log-spaced in
    [nmin, nmax], recipes cycling banded / Laplacian lower (the nearest
    square grid) / power-law, seeds seed + i."""
    facs = []
    for i in range(count):
        n = int(round(nmin * (nmax / nmin) ** (i / max(1, count - 1))))
        if i % 3 == 0:
            facs.append(banded(n, seed=seed + i))
        elif i % 3 == 1:
            facs.append(laplacian(int(round(n ** 0.5))))
        else:
            facs.append(powerlaw(n, seed=seed + i))
    return facs


def rhs(n, seed=0, k=1):
    """Seeded right-hand sides U(-1,1), column-major (n, k)."""
    rng = np.random.default_rng(seed + 7919)
    b = rng.uniform(-1.0, 1.0, (k, n)).T
    return np.ascontiguousarray(b[:, 0]) if k == 1 else np.asfortranarray(b)


WORKLOADS = {
    "banded": lambda: banded(100_000, seed=0),
    "laplacian": lambda: laplacian(1500),
    "powerlaw": lambda: powerlaw(4_000_000, seed=0),
}
