"""Shared test helpers: small matrices with known structure and checks that
characterise the reference's level order independently of any simulation of
its queue."""
import numpy as np

from rbsptrsv import gen


def csc_from_deps(n, deps, lower=True, diag=None, vals=None, seed=0):
    """CSC of an n x n triangular matrix with the given off-diagonal
    dependency lists (deps[i] = columns of row i)."""
    rng = np.random.default_rng(seed)
    rows, cols, vv = [], [], []
    for i in range(n):
        for c in deps.get(i, []):
            rows.append(i)
            cols.append(c)
            vv.append(rng.uniform(-1, 1) if vals is None else vals)
        rows.append(i)
        cols.append(i)
        vv.append(4.0 if diag is None else diag)
    f = gen._from_rows(n, np.array(rows), np.array(cols), np.array(vv), True,
                       "deps")
    return f if lower else f.transpose()


def levels_by_longest_path(rowptr, colidx, m, lower=True):
    """level(v) = 1 + max level over off-diagonal dependencies (0 if none),
    by a pass in dependency order."""
    lev = np.zeros(m, np.int64)
    order = range(m) if lower else range(m - 1, -1, -1)
    for i in order:
        best = -1
        for q in range(rowptr[i], rowptr[i + 1]):
            c = colidx[q]
            if c != i:
                best = max(best, lev[c])
        lev[i] = best + 1
    return lev


def check_bfs_order(colptr, rowidx, m, levelItem, levelPtr, lower=True):
    """Characterisation of findlevel's discovery order (R/findlevel.h:7-83)
    that does not replay its queue:
      * levelItem is a permutation, levels match the longest-path levels;
      * level 0 is ascending;
      * inside level k+1 rows are sorted by (pos(u*), offset of the row in
        column u*), u* = the dependency with the largest position."""
    colptr = np.asarray(colptr)
    rowidx = np.asarray(rowidx)
    levelItem = np.asarray(levelItem)
    assert sorted(levelItem.tolist()) == list(range(m))
    pos = np.empty(m, np.int64)
    pos[levelItem] = np.arange(m)
    # CSR of dependencies
    deps = [[] for _ in range(m)]
    off_in_col = {}
    for c in range(m):
        for k, q in enumerate(range(colptr[c], colptr[c + 1])):
            r = int(rowidx[q])
            if r != c:
                deps[r].append(c)
                off_in_col[(r, c)] = k
    lev_expect = np.zeros(m, np.int64)
    order = range(m) if lower else range(m - 1, -1, -1)
    for i in order:
        lev_expect[i] = 1 + max([lev_expect[c] for c in deps[i]], default=-1)
    nl = len(levelPtr) - 1
    for L in range(nl):
        seg = levelItem[levelPtr[L]:levelPtr[L + 1]]
        assert len(seg) > 0
        assert np.all(lev_expect[seg] == L), "level mismatch"
        if L == 0:
            assert np.all(np.diff(seg) > 0), "roots must be ascending"
        else:
            keys = []
            for v in seg:
                ustar = max(deps[v], key=lambda u: pos[u])
                keys.append((pos[ustar], off_in_col[(int(v), ustar)]))
            assert keys == sorted(keys), "discovery order violated"
    assert levelPtr[-1] == m


def block_sets(colptr, rowidx, val, bounds, i, subst):
    """Entries of block i of a permuted CSC as {(local_row, local_col): val},
    by the block rules of R/recblocking_preprocess.h:83-95,151-161."""
    up, down, left, right = (bounds[k][i] for k in ("up", "down", "left", "right"))
    out = {}
    for j in range(left, right):
        for q in range(colptr[j], colptr[j + 1]):
            r = int(rowidx[q])
            if i % 2 == 0:
                ok = r < down if subst == 0 else r >= up
            else:
                ok = up <= r < down
            if ok:
                out[(r - up, j - left)] = float(val[q])
    return out


def rel_err(x, xref):
    return float(np.abs(np.asarray(x) - xref).max() / max(np.abs(xref).max(), 1e-300))


def residual(f, x, b):
    r = np.asarray(b) - f.matvec(np.asarray(x))
    return float(np.linalg.norm(r) / max(np.linalg.norm(b), 1e-300))


def selector_factor(case):
    """Lower-triangular factor (lv = 1 layout, see tests/golden/selector.json)
    for one hand-worked selector case:
      fan_chain(R, K): rows 0..R-1 have no dependency; chain rows R..R+K-1,
          the first depending on every root, each later one on every root
          and on the chain row before it; then s = R + K rows that each
          depend on the last chain row (they form block tri1);
      chain(n): row i depends on row i-1 for 0 < i < n; then n rows that
          each depend on row n-1;
      long_chain(n): row i depends on row i-1 for every 0 < i < n;
      square(s, lens): rows 0..s-1 have no dependency; row s + a depends on
          columns 0..lens[a]-1; the remaining s - len(lens) rows each depend
          on row s."""
    k = case["kind"]
    deps = {}
    if k == "fan_chain":
        R, K = case["R"], case["K"]
        for j in range(K):
            row = R + j
            deps[row] = list(range(R)) + ([row - 1] if j > 0 else [])
        s = R + K
        for i in range(s, 2 * s):
            deps[i] = [s - 1]
        n = 2 * s
    elif k == "chain":
        c = case["n"]
        for i in range(1, c):
            deps[i] = [i - 1]
        for i in range(c, 2 * c):
            deps[i] = [c - 1]
        n = 2 * c
    elif k == "long_chain":
        n = case["n"]
        for i in range(1, n):
            deps[i] = [i - 1]
    elif k == "square":
        s, lens = case["s"], case["lens"]
        for a, L in enumerate(lens):
            deps[s + a] = list(range(L))
        for i in range(s + len(lens), 2 * s):
            deps[i] = [s]
        n = 2 * s
    else:
        raise ValueError(k)
    return csc_from_deps(n, deps, seed=7)
