"""The `sptrsv-double` driver (recblock-sptrsv/main.cu restated on the C
ABI) run end to end on Matrix Market files.

Each case is its own process, as the reference's program is.  This file is
collected before the other GPU tests: a process that has made cooperative
launches keeps other processes from making them, so the driver runs while
the test process has not touched the GPU yet.
"""
import os
import re
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

from rbsptrsv import gen, mmio

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "recblock-sptrsv_amd", "sptrsv-double")


def _laplacian_file(path, nx, symmetric):
    f = gen.laplacian(nx)
    rp, ci, v = f.csr()
    L = sp.csr_matrix((v, ci, rp), shape=(f.n, f.n))
    if symmetric:  # the lower triangle, as a symmetric file stores it
        A = L.tocoo()
        mmio.write_mtx(path, A.row, A.col, A.data, f.n, f.n, "symmetric")
    else:
        A = (L + L.T - sp.diags(L.diagonal())).tocoo()
        perm = np.random.default_rng(0).permutation(A.nnz)  # any file order
        mmio.write_mtx(path, A.row[perm], A.col[perm], A.data[perm], f.n, f.n)
    return f.n


def _run(args):
    print("driver:", " ".join(args), flush=True)
    out = subprocess.run([CLI] + args, capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    m = re.search(r"\|x-xref\|/\|xref\| = +([0-9.eE+-]+)", out.stdout)
    assert m, out.stdout + out.stderr
    return float(m.group(1)), out.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("direction", ["-forward", "-backward"])
@pytest.mark.parametrize("symmetric", [True, False])
@pytest.mark.parametrize("rhs,extra", [
    (1, []),                                   # the reference's entry point
    (1, ["--reference-values"]),               # the reference's rewrite
    (3, ["--schedule", "fused"]),
    (2, ["--schedule", "blocked", "--reference-values"]),
])
def test_driver_solves_mtx(tmp_path, direction, symmetric, rhs, extra):
    """x check passed: file values (the Laplacian, well conditioned) to
    1e-12; the reference's integer rewrite exactly (every value and
    intermediate is an integer below 2^53)."""
    assert os.path.exists(CLI), "sptrsv-double is not built"
    p = os.path.join(str(tmp_path), "lap.mtx")
    n = _laplacian_file(p, 30, symmetric)
    err, text = _run(["-d", "0", "-rhs", str(rhs), "-lv", "-1", direction,
                      "-mtx", p, "--seed", "5", "--repeat", "3"] + extra)
    assert "input matrix A: ( %d, %d )" % (n, n) in text
    assert "x check passed" in text
    if "--reference-values" in extra:
        assert err == 0.0
    else:
        assert err < 1e-12


@pytest.mark.gpu
def test_driver_non_square_and_bad_file(tmp_path):
    p = os.path.join(str(tmp_path), "rect.mtx")
    mmio.write_mtx(p, [0, 1], [0, 2], [1.0, 2.0], 2, 3)
    out = subprocess.run([CLI, "-d", "0", "-rhs", "1", "-lv", "1", "-forward",
                          "-mtx", p], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "we need square matrix. Exit!" in out.stdout
    out = subprocess.run([CLI, "-d", "0", "-rhs", "1", "-lv", "1", "-forward",
                          "-mtx", os.path.join(str(tmp_path), "none.mtx")],
                         capture_output=True, text=True, timeout=60)
    assert out.returncode != 0 and "cannot read" in out.stderr
