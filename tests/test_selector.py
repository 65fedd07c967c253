"""The adaptive selector at every threshold, against hand-worked answers.

tests/golden/selector.json lists structures whose selector inputs (blk_nnz,
nlv, nnzr, m_new, lenmax, longrow, longlen, empty ratio, nnz per level) and
decisions were derived on paper from the reference's rules
(R/recblocking_solver_cuda.h:176-367 triangular, :441-546 square;
R/utils_cuda.h:781-830 find_level_nnz / cal_longrow); the derivations are in
tests/golden/README.md.  Each threshold is hit from both sides (nnzr 15/16
with nlv 20/21, nnzr 1 with nlv 100/101, nlv 20000/20001, square nnzr 12/13
against empty 15/50 and just above, a row of 2048 against 2049 entries), so
an off-by-one, a swapped comparison or a wrong divisor in the oracle fails
here.  The CPU tests pin the oracle; the GPU tests hold the device pipeline
to the same answers and solve the method-1 case (the level-set kernel that
replaces cuSPARSE csrsv2) against the oracle."""
import json
import os

import numpy as np
import pytest

from helpers import rel_err, residual, selector_factor

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "selector.json")))
CASES = FIX["tri"] + FIX["sq"]
BLOCKS = (("tri0", 0), ("sq1", 1), ("tri1", 2))


def _got(part, bnz, an, blk, i, key):
    if key == "blk_m":
        return int(part["blk_m"][i])
    if key == "blk_nnz":
        return int(bnz[i])
    if key == "nnz_lv":
        return [int(v) for v in an["nnz_lv"][i // 2]]
    v = an[key][i]
    return float(v) if key == "empty_ratio" else int(v)


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_selector_handworked(oracle, case):
    f = selector_factor(case)
    re, part, bnz, an = oracle.full_pipeline(f.colptr, f.rowidx, f.val, f.n,
                                             0, 1)
    for blk, i in BLOCKS:
        for key, want in case.get(blk, {}).items():
            assert _got(part, bnz, an, blk, i, key) == want, (blk, key)


def test_oracle_find_level_nnz_position_quirk(oracle):
    """find_level_nnz sums the row lengths of the block-local CSR over the
    POSITION range [levelPtr[l], levelPtr[l+1]) -- rows in row-id order --
    not over the rows that sit in level l (R/utils_cuda.h:781-796).  On a
    block whose discovery order is not its row order the two differ; the
    fixture holds both answers so the test tells them apart."""
    q = FIX["find_level_nnz_quirk"]
    from helpers import csc_from_deps
    n = q["n"]
    f = csc_from_deps(n, {int(k): v for k, v in q["deps"].items()})
    one = np.zeros(1, np.int32)
    part = {"blk_m": one + n, "blk_n": one + n, "up": one.copy(),
            "down": one + n, "left": one.copy(), "right": one + n}
    an = oracle.analyze(f.colptr, f.rowidx, n, 0, 0, part,
                        np.array([f.nnz], np.int32))
    assert int(an["nlv"][0]) == q["nlv"]
    assert an["lvptr_local"][0].tolist() == q["lvptr"]
    assert an["lvitem_local"][0].tolist() == q["lvitem"]
    assert an["nnz_lv"][0].tolist() == q["nnz_lv_position_ranges"]
    assert q["nnz_lv_position_ranges"] != q["nnz_lv_true_levels"]
    assert int(an["method"][0]) == q["method"]
    assert int(an["nnzr"][0]) == q["nnzr"]


# ----------------------------------------------------------------- GPU
@pytest.fixture(scope="module")
def torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a device"
    return torch


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_gpu_selector_handworked(torch, case):
    from rbsptrsv import Plan
    print("selector case", case["name"], flush=True)
    f = selector_factor(case)
    p = Plan.from_factor(f, lv=1, schedule=1)
    try:
        bl = p.blocks()
        for blk, i in BLOCKS:
            for key, want in case.get(blk, {}).items():
                if key == "nnz_lv":
                    got = p.tri_levels(i // 2)[2].tolist()
                elif key == "empty_ratio":
                    got = float(bl[key][i])
                else:
                    got = int(bl[key][i])
                assert got == want, (blk, key)
    finally:
        p.close()


@pytest.mark.gpu
@pytest.mark.parametrize("schedule", [1, 0])
def test_gpu_method1_chain_solve(torch, oracle, schedule):
    """A 40,002-row chain at lv = 1: both triangular blocks have 20,001
    local levels, so the reference would hand them to cuSPARSE csrsv2
    (method 1, R/recblocking_solver_cuda.h:209-291); here they run this
    library's own kernel, which honours the stored diagonal."""
    from rbsptrsv import Plan, gen
    case = [c for c in FIX["tri"] if c["name"] == "long_chain_40002"][0]
    print("method-1 chain, schedule", schedule, flush=True)
    f = selector_factor(case)
    b = gen.rhs(f.n, seed=3)
    re, part, bnz, an = oracle.full_pipeline(f.colptr, f.rowidx, f.val, f.n,
                                             0, 1)
    xo = oracle.solve_blocks(re, f.n, 1, 0, part, b)
    p = Plan.from_factor(f, lv=1, schedule=schedule)
    try:
        assert p.blocks()["method"].tolist() == [1, 1, 1]
        x = p.solve(torch.as_tensor(b, device="cuda"))
        p.check()
        x = x.cpu().numpy()
    finally:
        p.close()
    assert rel_err(x, xo) < 1e-10
    assert residual(f, x, b) < 1e-12
