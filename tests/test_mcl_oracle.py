"""The MCL prune / recover / select restatement (oracle/cbx_oracle.cpp,
MCLPruneRecoverySelect of ParFriends.h:185-345) pinned against the genuine
reference (oracle/_ref), against the committed fixtures written from it, and
against what the definition fixes: survivors are a subset of the input with
their order kept, every kept value of a column is >= every dropped one, a
column that is neither recovered nor selected keeps exactly its values >= the
hard threshold, and a selected column keeps at least `select` entries and
every value tied with the smallest survivor."""
from pathlib import Path

import numpy as np
import pytest

import combblas_amd as cb
import oracle
from helpers import MCL_PARAMS, bit_exact, mcl_columns

GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (needs the reference tree)")
@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("params", MCL_PARAMS)
def test_mcl_prune_matches_reference(dtype, params):
    for seed in (1, 2, 3):
        A = mcl_columns(seed, dtype=dtype)
        bit_exact(oracle.mcl_prune(A, *params), oracle.ref_mcl_prune(A, *params))


def test_mcl_prune_golden():
    files = sorted(GOLDEN.glob("mcl_prune_*.npz"))
    assert files, "golden MCL fixtures missing"
    for f in files:
        z = np.load(f)
        A = cb.Dcsc(int(z["m"]), int(z["n"]), z["a_cp"], z["a_jc"], z["a_ir"].astype(np.int64), z["a_val"], str(z["dtype"]))
        params = (float(z["thr"]), int(z["select"]), int(z["recover"]), float(z["pct"]))
        R = cb.Dcsc(int(z["m"]), int(z["n"]), z["c_cp"], z["c_jc"], z["c_ir"].astype(np.int64), z["c_val"], str(z["dtype"]))
        bit_exact(oracle.mcl_prune(A, *params), R)


@pytest.mark.parametrize("params", MCL_PARAMS)
def test_mcl_prune_properties(params):
    thr, select, recover, pct = params
    A = mcl_columns(7, dtype="float32")
    X, cut = oracle.mcl_prune(A, *params, return_cut=True)
    T = np.float32
    kept_cols = set(X.jc.tolist())
    xcol = {int(j): (X.ir[X.cp[q]:X.cp[q + 1]], X.val[X.cp[q]:X.cp[q + 1]]) for q, j in enumerate(X.jc)}
    for q, j in enumerate(A.jc):
        ir = A.ir[A.cp[q]:A.cp[q + 1]]
        v = A.val[A.cp[q]:A.cp[q + 1]]
        keep = ~(v < cut[q])
        if keep.any():
            r2, v2 = xcol[int(j)]
            assert np.array_equal(r2, ir[keep]) and np.array_equal(v2, v[keep])
            if (~keep).any():
                assert v[keep].min() >= v[~keep].max()
        else:
            assert int(j) not in kept_cols
        nP = int((v > T(thr)).sum())
        sP = float(v[v > T(thr)].astype(np.float64).sum())
        recovered = nP < recover and v.size > nP and sP < pct
        selected = (not recovered) and select > 0 and nP > select
        if not recovered and not selected:
            assert cut[q] == T(thr), (j, cut[q])
        if selected:
            assert keep.sum() >= select
            srt = np.sort(v)[::-1]
            if cut[q] == srt[select - 1]:
                assert keep.sum() == int((v >= srt[select - 1]).sum())  # ties kept
        if recovered and v.size >= recover:
            assert keep.sum() >= recover
        if recovered and v.size < recover:
            assert keep.all()


def test_powerlaw_generator():
    A = cb.gen_powerlaw_stochastic(20000, seed=9)
    B = cb.gen_powerlaw_stochastic(20000, seed=9)
    assert np.array_equal(A.ir, B.ir) and np.array_equal(A.val, B.val)
    assert A.dtype == "float32" and A.nzc == A.n  # every vertex has degree >= 1
    sums = np.add.reduceat(A.val.astype(np.float64), A.cp[:-1])
    assert np.allclose(sums, 1.0, atol=1e-5)
    deg = np.diff(A.cp)
    assert 40 < deg.mean() <= 50.5  # duplicates merge a few of the 50 stubs per column on average
    rowdeg = np.bincount(A.ir, minlength=A.m)
    assert rowdeg.max() > 10 * rowdeg.mean()  # heavy tail on both sides
    for q in range(0, A.nzc, 997):
        assert np.all(np.diff(A.ir[A.cp[q]:A.cp[q + 1]]) > 0)
