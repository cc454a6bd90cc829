"""Learned single-pass SpGEMM (CBX_SPGEMM_LEARNED) must equal the exact result
whatever the predictor says.  Constant predictors steer every path: units far
larger than predicted (re-run exactly in rank windows, counted), a predicted
total too small for the output (fallback to the two-phase path), and
predictions that fit; the trained model (if present) must need no fallback."""
from __future__ import annotations

import numpy as np
import pytest

import combblas_amd as cb
import oracle
from combblas_amd import predictor
from combblas_amd.checks import product_matches
from tests.helpers import bit_exact, same_pattern

pytestmark = pytest.mark.gpu
B3 = 64 * 6 + 64 + 64 * 64 + 64 + 64  # index of the output bias


def constant_predictor(log2_nnz: float) -> np.ndarray:
    P = np.zeros(cb.MLP_PARAMS, np.float32)
    P[B3] = log2_nnz
    P[-6:] = 1.0  # feature scale
    return P


@pytest.fixture(autouse=True)
def _reset_predictor():
    yield
    cb.set_predictor(None)


def learned(A, B, semiring="plus_times"):
    dC, st = cb.spgemm(A.to_device(), B.to_device(), semiring, "learned")
    return dC.to_host(), st


def test_learned_requires_predictor():
    cb.set_predictor(None)
    A = cb.gen_rmat(10, 8)
    with pytest.raises(cb.CbxError) as ei:
        learned(A, A)
    assert ei.value.status == cb.ERR_INVALID


def unit_tier_nnz(A, B):
    """Exact output sizes of the columns that go to the unit tier (flop_j > 4096)."""
    flop, nnz = oracle.symbolic(A, B)
    return nnz[flop > 4096]


@pytest.mark.parametrize("expect", ["fallback", "fits"])
def test_learned_constant_predictors_exact(expect):
    """fallback: 1 output per column, far below the total; fits: every column
    predicted at the largest size."""
    A = cb.gen_rmat(14, 8)
    sizes = unit_tier_nnz(A, A)
    log2_nnz = {"fallback": 1.0, "fits": float(np.log2(1 + sizes.max()))}[expect]
    cb.set_predictor(constant_predictor(log2_nnz))
    got, st = learned(A, A)
    want, _ = oracle.spgemm(A, A)
    bit_exact(got, want)
    if expect == "fallback":
        assert st["learned_fallback"] == 1
    else:
        assert st["learned_fallback"] == 0 and st["units_light"] > 0 and st["overflow_cols"] == 0


@pytest.mark.skipif(not predictor.MODEL_PATH.exists(), reason="no trained predictor in combblas_amd/models")
def test_learned_overflow_windows_exact():
    """Units cut three times larger than their value window: every large unit
    is re-swept in windows of ranks, counted, and the output stays exact."""
    A = cb.gen_rmat(16, 16)
    base = cb.get_config()
    cb.set_predictor(predictor.load())
    try:
        cb.set_config(learned_target=3.0)  # units three times their 4096-entry window
        got, st = learned(A, A)
    finally:
        cb.set_config(**base)
    assert st["learned_fallback"] == 0 and st["overflow_cols"] > 0
    want, _ = oracle.spgemm(A, A)
    bit_exact(got, want)


@pytest.mark.parametrize("semiring,dtype", [("plus_times", "int64"), ("plus_times", "bool"), ("min_plus", "int64"),
                                            ("select_max", "int64"), ("plus_times", "float64")])
def test_learned_semirings(semiring, dtype):
    rng = np.random.default_rng(2)
    A = cb.gen_rmat(13, 8).astype(dtype)
    if dtype == "int64":
        A.val = rng.integers(-9, 10, A.nnz).astype(np.int64)
    cb.set_predictor(constant_predictor(float(np.log2(1 + unit_tier_nnz(A, A).max()))))
    got, st = learned(A, A, semiring)
    assert st["learned_fallback"] == 0
    want, _ = oracle.spgemm(A, A, semiring)
    if dtype == "float64":
        same_pattern(got, want)
        assert np.allclose(got.val, want.val, rtol=1e-12, atol=0)
    else:
        bit_exact(got, want)


@pytest.mark.skipif(not predictor.MODEL_PATH.exists(), reason="no trained predictor in combblas_amd/models")
def test_learned_trained_model_s16():
    A = cb.gen_rmat(16, 16)
    cb.set_predictor(predictor.load())
    got, st = learned(A, A)
    assert st["learned_fallback"] == 0
    want, _ = oracle.spgemm(A, A)
    bit_exact(got, want)
    # the columns the single pass sized by prediction: their exact total is a
    # part of the product, and the trained network's total lands near it
    assert st["units_light"] > 0
    assert 0 < st["actual_nnz_predicted"] <= st["nnz"]
    assert 0.5 < st["predicted_nnz"] / st["actual_nnz_predicted"] < 2.0


@pytest.mark.skipif(not predictor.MODEL_PATH.exists(), reason="no trained predictor in combblas_amd/models")
def test_learned_trained_model_s18_checksums():
    A = cb.gen_rmat(18, 16)
    cb.set_predictor(predictor.load())
    dA = A.to_device()
    dC, st = cb.spgemm(dA, dA, "plus_times", "learned")
    assert st["learned_fallback"] == 0
    dC.check()
    got = dC.column_checksums(salt=99)
    _, jcC = dC.structure()
    jcB, want = oracle.product_checksums(A, A, salt=99)
    pos = np.searchsorted(jcB, jcC)
    assert np.array_equal(jcB[pos], jcC) and np.array_equal(got, want[pos])


@pytest.mark.skipif(not predictor.MODEL_PATH.exists(), reason="no trained predictor in combblas_amd/models")
def test_trained_network_beats_sampling_baseline_s18():
    """What the trained network gains over the method without it, the
    reference's own estimator (estimateNNZ_sampling's min-hash estimates,
    5 rounds, per column): per-column sizes of the unit-tier columns at least
    twice as accurate; both drive the single pass to the exact product (the
    estimator may over- or under-size the capacity and fall back)."""
    A = cb.gen_rmat(18, 16)
    dA = A.to_device()
    X = cb.spgemm_features(dA, dA)
    nnz = cb.spgemm_symbolic(dA, dA).astype(np.float64)
    flop = np.exp2(X[:, 1].astype(np.float64)) - 1
    big = flop > 4096
    err = lambda pred: float(np.mean(np.abs(pred[big] - nnz[big]) / np.maximum(nnz[big], 1)))
    mlp_pred = np.exp2(cb.mlp_forward(predictor.load(), X).astype(np.float64)) - 1
    est, _ = cb.spgemm_estimate_sampling(dA, dA, 5)
    assert err(mlp_pred) < 0.5 * err(est.astype(np.float64)), (err(mlp_pred), err(est))
    counts = nnz.astype(np.int64)
    for name in ("baseline", "trained"):
        if name == "baseline":
            cb.set_predictor(None)
            cb.set_sampling_predictor(5)
        else:
            cb.set_sampling_predictor(0)
            cb.set_predictor(predictor.load())
        C, st = cb.spgemm(dA, dA, "plus_times", "learned")
        if name == "trained":
            assert st["learned_fallback"] == 0
        assert product_matches(C, A, A, col_nnz=counts)
        C.release()
    cb.set_sampling_predictor(0)


@pytest.mark.skipif(not predictor.MODEL_PATH.exists(), reason="no trained predictor in combblas_amd/models")
@pytest.mark.parametrize("scale,ef", [(14, 8), (16, 16)])
@pytest.mark.parametrize("share", [0.25, 0.0])
def test_learned_heavy_columns_exact(scale, ef, share):
    """heavy_nnz at its lowest (4096): columns predicted larger than that take
    the exact path for predicted-heavy columns, with row maps (share > 0) and
    without them; the whole product is bit-exact against the oracle."""
    A = cb.gen_rmat(scale, ef)
    base = cb.get_config()
    cb.set_predictor(predictor.load())
    try:
        cb.set_config(heavy_nnz=4096, row_map_share=share)
        got, st = learned(A, A)
    finally:
        cb.set_config(**base)
    assert st["learned_fallback"] == 0
    assert st["units_heavy"] > 0
    want, _ = oracle.spgemm(A, A)
    bit_exact(got, want)
