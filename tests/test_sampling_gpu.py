"""estimateNNZ_sampling on the device (cbx_spgemm_estimate_sampling) against
the oracle's restatement, itself pinned to the genuine template: every
column's estimate within a few units in the last place (the device and host
logf may round -log(1-u) differently; the minima and sums are the same
operations), and the learned single pass driven by these estimates (the
method without the network) still gives the exact product."""
from __future__ import annotations

import numpy as np
import pytest

import combblas_amd as cb
import oracle
from tests.helpers import bit_exact

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scale,ef,rounds", [(12, 8, 5), (14, 8, 5), (13, 16, 3), (14, 8, 8)])
def test_device_estimates_match_oracle(scale, ef, rounds):
    A = cb.gen_rmat(scale, ef)
    B = cb.gen_rmat(scale, ef, seed=321)
    for X, Y in ((A, A), (A, B)):
        want, wtot, _ = oracle.estimate_sampling(X, Y, rounds)
        got, gtot = cb.spgemm_estimate_sampling(X.to_device(), Y.to_device(), rounds)
        assert got.shape == want.shape
        assert np.allclose(got, want, rtol=2e-6, atol=0)
        # the device sums the column estimates in FP64, the reference in a
        # running FP32 sum over the N columns: the device total is the exact
        # sum of its columns, and the reference's lies within gamma_(N-1) of it
        exact = float(np.sum(want.astype(np.float64)))
        assert abs(gtot - exact) <= 2e-6 * exact
        k = (want.size - 1) * 2.0**-24
        assert abs(wtot - exact) <= k / (1 - k) * exact


def test_sampling_predictor_learned_pass_exact():
    A = cb.gen_rmat(14, 8)
    cb.set_predictor(None)
    cb.set_sampling_predictor(5)
    try:
        dC, st = cb.spgemm(A.to_device(), A.to_device(), "plus_times", "learned")
    finally:
        cb.set_sampling_predictor(0)
    want, _ = oracle.spgemm(A, A)
    bit_exact(dC.to_host(), want)
    assert st["predicted_nnz"] > 0 or st["learned_fallback"] == 1


def test_learned_mode_needs_a_predictor():
    cb.set_predictor(None)
    cb.set_sampling_predictor(0)
    A = cb.gen_rmat(10, 8).to_device()
    with pytest.raises(cb.CbxError) as ei:
        cb.spgemm(A, A, "plus_times", "learned")
    assert ei.value.status == cb.ERR_INVALID
    with pytest.raises(cb.CbxError):
        cb.set_sampling_predictor(1)
