"""The benchmark workloads at their full sizes, against the oracle:
256 batched Galerkin products (grid edges 16..96) entry for entry; the MCL
expansion step at n = 10^6 on a column sample within the FP32 bound of the
FP64 oracle, with the exact two-phase sizing and with the trained predictor's
learned sizing; the learned single pass at R-MAT s20 ef16 through the
closed-form column checksums and the exact per-column entry counts."""
from __future__ import annotations

import numpy as np
import pytest

import combblas_amd as cb
import oracle
from combblas_amd import predictor
from combblas_amd.checks import product_matches
from tests.helpers import bit_exact

pytestmark = pytest.mark.gpu


def test_galerkin_256_full_size_exact():
    rng = np.random.default_rng(0xC5)
    grids = [int(g) for g in rng.integers(16, 97, 256)]
    As, Ts = [], []
    for g in grids:
        As.append(cb.gen_stencil27(g).to_device())
        Ts.append(cb.gen_aggregation(g, True).to_device())
    SATs, st = cb.galerkin_batched(As, Ts)
    for X in As + Ts:
        X.release()
    for g, S in zip(grids, SATs):
        bit_exact(S.to_host(), oracle.galerkin_direct(g))  # integers in FP64: exact in any order
        S.release()


def mcl_sample_check(mode: str) -> dict:
    """The MCL expansion step at n = 10^6 on 48 sampled columns of B, against
    the FP64 oracle: every kept entry within gamma_k of the FP64 product, and
    the kept set equal to the oracle's except for entries whose FP64 value
    lies within 2 gamma_k of their column's cut."""
    A = cb.gen_powerlaw_stochastic(1_000_000)
    rng = np.random.default_rng(11)
    qs = np.sort(rng.choice(A.nzc, 48, replace=False))
    idx = np.concatenate([np.arange(A.cp[q], A.cp[q + 1]) for q in qs])
    Bs = cb.Dcsc.from_coo(A.m, A.n, A.ir[idx], np.repeat(A.jc[qs], np.diff(A.cp)[qs]), A.val[idx], "float32")
    params = (1e-4, 1000, 0, 0.9)
    dA = A.to_device()
    P, st = cb.mcl_expand(dA, Bs.to_device(), *params, mode=mode)
    G = P.to_host()
    C64, _ = oracle.spgemm(A.astype("float64"), Bs.astype("float64"))
    W, cut = oracle.mcl_prune(C64, *params, return_cut=True)
    flop, _ = oracle.symbolic(A, Bs)
    gam_col = flop * 2.0 ** -24 / (1 - flop * 2.0 ** -24)
    assert np.array_equal(C64.jc, Bs.jc)
    gam = np.repeat(gam_col, np.diff(C64.cp))
    cut_e = np.repeat(cut, np.diff(C64.cp))
    k64 = C64.col_ids() * C64.m + C64.ir
    kg = G.col_ids() * G.m + G.ir
    i = np.searchsorted(k64, kg)
    assert np.all(k64[i] == kg)
    assert np.all(np.abs(G.val.astype(np.float64) - C64.val[i]) <= gam[i] * C64.val[i])
    diff = np.setxor1d(kg, W.col_ids() * W.m + W.ir)
    j = np.searchsorted(k64, diff)
    assert np.all(np.abs(C64.val[j] - cut_e[j]) <= 2 * gam[j] * np.maximum(C64.val[j], cut_e[j]))
    assert st["selected_cols"] > 0
    return st


def test_mcl_expansion_1m_sample_fp32_bound():
    mcl_sample_check("two_phase")


@pytest.mark.skipif(not predictor.MODEL_PATH.exists(), reason="no trained predictor in combblas_amd/models")
def test_mcl_expansion_1m_learned_sizing_fp32_bound():
    """The same step with learned nnz sizing (the shipped R-MAT-trained
    predictor on the power-law input): the result obeys the same bound
    whether the predictions fit or the capacity check falls back."""
    cb.set_predictor(predictor.load())
    try:
        st = mcl_sample_check("learned")
    finally:
        cb.set_predictor(None)
    assert st["phases"] >= 1


@pytest.mark.skipif(not predictor.MODEL_PATH.exists(), reason="no trained predictor in combblas_amd/models")
def test_learned_s20_full_size_checksums():
    """Learned single pass at R-MAT s20 ef16 with the stored tunables of that
    workload: device structure check, every column's checksum against its
    closed form, and every column's entry count against the exact symbolic
    pass (a zero-valued spurious or dropped entry changes the count)."""
    from combblas_amd import autotune
    A = cb.gen_rmat(20, 16)
    dA = A.to_device()
    counts = cb.spgemm_symbolic(dA, dA)
    base = cb.get_config()
    cb.set_predictor(predictor.load())
    try:
        autotune.apply("rmat_s20_ef16_learned")
        C, st = cb.spgemm(dA, dA, "plus_times", "learned")
    finally:
        cb.set_predictor(None)
        cb.set_config(**base)
    assert st["learned_fallback"] == 0
    assert product_matches(C, A, A, col_nnz=counts)
    C.release()
