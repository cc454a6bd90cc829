"""MCL expansion on the GPU (cbx_mcl_prune, cbx_mcl_expand) against the
oracle: the epilogue bit-exact on identical inputs, the phased product plus
epilogue bit-exact where FP32 arithmetic is exact (dyadic values), and on the
power-law workload within the FP32 error bound of an FP64 oracle."""
from __future__ import annotations

import numpy as np
import pytest

import combblas_amd as cb
import oracle
from tests.helpers import MCL_PARAMS, bit_exact, mcl_columns

pytestmark = pytest.mark.gpu


def float_sum_decisions_agree(A: cb.Dcsc, params) -> bool:
    """The device sums column statistics in FP64, the reference sequentially in
    the value type (SpParMat.cpp:961-1010); the recovery tests `sum < pct`
    must come out the same both ways for the comparison to be exact."""
    thr, select, recover, pct = params
    T = np.float32 if A.dtype == "float32" else np.float64
    p = T(pct)
    for q in range(A.nzc):
        v = A.val[A.cp[q]:A.cp[q + 1]]
        big = v[v > T(thr)]
        seq = np.cumsum(big, dtype=T)[-1] if big.size else T(0)  # sequential, in the value type
        if (seq < p) != (float(big.astype(np.float64).sum()) < float(p)):
            return False
    return True


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("params", MCL_PARAMS)
def test_mcl_prune_bit_exact(dtype, params):
    for seed in (1, 2):
        A = mcl_columns(seed, dtype=dtype)
        assert float_sum_decisions_agree(A, params)
        got, st = cb.mcl_prune(A.to_device(), *params)
        want = oracle.mcl_prune(A, *params)
        bit_exact(got.to_host(), want)
        assert st["nnz"] == want.nnz and st["nnz_product"] == A.nnz


def test_mcl_prune_edge_cases():
    E = cb.Dcsc(50, 40, np.zeros(1, np.int64), np.zeros(0, np.int64), np.zeros(0, np.int64),
                np.zeros(0, np.float32), "float32")
    got, st = cb.mcl_prune(E.to_device())
    assert got.to_host().nnz == 0 and st["nnz"] == 0
    A = mcl_columns(3)
    got, _ = cb.mcl_prune(A.to_device(), thr=10.0, select=5, recover=0)  # every entry below the threshold
    assert got.to_host().nnz == 0
    got, _ = cb.mcl_prune(A.to_device(), thr=-1.0, select=0, recover=0)  # nothing pruned
    bit_exact(got.to_host(), A)
    with pytest.raises(cb.CbxError):
        cb.mcl_prune(cb.gen_rmat(8, 4).to_device())  # int64 values


def dyadic(n: int, seed: int) -> cb.Dcsc:
    """Values k/1024, k in 1..16: every product and sum of up to 4096 terms is
    exact in FP32, so results do not depend on accumulation order."""
    A = cb.gen_powerlaw_stochastic(n, mean=12.0, dmax=200, seed=seed)
    rng = np.random.default_rng(seed)
    A.val = (rng.integers(1, 17, A.nnz) / 1024.0).astype(np.float32)
    return A


@pytest.mark.parametrize("phases", [1, 3])
@pytest.mark.parametrize("params", [(2.0 ** -14, 40, 0, 0.9), (2.0 ** -12, 30, 60, 0.05), (2.0 ** -16, 0, 0, 0.9)])
def test_mcl_expand_exact(phases, params):
    A = dyadic(6000, 11)
    C, _ = oracle.spgemm(A, A)
    want = oracle.mcl_prune(C, *params)
    dA = A.to_device()
    got, st = cb.mcl_expand(dA, dA, *params, phases=phases)
    bit_exact(got.to_host(), want)
    assert st["phases"] == phases and st["nnz_product"] == C.nnz and st["nnz"] == want.nnz


@pytest.mark.parametrize("tiers", ["default", "units"])
@pytest.mark.parametrize("tiny_algo", [0, 1])
def test_mcl_fused_threshold_matches_unfused(tiers, tiny_algo):
    """With recover_num 0 the two-phase expansion fuses the hard threshold into
    the product's emission (mcl_fuse 1): the kernels write only entries with
    !(v < threshold), compacted per output range, and the ranges are gathered.
    On dyadic values (every sum exact in FP32, so the LDS atomics' order does
    not show) the result equals the unfused expansion's bit for bit and the
    oracle's, for a sparse power-law matrix and a denser one (mean 50 entries
    per column, at most 500), with the unit tier forced for most columns or not, in both
    tiny-tier forms, in one and in three phases."""
    A = dyadic(6000, 13)
    B = cb.gen_powerlaw_stochastic(5000, seed=6)
    # k/64, k in 1..4: products are multiples of 2^-12, at most 2^-8, and a
    # column's at most 500 * 500 of them sum below 2^10: exact in 22 bits
    B.val = (np.random.default_rng(6).integers(1, 5, B.nnz) / 64.0).astype(np.float32)
    params = (2.0 ** -14, 40, 0, 0.9)
    base = cb.get_config()
    extra = {"tiny_flop": 128, "hash_mid_flop": 256, "hash_max_flop": 256} if tiers == "units" else {}
    try:
        for X in (A, B):
            dX = X.to_device()
            out = {}
            for fuse in (0, 1):
                cb.set_config(**dict(base, mcl_fuse=fuse, tiny_algo=tiny_algo, **extra))
                for phases in (1, 3):
                    got, st = cb.mcl_expand(dX, dX, *params, phases=phases)
                    out[fuse, phases] = (got.to_host(), st)
            for key in out:
                bit_exact(out[key][0], out[0, 1][0])
                assert out[key][1]["nnz_product"] == out[0, 1][1]["nnz_product"]
                assert out[key][1]["nnz"] == out[0, 1][1]["nnz"]
            C, _ = oracle.spgemm(X, X)
            bit_exact(out[1, 1][0], oracle.mcl_prune(C, *params))
            assert out[1, 3][1]["nnz_product"] == C.nnz
    finally:
        cb.set_config(**base)


def test_mcl_fused_threshold_on_row_list_units():
    """The fused threshold through the row-list units' emission: a uniform
    random 300,007-column matrix (12 entries per column, dyadic values) with
    the unit tier forced for every column, so that each column's ~140
    outputs over 74 row blocks form one list unit.  Lists on and off, fused
    and not: bit for bit the same, and the oracle's."""
    rng = np.random.default_rng(8)
    n = 300_007
    key = np.unique(rng.integers(0, n, 12 * n) * n + rng.integers(0, n, 12 * n))
    vals = (rng.integers(1, 17, key.size) / 1024.0).astype(np.float32)  # <= 900 terms of <= 2^-12: exact
    A = cb.Dcsc.from_coo(n, n, key % n, key // n, vals, "float32")
    params = (2.0 ** -14, 40, 0, 0.9)
    base = cb.get_config()
    dA = A.to_device()
    out = {}
    try:
        for lists in (0, 1):
            for fuse in (0, 1):
                cb.set_config(**dict(base, tiny_flop=64, hash_mid_flop=64, hash_max_flop=64, unit_lists=lists,
                                     mcl_fuse=fuse))
                got, st = cb.mcl_expand(dA, dA, *params, phases=2)
                out[lists, fuse] = (got.to_host(), st)
    finally:
        cb.set_config(**base)
    C, _ = oracle.spgemm(A, A)
    want = oracle.mcl_prune(C, *params)
    for key_, (got, st) in out.items():
        bit_exact(got, want)
        assert st["nnz_product"] == C.nnz, key_


def test_mcl_expand_learned_exact():
    from tests.test_learned_gpu import constant_predictor
    A = dyadic(6000, 12)
    C, _ = oracle.spgemm(A, A)
    want = oracle.mcl_prune(C, 2.0 ** -14, 40, 0, 0.9)
    cb.set_predictor(constant_predictor(10.0))
    try:
        dA = A.to_device()
        got, st = cb.mcl_expand(dA, dA, 2.0 ** -14, 40, 0, 0.9, phases=2, mode="learned")
    finally:
        cb.set_predictor(None)
    bit_exact(got.to_host(), want)


def test_mcl_expand_learned_powerlaw_network():
    """The learned expansion sized by the committed power-law network (the
    one the MCL benchmark installs) equals the oracle's pruned product, in
    one and in two phases."""
    from combblas_amd import predictor
    if not predictor.model_path("powerlaw").exists():
        pytest.skip("power-law network not trained (python -m combblas_amd.predictor train powerlaw)")
    A = dyadic(6000, 14)
    C, _ = oracle.spgemm(A, A)
    params = (2.0 ** -14, 40, 0, 0.9)
    want = oracle.mcl_prune(C, *params)
    cb.set_predictor(predictor.load("powerlaw"))
    try:
        dA = A.to_device()
        for phases in (1, 2):
            got, st = cb.mcl_expand(dA, dA, *params, phases=phases, mode="learned")
            bit_exact(got.to_host(), want)
            assert st["nnz_product"] == C.nnz
    finally:
        cb.set_predictor(None)


def _keys(X: cb.Dcsc) -> np.ndarray:
    return X.col_ids().astype(np.int64) * X.m + X.ir


def test_mcl_expand_powerlaw_fp32_bound():
    """FP32 expansion of a column-stochastic power-law matrix against the FP64
    oracle.  Every kept value is within Higham's bound |c^ - c| <= gamma_k c
    (all terms positive, gamma_k = k u / (1 - k u), u = 2^-24, k = terms of
    the column); an entry kept by only one side lies within that bound of the
    column's cut."""
    A = cb.gen_powerlaw_stochastic(5000, seed=5)
    params = (1e-4, 100, 0, 0.9)
    dA = A.to_device()
    got, st = cb.mcl_expand(dA, dA, *params, phases=2)
    G = got.to_host()
    A64 = A.astype("float64")
    C64, _ = oracle.spgemm(A64, A64)
    W, cut = oracle.mcl_prune(C64, *params, return_cut=True)
    flop, _ = oracle.symbolic(A, A)
    assert np.array_equal(C64.jc, A.jc)  # every column of A is stored, so C's columns line up with flop
    u = 2.0 ** -24
    gam_col = np.minimum(flop * u / (1 - flop * u), 1.0)
    gam = np.repeat(gam_col, np.diff(C64.cp))
    cut_e = np.repeat(cut, np.diff(C64.cp))
    k64 = _keys(C64)
    kg = _keys(G)
    idx = np.searchsorted(k64, kg)
    assert np.all(k64[idx] == kg), "GPU kept an entry that is not in the product"
    c = C64.val[idx]
    assert np.all(np.abs(G.val.astype(np.float64) - c) <= gam[idx] * c)
    diff = np.setxor1d(kg, _keys(W))
    j = np.searchsorted(k64, diff)
    cd = C64.val[j]
    assert np.all(np.abs(cd - cut_e[j]) <= 2 * gam[j] * np.maximum(cd, cut_e[j]))
    assert diff.size <= 1e-3 * max(W.nnz, 1)
    assert st["nnz"] == G.nnz and st["flops"] == int(flop.sum())
