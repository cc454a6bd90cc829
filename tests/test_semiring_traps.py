"""The semiring corner cases the reference's own definitions fix
(include/CombBLAS/Semirings.h), on the oracle, the genuine reference
(oracle/_ref) and the device:

* MinPlusSRing's multiply is inf_plus (Semirings.h:40-47): an operand equal
  to numeric_limits<T>::max() makes the product max() itself, whatever the
  other operand, and min() over products then keeps max() only where every
  product saturated.  Exercised with int64, float64 and float32 operands of
  which about a quarter are max().
* SelectMaxSRing's multiply is a*b (Semirings.h:176-182) and its add is max;
  its id() is -1 (:168), which is not an identity of max for products below
  -1.  The reference never seeds an accumulator with id() in the local
  product (first insertion, then add), so a column whose products are all
  below -1 keeps its largest product.  Exercised with every product in
  [-81, -2]: every output must be below -1.
"""
from __future__ import annotations

import numpy as np
import pytest

import combblas_amd as cb
import oracle
from tests.helpers import bit_exact

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (no reference tree)")
TYPES = {"int64": np.int64, "float64": np.float64, "float32": np.float32}


def saturated(X: cb.Dcsc, dtype: str, rng, share: float = 0.25) -> cb.Dcsc:
    """Values uniform in [1, 9] with `share` of them set to numeric_limits<T>::max()."""
    t = TYPES[dtype]
    Y = X.astype(dtype)
    v = rng.integers(1, 10, X.nnz).astype(t)
    big = np.iinfo(t).max if dtype == "int64" else np.finfo(t).max
    v[rng.random(X.nnz) < share] = big
    Y.val = v
    return Y


def negative_products(A: cb.Dcsc, B: cb.Dcsc, dtype: str, rng) -> tuple[cb.Dcsc, cb.Dcsc]:
    """A's values in [-9, -2], B's in [1, 9]: every product lies in [-81, -2]."""
    t = TYPES[dtype]
    A2, B2 = A.astype(dtype), B.astype(dtype)
    A2.val = rng.integers(-9, -1, A.nnz).astype(t)
    B2.val = rng.integers(1, 10, B.nnz).astype(t)
    return A2, B2


def minplus_operands(dtype: str, scale: int = 10):
    rng = np.random.default_rng(41)
    A = saturated(cb.gen_rmat(scale, 8), dtype, rng)
    B = saturated(cb.gen_rmat(scale, 8, seed=77), dtype, rng)
    return A, B


def selectmax_operands(dtype: str, scale: int = 10):
    rng = np.random.default_rng(43)
    return negative_products(cb.gen_rmat(scale, 8), cb.gen_rmat(scale, 8, seed=78), dtype, rng)


def _max_of(dtype):
    t = TYPES[dtype]
    return np.iinfo(t).max if dtype == "int64" else np.finfo(t).max


@pytest.mark.parametrize("dtype", ["int64", "float64", "float32"])
def test_oracle_minplus_saturation_brute_force(dtype):
    """The oracle's min-plus product equals a brute-force dense min over
    inf_plus products, and holds max() exactly where every product saturated."""
    rng = np.random.default_rng(5)
    A = saturated(cb.gen_rmat(6, 4), dtype, rng, 0.3)
    B = saturated(cb.gen_rmat(6, 4, seed=9), dtype, rng, 0.3)
    C, _ = oracle.spgemm(A, B, "min_plus")
    big = _max_of(dtype)
    t = TYPES[dtype]
    Ad, Bd = A.to_dense(), B.to_dense()
    Am = np.zeros(Ad.shape, bool)
    Am[A.ir, A.col_ids()] = True
    Bm = np.zeros(Bd.shape, bool)
    Bm[B.ir, B.col_ids()] = True
    want = np.full((A.m, B.n), big, t)
    hit = np.zeros((A.m, B.n), bool)
    for k in range(A.n):
        rows, cols = np.nonzero(Am[:, k])[0], np.nonzero(Bm[k, :])[0]
        if not rows.size or not cols.size:
            continue
        a, b = Ad[rows, k][:, None], Bd[k, cols][None, :]
        with np.errstate(over="ignore"):
            p = np.where((a == big) | (b == big), big, a + b).astype(t)
        sub = want[np.ix_(rows, cols)]
        want[np.ix_(rows, cols)] = np.minimum(sub, p)
        hit[np.ix_(rows, cols)] = True
    got_hit = np.zeros((A.m, B.n), bool)
    got_hit[C.ir, C.col_ids()] = True
    assert np.array_equal(got_hit, hit)
    assert np.array_equal(C.val, want[C.ir, C.col_ids()])
    assert np.any(C.val == big) and np.any(C.val != big)


@needs_ref
@pytest.mark.parametrize("dtype", ["int64", "float64", "float32"])
@pytest.mark.parametrize("mode", ["hybrid", "hash", "hash_unsorted"])
def test_oracle_minplus_saturation_matches_reference(dtype, mode):
    A, B = minplus_operands(dtype)
    want = oracle.ref_spgemm(A, B, "min_plus", mode)
    got, _ = oracle.spgemm(A, B, "min_plus", mode)
    bit_exact(got, want)
    assert np.any(want.val == _max_of(dtype))


@needs_ref
@pytest.mark.parametrize("dtype", ["int64", "float64", "float32"])
@pytest.mark.parametrize("mode", ["hybrid", "hash"])
def test_oracle_selectmax_below_minus_one_matches_reference(dtype, mode):
    A, B = selectmax_operands(dtype)
    want = oracle.ref_spgemm(A, B, "select_max", mode)
    got, _ = oracle.spgemm(A, B, "select_max", mode)
    bit_exact(got, want)
    assert want.nnz > 0 and np.all(want.val <= -2)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["int64", "float64", "float32"])
@pytest.mark.parametrize("mode", ["two_phase", "unsorted"])
def test_gpu_minplus_saturation_bit_exact(dtype, mode):
    """s12 operands: columns reach the tiny, hash and unit tiers."""
    A, B = minplus_operands(dtype, 12)
    dC, _ = cb.spgemm(A.to_device(), B.to_device(), "min_plus", mode)
    got = dC.to_host()
    want, _ = oracle.spgemm(A, B, "min_plus", "hybrid")
    if mode == "unsorted":
        from tests.helpers import sort_columns
        got = sort_columns(got)
    bit_exact(got, want)
    assert np.any(got.val == _max_of(dtype))


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["int64", "float64", "float32"])
def test_gpu_selectmax_below_minus_one_bit_exact(dtype):
    A, B = selectmax_operands(dtype, 12)
    dC, _ = cb.spgemm(A.to_device(), B.to_device(), "select_max")
    got = dC.to_host()
    want, _ = oracle.spgemm(A, B, "select_max")
    bit_exact(got, want)
    assert got.nnz > 0 and np.all(got.val <= -2)
