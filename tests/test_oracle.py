"""Pins the CPU oracle (oracle/cbx_oracle.cpp) before it is trusted.

* against the genuine CombBLAS kernels compiled from the reference tree
  (oracle/_ref) for every semiring, value type and mode, bit for bit;
* against brute-force dense semiring products on tiny inputs (pattern, and
  exact values for integer / boolean / min-plus / select-max);
* against closed forms: flop_j = sum of A-column lengths, nnz(C) = sum nnz_j,
  and sum(C) = sum_k colsum_k(A) * rowsum_k(B) for plus-times;
* against the committed golden fixtures (written by tests/golden/make_golden.py
  from the genuine reference).
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import combblas_amd as cb
import oracle
from tests.helpers import bit_exact, random_dcsc, same_pattern

GOLDEN = Path(__file__).resolve().parent / "golden"

SR_DTYPES = [("plus_times", "int64"), ("plus_times", "float64"), ("plus_times", "float32"),
             ("plus_times", "bool"), ("min_plus", "int64"), ("min_plus", "float64"), ("min_plus", "float32"),
             ("select_max", "int64"), ("select_max", "float64"), ("select_max", "float32")]


def _signed_values(X: cb.Dcsc, rng) -> cb.Dcsc:
    Y = X.astype(X.dtype)
    if X.dtype.startswith("float"):
        Y.val = rng.standard_normal(X.nnz).astype(Y.val.dtype)
    elif X.dtype == "int64":
        Y.val = rng.integers(-9, 10, X.nnz).astype(np.int64)
    return Y


needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (no reference tree)")


@needs_ref
@pytest.mark.parametrize("semiring,dtype", SR_DTYPES)
@pytest.mark.parametrize("mode", ["hybrid", "hash", "hash_unsorted"])
def test_oracle_matches_reference_rmat(semiring, dtype, mode):
    rng = np.random.default_rng(7)
    A = _signed_values(cb.gen_rmat(9, 8).astype(dtype), rng)
    B = _signed_values(cb.gen_rmat(9, 4, seed=99).astype(dtype), rng)
    R = oracle.ref_spgemm(A, B, semiring, mode)
    O, _ = oracle.spgemm(A, B, semiring, mode)
    bit_exact(O, R)


@needs_ref
@pytest.mark.parametrize("seed", range(4))
def test_oracle_matches_reference_rectangular(seed):
    rng = np.random.default_rng(seed)
    A = random_dcsc(rng, 70, 45, 0.08, "float64", -3, 3)
    B = random_dcsc(rng, 45, 33, 0.15, "float64", -3, 3)
    for mode in ("hybrid", "hash", "hash_unsorted"):
        bit_exact(oracle.spgemm(A, B, "plus_times", mode)[0], oracle.ref_spgemm(A, B, "plus_times", mode))


@needs_ref
def test_oracle_symbolic_matches_reference():
    A = cb.gen_rmat(12, 8)
    f1, n1 = oracle.ref_symbolic(A, A)
    f2, n2 = oracle.symbolic(A, A)
    assert np.array_equal(f1, f2) and np.array_equal(n1, n2)


def _dense_semiring(Ad, Am, Bd, Bm, semiring, dtype):
    m, k = Ad.shape
    n = Bd.shape[1]
    P = (Am.astype(np.int64) @ Bm.astype(np.int64)) > 0
    if semiring == "plus_times" and dtype == "bool":
        V = P & ((Ad.astype(np.int64) * Am) @ (Bd.astype(np.int64) * Bm) > 0)
        return P, V.astype(np.uint8)
    if semiring == "plus_times":
        return P, (Ad * Am) @ (Bd * Bm)
    out = np.zeros((m, n), Ad.dtype)
    for i in range(m):
        for j in range(n):
            ks = np.nonzero(Am[i] & Bm[:, j])[0]
            if ks.size == 0:
                continue
            if semiring == "min_plus":
                inf = np.iinfo(Ad.dtype).max if Ad.dtype.kind == "i" else np.finfo(Ad.dtype).max
                terms = [inf if (Ad[i, k] == inf or Bd[k, j] == inf) else Ad[i, k] + Bd[k, j] for k in ks]
                out[i, j] = min(terms)
            else:
                out[i, j] = max(Ad[i, k] * Bd[k, j] for k in ks)
    return P, out


@pytest.mark.parametrize("semiring,dtype", [s for s in SR_DTYPES if s[1] in ("int64", "bool")])
@pytest.mark.parametrize("seed", range(3))
def test_oracle_matches_bruteforce(semiring, dtype, seed):
    rng = np.random.default_rng(100 + seed)
    m, k, n = 40, 30, 35
    Am = rng.random((m, k)) < 0.12
    Bm = rng.random((k, n)) < 0.12
    lo = -5 if semiring != "plus_times" or dtype != "bool" else 0
    Ad = rng.integers(lo, 9, (m, k)).astype(np.int64)
    Bd = rng.integers(lo, 9, (k, n)).astype(np.int64)
    if dtype == "bool":
        Ad, Bd = (Ad > 2).astype(np.uint8), (Bd > 2).astype(np.uint8)
    A = cb.Dcsc.from_dense(Ad, dtype, mask=Am)
    B = cb.Dcsc.from_dense(Bd, dtype, mask=Bm)
    P, V = _dense_semiring(Ad, Am, Bd, Bm, semiring, dtype)
    for mode in ("hybrid", "hash"):
        Cc, _ = oracle.spgemm(A, B, semiring, mode)
        Pc = np.zeros_like(P)
        Pc[Cc.ir, Cc.col_ids()] = True
        assert np.array_equal(Pc, P), "pattern (explicit zeros included) differs from the boolean product"
        assert np.array_equal(Cc.val, V[Cc.ir, Cc.col_ids()].astype(Cc.val.dtype))


def test_oracle_keeps_explicit_zeros():
    # A(:,0) and A(:,1) hit row 0 with +1 and -1: C(0,0) = 0 must stay stored
    A = cb.Dcsc.from_coo(2, 2, [0, 0], [0, 1], [1, -1])
    B = cb.Dcsc.from_coo(2, 1, [0, 1], [0, 0], [1, 1])
    for mode in ("hybrid", "hash"):
        Cc, _ = oracle.spgemm(A, B, "plus_times", mode)
        assert Cc.nnz == 1 and Cc.ir[0] == 0 and Cc.val[0] == 0


def test_oracle_closed_forms():
    A = cb.gen_rmat(12, 8)
    flop, nnz = oracle.symbolic(A, A)
    # flop_j = sum over k in A(:,j) of nnz(A(:,k))
    lens = np.zeros(A.n, np.int64)
    lens[A.jc] = np.diff(A.cp)
    per_entry = lens[A.ir]
    assert np.array_equal(flop, np.add.reduceat(per_entry, A.cp[:-1]))
    Cc, fl = oracle.spgemm(A, A)
    assert fl == flop.sum()
    assert Cc.nnz == nnz.sum()
    assert np.array_equal(np.diff(Cc.cp), nnz[nnz > 0])
    # sum(C) = sum_k colsum_k(A) * rowsum_k(A)
    colsum = np.zeros(A.n, np.int64)
    colsum[A.jc] = np.add.reduceat(A.val, A.cp[:-1])
    rowsum = np.bincount(A.ir, weights=A.val, minlength=A.m).astype(np.int64)
    assert int(Cc.val.sum()) == int((colsum * rowsum).sum())


def test_oracle_column_range_equals_full():
    A = cb.gen_rmat(11, 8)
    full, _ = oracle.spgemm(A, A)
    half = A.nzc // 2
    lo, _ = oracle.spgemm(A, A, cols=(0, half))
    hi, _ = oracle.spgemm(A, A, cols=(half, A.nzc))
    assert lo.nnz + hi.nnz == full.nnz
    assert np.array_equal(np.concatenate([lo.ir, hi.ir]), full.ir)
    assert np.array_equal(np.concatenate([lo.val, hi.val]), full.val)


def test_oracle_against_golden_c1():
    g = np.load(GOLDEN / "spgemm_rmat_s14_ef8_plus_times_int64.npz")
    A = cb.gen_rmat(14, 8)
    assert A.nnz == int(g["nnz_a"])
    Cc, fl = oracle.spgemm(A, A)
    assert fl == int(g["flops"]) and Cc.nnz == int(g["nnz_c"])
    assert np.array_equal(Cc.jc, g["jc"])
    assert np.array_equal(np.diff(Cc.cp), g["col_nnz"])
    assert np.array_equal(oracle.column_digests(Cc), g["col_digest"])


@pytest.mark.parametrize("name", sorted(p.name for p in GOLDEN.glob("spgemm_tiny_*.npz")))
def test_oracle_against_golden_tiny(name):
    g = np.load(GOLDEN / name)
    semiring, dtype, mode = str(g["semiring"]), str(g["dtype"]), str(g["mode"])
    A = cb.Dcsc(int(g["a_m"]), int(g["a_n"]), g["a_cp"], g["a_jc"], g["a_ir"], g["a_val"], dtype)
    B = cb.Dcsc(int(g["b_m"]), int(g["b_n"]), g["b_cp"], g["b_jc"], g["b_ir"], g["b_val"], dtype)
    want = cb.Dcsc(A.m, B.n, g["c_cp"], g["c_jc"], g["c_ir"], g["c_val"], dtype)
    got, _ = oracle.spgemm(A, B, semiring, mode)
    bit_exact(got, want)


def test_generator_deterministic_and_pinned():
    A1 = cb.gen_rmat(14, 8)
    A2 = cb.gen_rmat(14, 8)
    bit_exact(A1, A2)
    assert (A1.nnz, A1.nzc) == (120002, 9205)
    S = cb.gen_rmat(10, 16, symmetrize=True, no_self_loops=True, pattern=True)
    D = S.to_dense()
    assert np.array_equal(D, D.T) and not np.any(np.diag(D)) and set(np.unique(S.val)) == {1}
    same_pattern(S, S)
