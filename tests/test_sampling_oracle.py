"""Pins the oracle's restatement of estimateNNZ_sampling
(include/CombBLAS/mtSpGEMM.h:936-1054) before it judges the device:

* bit for bit against the genuine template (oracle/_ref): the int64
  estimate of whole products, and of single columns (where the reference's
  sum is that column's estimate alone);
* the sample stream against its definition: minstd_rand0 from seed 1
  (x <- 16807 x mod 2^31-1), u = float(x - 1) / 2^31, -log(1 - u);
* the two-level minimum against the product's pattern: for every column j,
  min over B(:,j)'s rows of (min over A's column rows) equals the minimum
  over the rows of C(:,j) = A*B, so each estimate is
  (r - 1) / sum_k min_{i in rows(C(:,j))} sample[i][k] recomputed from the
  oracle's own product.
"""
from __future__ import annotations

import numpy as np
import pytest

import combblas_amd as cb
import oracle

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (no reference tree)")


@pytest.fixture
def one_thread():
    n = oracle.max_threads()
    oracle.set_threads(1)  # the reference's float reduction is sequential on one thread
    yield
    oracle.set_threads(n)


@needs_ref
@pytest.mark.parametrize("scale,ef,rounds", [(8, 8, 5), (10, 8, 5), (11, 16, 3), (10, 4, 8)])
def test_estimate_matches_reference(scale, ef, rounds, one_thread):
    A = cb.gen_rmat(scale, ef)
    B = cb.gen_rmat(scale, ef, seed=123)
    _, _, est = oracle.estimate_sampling(A, B, rounds)
    assert est == oracle.ref_estimate_sampling(A, B, rounds)
    _, _, est2 = oracle.estimate_sampling(A, A, rounds)
    assert est2 == oracle.ref_estimate_sampling(A, A, rounds)


@needs_ref
def test_per_column_estimates_match_reference(one_thread):
    A = cb.gen_rmat(11, 8)
    col, _, _ = oracle.estimate_sampling(A, A, 5)
    rng = np.random.default_rng(3)
    for q in rng.choice(A.nzc, 24, replace=False):
        Bq = oracle.column_slice(A, int(q), int(q) + 1)
        assert oracle.ref_estimate_sampling(A, Bq, 5) == int(np.float32(col[q]))


def test_sample_stream_definition():
    n = 4000
    got = oracle.exp_samples(n)
    x = np.zeros(n, np.int64)
    v = 1
    for i in range(n):
        v = (v * 16807) % 2147483647
        x[i] = v
    u = (x - 1).astype(np.float32) / np.float32(2.0 ** 31)
    want = -np.log(np.float32(1.0) - u)
    assert np.allclose(got, want, rtol=4e-7, atol=0)  # a few units in the last place: two logf implementations
    assert abs(float(got.mean()) - 1.0) < 0.05


@pytest.mark.parametrize("rounds", [3, 5])
def test_two_level_minimum_equals_product_pattern_minimum(rounds):
    A = cb.gen_rmat(9, 8)
    B = cb.gen_rmat(9, 4, seed=77)
    col, tot, est = oracle.estimate_sampling(A, B, rounds)
    S = oracle.exp_samples(A.m * rounds).reshape(A.m, rounds)
    Cp, _ = oracle.spgemm(A.astype("bool"), B.astype("bool"))
    pos = np.searchsorted(B.jc, Cp.jc)
    want = np.zeros(B.nzc, np.float32)  # columns of B with an empty product estimate 0
    for c in range(Cp.nzc):
        fin = S[Cp.ir[Cp.cp[c]:Cp.cp[c + 1]]].min(axis=0)
        acc = np.float32(0)
        for k in range(rounds):
            acc = np.float32(acc + fin[k])
        want[pos[c]] = np.float32(rounds - 1) / acc
    assert np.array_equal(col, want)
    assert est == int(tot)
