"""Dense-vector SpMV (dcsc_gespmv, Friends.h:63-77): the restatement against
the genuine template and against a dense brute force."""
import numpy as np
import pytest

import combblas_amd as cb
import oracle
from helpers import random_dcsc


def brute(A, x, y, semiring):
    d = A.to_dense().astype(np.float64 if x.dtype != np.int64 else np.int64)
    mask = np.zeros((A.m, A.n), bool)
    mask[A.ir, A.col_ids()] = True
    out = y.copy()
    for i in range(A.m):
        for j in np.nonzero(mask[i])[0]:
            a = x[j] if A.dtype == "bool" else d[i, j]
            if semiring == "plus_times":
                out[i] = out[i] + (a * x[j] if A.dtype != "bool" else a)
            elif semiring == "min_plus":
                out[i] = min(out[i], a + x[j])
            else:
                out[i] = max(out[i], a * x[j] if A.dtype != "bool" else a)
    return out


@pytest.mark.parametrize("semiring", ["plus_times", "min_plus", "select_max"])
def test_spmv_dense_oracle_vs_brute(semiring):
    rng = np.random.default_rng(3)
    A = random_dcsc(rng, 40, 30, 0.2, lo=-5, hi=9)
    x = rng.integers(-7, 8, A.n).astype(np.int64)
    y = rng.integers(-3, 4, A.m).astype(np.int64)
    assert np.array_equal(oracle.spmv_dense(A, x, y, semiring), brute(A, x, y, semiring))


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (needs the reference tree)")
@pytest.mark.parametrize("case", ["plus_times-int64", "min_plus-int64", "select_max-int64", "plus_times-float64",
                                  "select_max-bool"])
def test_spmv_dense_oracle_vs_reference(case):
    semiring, dt = case.split("-")
    rng = np.random.default_rng(11)
    A = cb.gen_rmat(12, 8, seed=5)
    if dt == "bool":
        A = A.astype("bool")
    elif dt == "float64":
        A = A.astype("float64")
        A.val = rng.standard_normal(A.nnz)
    else:
        A.val = rng.integers(-9, 10, A.nnz).astype(np.int64)
    xd = np.float64 if dt == "float64" else np.int64
    x = (rng.standard_normal(A.n) if xd == np.float64 else rng.integers(-20, 20, A.n)).astype(xd)
    y = (rng.standard_normal(A.m) if xd == np.float64 else rng.integers(-5, 5, A.m)).astype(xd)
    got = oracle.spmv_dense(A, x, y, semiring)
    want = oracle.ref_spmv_dense(A, x, y, semiring)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
