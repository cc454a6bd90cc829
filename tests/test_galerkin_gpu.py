"""Device transpose, batched products and the batched Galerkin triple product
against the oracle, bit-exact (integer-valued FP64 is exact in any order)."""
from __future__ import annotations

import numpy as np
import pytest

import combblas_amd as cb
import oracle
from tests.helpers import bit_exact, random_dcsc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", ["int64", "float64", "float32", "bool"])
def test_transpose(dtype):
    rng = np.random.default_rng(4)
    for m, n, dens in [(70, 50, 0.1), (1, 300, 0.5), (300, 1, 0.5), (40, 40, 0.0)]:
        X = random_dcsc(rng, m, n, dens, "int64" if dtype == "bool" else dtype)
        if dtype == "bool":
            X = X.astype("bool")
        got = cb.transpose(X.to_device()).to_host()
        bit_exact(got, oracle.transpose(X))
    A = cb.gen_rmat(14, 8)
    bit_exact(cb.transpose(A.to_device()).to_host(), oracle.transpose(A))


@pytest.mark.parametrize("semiring", ["plus_times", "min_plus", "select_max"])
def test_spgemm_batched_matches_single(semiring):
    rng = np.random.default_rng(5)
    As, Bs = [], []
    for i, (m, k, n) in enumerate([(60, 40, 30), (1, 1, 1), (200, 150, 120), (30, 30, 0), (500, 80, 700)]):
        As.append(random_dcsc(rng, m, k, 0.08, lo=-9, hi=9))
        Bs.append(random_dcsc(rng, k, n, 0.08, lo=-9, hi=9))
    As.append(cb.gen_rmat(12, 8))
    Bs.append(cb.gen_rmat(12, 8, seed=3))
    Cs, st = cb.spgemm_batched([A.to_device() for A in As], [B.to_device() for B in Bs], semiring)
    for A, B, Cd in zip(As, Bs, Cs):
        want, _ = oracle.spgemm(A, B, semiring)
        bit_exact(Cd.to_host(), want)
    assert st["nnz"] == sum(C.nnz for C in Cs)


def test_galerkin_batched_small_grids():
    grids = [1, 2, 3, 4, 5, 6, 7, 9, 12, 16, 17]
    As = [cb.gen_stencil27(g).to_device() for g in grids]
    Ts = [cb.gen_aggregation(g, True).to_device() for g in grids]
    SATs, st = cb.galerkin_batched(As, Ts)
    for g, S in zip(grids, SATs):
        bit_exact(S.to_host(), oracle.galerkin_direct(g))


def test_galerkin_batched_256():
    """256 grids in one call (edges 8..40, seeded), every product bit-exact."""
    rng = np.random.default_rng(256)
    grids = [int(g) for g in rng.integers(8, 41, 256)]
    As = [cb.gen_stencil27(g).to_device() for g in grids]
    Ts = [cb.gen_aggregation(g, True).to_device() for g in grids]
    SATs, st = cb.galerkin_batched(As, Ts)
    for g, S in zip(grids, SATs):
        bit_exact(S.to_host(), oracle.galerkin_direct(g))
