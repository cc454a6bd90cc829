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


def test_batched_operands_checked_against_their_own_bounds():
    """Each batch member is checked against its own shape, not only the packed
    block-diagonal matrix: a row equal to the member's m (a valid row of the
    packed matrix, in the next member's block), a column id equal to its n,
    rows descending inside a column, and a cp that does not close at nnz are
    rejected, and the valid batch still gives the single products."""
    rng = np.random.default_rng(9)
    As = [random_dcsc(rng, 50, 40, 0.1, lo=1, hi=9) for _ in range(3)]
    Bs = [random_dcsc(rng, 40, 30, 0.1, lo=1, hi=9) for _ in range(3)]
    Cs, _ = cb.spgemm_batched([A.to_device() for A in As], [B.to_device() for B in Bs])
    for A, B, Cd in zip(As, Bs, Cs):
        bit_exact(Cd.to_host(), oracle.spgemm(A, B)[0])
    X = As[1]
    c = int(np.argmax(np.diff(X.cp) >= 2))

    def variant(edit):
        Y = cb.Dcsc(X.m, X.n, X.cp.copy(), X.jc.copy(), X.ir.copy(), X.val.copy())
        edit(Y)
        return Y

    def row_at_m(Y):
        Y.ir[-1] = Y.m

    def column_at_n(Y):
        Y.jc[-1] = Y.n

    def rows_descend(Y):
        b = int(Y.cp[c])
        Y.ir[b], Y.ir[b + 1] = Y.ir[b + 1], Y.ir[b]

    def cp_short(Y):
        Y.cp[-1] -= 1

    for edit in (row_at_m, column_at_n, rows_descend, cp_short):
        bad = variant(edit)
        with pytest.raises(cb.CbxError):
            cb.spgemm_batched([As[0].to_device(), bad.to_device(), As[2].to_device()],
                              [B.to_device() for B in Bs])
        Ts = [random_dcsc(rng, 50, 12, 0.2, lo=1, hi=3) for _ in range(3)]
        with pytest.raises(cb.CbxError):
            cb.galerkin_batched([As[0].to_device(), bad.to_device(), As[2].to_device()],
                                [T.to_device() for T in Ts])
