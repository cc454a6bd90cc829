"""Galerkin operands and the coarse operator (ReleaseTests/GalerkinNew.cpp):
the generators' structure, the direct aggregation sum against the SpGEMM
restatement S*(A*T) and the genuine reference, and the reference's own check,
the splitting identity S*L*T + S*D*T == S*A*T."""
import numpy as np
import pytest

import combblas_amd as cb
import oracle
from helpers import bit_exact


@pytest.mark.parametrize("g", [1, 2, 3, 5, 8])
def test_stencil_and_aggregation_structure(g):
    A = cb.gen_stencil27(g)
    n = g ** 3
    assert A.m == A.n == n and A.dtype == "float64"
    d = A.to_dense()
    assert np.array_equal(d, d.T)
    assert np.all(np.diag(d) == 26.0)
    per_dim = np.array([1 + (x > 0) + (x < g - 1) for x in range(g)])
    assert A.nnz == int(per_dim.sum() ** 3)
    T = cb.gen_aggregation(g, True)
    R = cb.gen_aggregation(g, False)
    gc = (g + 1) // 2
    assert (T.m, T.n, R.m, R.n) == (n, gc ** 3, gc ** 3, n)
    bit_exact(oracle.transpose(T), R)
    assert np.all(np.add.reduceat(T.val, T.cp[:-1]) == np.diff(T.cp))  # aggregate sizes
    assert np.all(np.diff(R.cp) == 1)


@pytest.mark.parametrize("g", [1, 2, 3, 4, 5, 7, 10])
def test_direct_equals_spgemm_chain(g):
    A, T = cb.gen_stencil27(g), cb.gen_aggregation(g, True)
    S = oracle.transpose(T)
    AT, _ = oracle.spgemm(A, T)
    SAT, _ = oracle.spgemm(S, AT)
    bit_exact(SAT, oracle.galerkin_direct(g))


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (needs the reference tree)")
@pytest.mark.parametrize("g", [3, 5, 8])
def test_direct_equals_reference_chain(g):
    A, T = cb.gen_stencil27(g), cb.gen_aggregation(g, True)
    S = oracle.transpose(T)
    SAT = oracle.ref_spgemm(S, oracle.ref_spgemm(A, T))
    bit_exact(SAT, oracle.galerkin_direct(g))


def test_splitting_identity():
    """GalerkinNew.cpp:105-118: with A = L + D, S*L*T + S*D*T equals S*A*T."""
    g = 6
    A, T = cb.gen_stencil27(g), cb.gen_aggregation(g, True)
    S = oracle.transpose(T)
    cols = A.col_ids()
    off = A.ir != cols
    L = cb.Dcsc.from_coo(A.m, A.n, A.ir[off], cols[off], A.val[off], "float64")
    dvec = np.zeros(A.n)
    dvec[cols[~off]] = A.val[~off]
    SD = cb.Dcsc(S.m, S.n, S.cp, S.jc, S.ir, S.val * dvec[S.col_ids()], "float64")  # DimApply(Column, dvec, *)
    SLT, _ = oracle.spgemm(S, oracle.spgemm(L, T)[0])
    SDT, _ = oracle.spgemm(SD, T)
    total = SLT.to_dense() + SDT.to_dense()
    assert np.array_equal(total, oracle.galerkin_direct(g).to_dense())
