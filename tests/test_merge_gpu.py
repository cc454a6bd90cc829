"""Device stage merge (cbx_dcsc_merge, MultiwayMerge<SR>) against the oracle:
integer, bool, min-plus and select-max bit-exact, FP within 1e-12 relative
(duplicates summed in another order); and the SUMMA identity on the device."""
from __future__ import annotations

import numpy as np
import pytest

import combblas_amd as cb
import oracle
from helpers import stage_slices
from test_merge_oracle import random_part, same

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("semiring,dtype", [("plus_times", "int64"), ("min_plus", "int64"), ("select_max", "int64"),
                                            ("plus_times", "bool"), ("plus_times", "float64"),
                                            ("min_plus", "float64"), ("select_max", "float64")])
def test_merge_matches_oracle(semiring, dtype):
    rng = np.random.default_rng(12)
    for k in (1, 3, 8):
        parts = [random_part(rng, 5000, 700, 30000, dtype) for _ in range(k)]
        parts.append(random_part(rng, 5000, 700, 0, dtype))
        dparts = [p.to_device() for p in parts]
        got = cb.merge(dparts, semiring).to_host()
        want = oracle.merge(parts, semiring)
        same(got, want, rtol=1e-12 if (dtype == "float64" and semiring == "plus_times") else 0.0)


def test_stage_products_merge_to_the_product_on_device():
    """Local products of 5 SUMMA stages, merged on the device, equal the
    product computed at once (PlusTimes<int64>, R-MAT s14)."""
    A = cb.gen_rmat(14, 8, seed=7)
    dA = A.to_device()
    full, _ = cb.spgemm(dA, dA)
    parts = []
    for As, Bs in stage_slices(A, A, 5):
        C, _ = cb.spgemm(As.to_device(), Bs.to_device())
        parts.append(C)
    got = cb.merge(parts).to_host()
    same(got, full.to_host())


def test_merge_edge_cases():
    rng = np.random.default_rng(2)
    empty = [random_part(rng, 50, 40, 0).to_device() for _ in range(3)]
    got = cb.merge(empty).to_host()
    assert got.nnz == 0 and got.nzc == 0 and (got.m, got.n) == (50, 40)
    with pytest.raises(cb.CbxError):
        cb.merge([random_part(rng, 50, 40, 10).to_device(), random_part(rng, 50, 41, 10).to_device()])
    with pytest.raises(ValueError):
        cb.merge([])


def test_merge_rejects_parts_outside_their_shape():
    """A part's column id equal to n lies inside the concatenation's next
    block (the merge's own operand check would accept it and merge the
    column into the wrong output column), and a cp that does not close at
    the part's nnz: both are rejected; the parts as given still merge."""
    rng = np.random.default_rng(3)
    parts = [random_part(rng, 50, 40, 60) for _ in range(3)]
    want = oracle.merge(parts)
    same(cb.merge([P.to_device() for P in parts]).to_host(), want)

    def column_at_n(P):
        P.jc[-1] = P.n

    def cp_short(P):
        P.cp[-1] -= 1

    for edit in (column_at_n, cp_short):
        bad = cb.Dcsc(parts[0].m, parts[0].n, parts[0].cp.copy(), parts[0].jc.copy(), parts[0].ir.copy(),
                      parts[0].val.copy(), parts[0].dtype)
        edit(bad)
        with pytest.raises(cb.CbxError):
            cb.merge([bad.to_device(), parts[1].to_device(), parts[2].to_device()])


@pytest.mark.parametrize("semiring,dtype,loops", [("plus_times", "int64", True), ("plus_times", "int64", False),
                                                  ("min_plus", "int64", False), ("select_max", "float64", True),
                                                  ("plus_times", "bool", True), ("plus_times", "float64", False)])
def test_from_coo_matches_oracle(semiring, dtype, loops):
    rng = np.random.default_rng(17)
    m, n, k = 70000, 50000, 400000
    rows = rng.integers(0, m, k)
    cols = rng.integers(0, n, k)
    rows[:5000] = cols[:5000]
    rows[5000:9000], cols[5000:9000] = rows[:4000], cols[:4000]  # duplicates
    vals = {"int64": rng.integers(-7, 8, k).astype(np.int64), "float64": rng.uniform(-1, 1, k),
            "bool": np.ones(k, np.uint8)}[dtype]
    dv = cb.DeviceArray.from_host(vals) if dtype != "bool" else None
    dr, dc = cb.DeviceArray.from_host(rows), cb.DeviceArray.from_host(cols)
    got = cb.from_coo_device(m, n, dr, dc, dv, dtype, semiring, loops).to_host()
    want = oracle.from_coo(m, n, rows, cols, vals if dtype != "bool" else None, dtype, semiring, loops)
    same(got, want, rtol=1e-12 if (dtype == "float64" and semiring == "plus_times") else 0.0)


def test_from_coo_edge_cases():
    rows = cb.DeviceArray.from_host(np.array([0, 1, 2], np.int64))
    cols = cb.DeviceArray.from_host(np.array([0, 1, 2], np.int64))
    got = cb.from_coo_device(3, 3, rows, cols, None, "int64", "plus_times", True).to_host()
    assert got.nnz == 0 and got.nzc == 0  # only loops
    bad = cb.DeviceArray.from_host(np.array([0, 5, 1], np.int64))
    with pytest.raises(cb.CbxError):
        cb.from_coo_device(3, 3, bad, cols)
    empty = cb.DeviceArray.from_host(np.zeros(0, np.int64))
    assert cb.from_coo_device(4, 4, empty, empty).to_host().nnz == 0


def test_from_coo_builds_the_rmat_benchmark_matrix():
    """The host generator's R-MAT edges (s16 ef16, values 1..9 summed over
    duplicates) rebuilt on the device give the generator's matrix."""
    A = cb.gen_rmat(16, 16)
    rows, cols = A.ir, A.col_ids()
    dA = cb.from_coo_device(A.m, A.n, cb.DeviceArray.from_host(rows), cb.DeviceArray.from_host(cols),
                            cb.DeviceArray.from_host(A.val)).to_host()
    same(dA, A)
