"""The stage-merge oracle (MultiwayMerge<SR>, MultiwayMerge.h:428-541) pinned
two ways: against the genuine reference merge where oracle/_ref is built, and
against the identity SUMMA relies on -- merging the local products of the
stages A(:, K_s) * B(K_s, :) gives A * B -- which holds with any reference
absent."""
from __future__ import annotations

import numpy as np
import pytest

import combblas_amd as cb
import oracle
from helpers import stage_slices


def random_part(rng, m, n, nnz, dtype="int64"):
    key = np.unique(rng.integers(0, n, nnz) * m + rng.integers(0, m, nnz))
    vals = {"int64": rng.integers(-5, 6, key.size).astype(np.int64),
            "float64": rng.uniform(-2, 2, key.size),
            "bool": np.ones(key.size, np.uint8)}[dtype]
    return cb.Dcsc.from_coo(m, n, key % m, key // m, vals, dtype)


def same(a: cb.Dcsc, b: cb.Dcsc, rtol=0.0):
    assert (a.m, a.n) == (b.m, b.n)
    assert np.array_equal(a.cp, b.cp) and np.array_equal(a.jc, b.jc) and np.array_equal(a.ir, b.ir)
    if rtol:
        assert np.allclose(a.val, b.val, rtol=rtol, atol=1e-15)
    else:
        assert np.array_equal(a.val, b.val)


@pytest.mark.parametrize("semiring", ["plus_times", "min_plus", "select_max"])
def test_merge_matches_reference(semiring):
    if not oracle.ref_available():
        pytest.skip("oracle/_ref is not built (needs the reference tree)")
    rng = np.random.default_rng(3)
    for k in (1, 2, 5):
        parts = [random_part(rng, 300, 200, 4000) for _ in range(k)]
        parts.append(random_part(rng, 300, 200, 0))  # an empty part
        same(oracle.merge(parts, semiring), oracle.ref_merge(parts, semiring))


def test_merge_bool_and_fp_match_reference():
    if not oracle.ref_available():
        pytest.skip("oracle/_ref is not built (needs the reference tree)")
    rng = np.random.default_rng(4)
    parts = [random_part(rng, 120, 90, 900, "bool") for _ in range(3)]
    same(oracle.merge(parts), oracle.ref_merge(parts))
    parts = [random_part(rng, 120, 90, 900, "float64") for _ in range(4)]
    same(oracle.merge(parts), oracle.ref_merge(parts), rtol=1e-12)


def test_merge_keeps_explicit_zeros_and_sums_duplicates():
    a = cb.Dcsc.from_coo(4, 3, [0, 2, 1], [0, 0, 2], [5, 7, 1], "int64")
    b = cb.Dcsc.from_coo(4, 3, [0, 3, 1], [0, 0, 2], [-5, 2, 4], "int64")
    got = oracle.merge([a, b])
    assert got.cp.tolist() == [0, 3, 4] and got.jc.tolist() == [0, 2]
    assert got.ir.tolist() == [0, 2, 3, 1] and got.val.tolist() == [0, 7, 2, 5]
    assert oracle.merge([a, b], "min_plus").val.tolist() == [-5, 7, 2, 1]


@pytest.mark.parametrize("nstages", [1, 3, 7])
def test_stage_products_merge_to_the_product(nstages):
    """SUMMA's identity: sum over stages s of A(:, K_s) * B(K_s, :) = A * B."""
    A = cb.gen_rmat(9, 8, seed=11)
    want, _ = oracle.spgemm(A, A)
    parts = [oracle.spgemm(As, Bs)[0] for As, Bs in stage_slices(A, A, nstages)]
    same(oracle.merge(parts), want)


def test_merge_rejects_mismatched_parts():
    rng = np.random.default_rng(5)
    with pytest.raises(ValueError):
        oracle.merge([random_part(rng, 10, 10, 20), random_part(rng, 10, 11, 20)])


def test_column_and_row_ranges():
    A = cb.gen_rmat(8, 4)
    d = A.to_dense()
    idx = np.arange(A.n)
    for lo, hi in [(0, 10), (50, 120), (200, 256), (0, 256), (5, 5)]:
        keep = (idx >= lo) & (idx < hi)
        assert np.array_equal(A.column_range(lo, hi).to_dense(), np.where(keep[None, :], d, 0))
        assert np.array_equal(A.row_range(lo, hi).to_dense(), np.where(keep[:, None], d, 0))


@pytest.mark.parametrize("remove_loops", [True, False])
def test_from_coo_matches_reference_graph500_constructor(remove_loops):
    if not oracle.ref_available():
        pytest.skip("oracle/_ref is not built (needs the reference tree)")
    rng = np.random.default_rng(8)
    m = n = 500
    rows = rng.integers(0, m, 20000)
    cols = rng.integers(0, n, 20000)
    rows[:300] = cols[:300]  # self loops
    same(oracle.from_coo(m, n, rows, cols, remove_loops=remove_loops),
         oracle.ref_from_coo_graph500(m, n, rows, cols, remove_loops))


def test_from_coo_of_concatenated_parts_is_their_merge():
    """Duplicates combined by the semiring: building from the concatenation of
    several matrices' entries equals merging them (pinned to the reference
    MultiwayMerge above), for each semiring."""
    rng = np.random.default_rng(9)
    parts = [random_part(rng, 200, 150, 2500) for _ in range(4)]
    rows = np.concatenate([p.ir for p in parts])
    cols = np.concatenate([p.col_ids() for p in parts])
    vals = np.concatenate([p.val for p in parts])
    perm = rng.permutation(rows.size)  # any input order
    for sr in ("plus_times", "min_plus", "select_max"):
        same(oracle.from_coo(200, 150, rows[perm], cols[perm], vals[perm], semiring=sr), oracle.merge(parts, sr))
