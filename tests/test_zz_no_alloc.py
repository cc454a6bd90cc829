"""cbx_spmspv_noalloc (scratch in the library's static per-GPU array, no
device allocation) against cbx_spmspv through the C ABI: bit for bit on two
calls in a row, the second with a smaller matrix and vector, for integer,
min-plus and select-max semirings (floating-point plus-times within the
bound of two summation orders); a matrix whose scratch exceeds the array is refused
with CBX_ERR_LIMIT before any device work."""
from __future__ import annotations

import numpy as np
import pytest

import combblas_amd as cb

pytestmark = pytest.mark.gpu


def graph(scale, ef=16):
    return cb.gen_rmat(scale, ef, symmetrize=True, no_self_loops=True, pattern=True).astype("bool")


def frontier(rng, n, count):
    return np.sort(rng.choice(n, count, replace=False)).astype(np.int32)


def same(a, b):
    assert np.array_equal(a[0], b[0]), "indices differ"
    assert a[1].dtype == b[1].dtype and np.array_equal(a[1].view(np.uint8), b[1].view(np.uint8)), "values differ"


@pytest.mark.parametrize("semiring,adtype", [("select_max", "bool"), ("plus_times", "int64"),
                                             ("min_plus", "float64"), ("plus_times", "float32")])
def test_noalloc_equals_spmspv_on_two_calls(semiring, adtype):
    rng = np.random.default_rng(91)
    big, small = graph(20, 16), graph(12, 8)  # the second call is smaller
    for G, nx in ((big, 40000), (small, 300)):
        A = G.astype(adtype)
        xdt = {"bool": np.int64, "int64": np.int64, "float64": np.float64, "float32": np.float32}[adtype]
        if adtype == "int64":
            A.val = rng.integers(-9, 10, A.nnz).astype(np.int64)
        elif adtype.startswith("float"):
            A.val = rng.uniform(0.5, 2.0, A.nnz).astype(xdt)
        dA = A.to_device()
        x = frontier(rng, A.n, nx)
        xv = (rng.integers(-50, 50, x.size) if xdt == np.int64 else rng.uniform(0.5, 2.0, x.size)).astype(xdt)
        want = cb.spmspv(dA, x, xv, semiring)
        got = cb.spmspv(dA, x, xv, semiring, noalloc=True)
        assert want[0].size > 0
        if semiring == "plus_times" and adtype.startswith("float"):
            # device atomics sum in no fixed order: all terms are positive, so two
            # orders differ by at most 2 gamma_k of the sum, k <= the row's degree
            # (the pattern is symmetric: row degree = column length)
            assert np.array_equal(got[0], want[0])
            deg = np.zeros(A.n, np.int64)
            deg[A.jc] = np.diff(A.cp)
            u = np.finfo(xdt).eps / 2
            k = deg[want[0]].astype(np.float64)
            assert np.all(np.abs(got[1].astype(np.float64) - want[1]) <= 2 * k * u / (1 - k * u) * want[1])
        else:
            same(got, want)
        dA.release()


def test_noalloc_empty_vector_and_repeat():
    A = graph(10, 8)
    dA = A.to_device()
    gi, gv = cb.spmspv(dA, np.zeros(0, np.int32), np.zeros(0, np.int64), "select_max", noalloc=True)
    assert gi.size == 0 and gv.size == 0
    x = np.arange(0, A.n, 7, dtype=np.int32)
    xv = x.astype(np.int64)
    first = cb.spmspv(dA, x, xv, "select_max", noalloc=True)
    second = cb.spmspv(dA, x, xv, "select_max", noalloc=True)  # the array keeps the last call's contents
    same(first, second)
    same(first, cb.spmspv(dA, x, xv, "select_max"))
    dA.release()


def test_noalloc_rejects_matrix_beyond_its_scratch():
    m = n = 1 << 23  # 8 MiB bitmap words + 64 MiB accumulator + 32 MiB column map: beyond 16 MiB
    A = cb.Dcsc.from_coo(m, n, np.array([0, 5]), np.array([1, 3]), np.array([1, 1]), "int64")
    dA = A.to_device()
    with pytest.raises(cb.CbxError) as ei:
        cb.spmspv(dA, np.array([1], np.int32), np.array([2], np.int64), "plus_times", noalloc=True)
    assert ei.value.status == cb.ERR_LIMIT
    gi, gv = cb.spmspv(dA, np.array([1], np.int32), np.array([2], np.int64), "plus_times")
    assert gi.tolist() == [0] and gv.tolist() == [2]
    dA.release()
