"""GPU SpMSpV (cbx_spmspv) and BFS (cbx_bfs_*) against the oracle: integer
semirings and BFS parents bit-exact; FP plus-times within the a-priori bound."""
from __future__ import annotations

import numpy as np
import pytest

import combblas_amd as cb
import oracle

pytestmark = pytest.mark.gpu


def graph(scale, ef=16):
    return cb.gen_rmat(scale, ef, symmetrize=True, no_self_loops=True, pattern=True).astype("bool")


@pytest.mark.parametrize("semiring", ["select_max", "plus_times", "min_plus"])
@pytest.mark.parametrize("adtype", ["bool", "int64", "float64"])
def test_spmspv_matches_oracle(semiring, adtype):
    rng = np.random.default_rng(17)
    G = graph(12, 8)
    A = G.astype(adtype)
    if adtype == "int64":
        A.val = rng.integers(-9, 10, A.nnz).astype(np.int64)
    elif adtype == "float64":
        A.val = rng.uniform(0.5, 2.0, A.nnz)
    xdt = np.float64 if adtype == "float64" else np.int64
    dA = A.to_device()
    for frac in (0.0005, 0.02, 0.3):
        x = np.sort(rng.choice(A.n, max(1, int(frac * A.n)), replace=False)).astype(np.int32)
        xv = (rng.uniform(0.5, 2.0, x.size) if xdt == np.float64 else rng.integers(-50, 50, x.size)).astype(xdt)
        gi, gv = cb.spmspv(dA, x, xv, semiring)
        oi, ov = oracle.spmspv(A, x, xv, semiring)
        assert np.array_equal(gi, oi)
        if xdt == np.float64 and semiring == "plus_times":
            assert np.allclose(gv, ov, rtol=1e-12, atol=0)
        else:
            assert np.array_equal(gv, ov)


def test_spmspv_empty_and_invalid():
    A = graph(8, 4)
    dA = A.to_device()
    yi, yv = cb.spmspv(dA, np.zeros(0, np.int32), np.zeros(0, np.int64))
    assert yi.size == 0
    with pytest.raises(cb.CbxError):
        cb.spmspv(dA, np.array([5, 3], np.int32), np.array([1, 2], np.int64))  # not ascending


@pytest.mark.parametrize("scale", [12, 16])
def test_bfs_parents_bit_exact(scale):
    A = graph(scale)
    plan = cb.BfsPlan(A.to_device())
    for root in cb.bfs_roots(A, 8, seed=scale):
        got, st = plan.run(int(root))
        want, it = oracle.bfs(A, int(root))
        assert st["levels"] == it
        assert st["reached"] == int((want >= 0).sum())
        assert np.array_equal(got, want)


def test_bfs_isolated_root_and_tiny_graph():
    A = cb.Dcsc.from_coo(6, 6, [0, 1, 1, 2], [1, 0, 2, 1], [1, 1, 1, 1], "bool")
    plan = cb.BfsPlan(A.to_device())
    got, st = plan.run(4)  # isolated
    assert st["levels"] == 1 and np.all(got == -1)
    got, st = plan.run(0)
    want, it = oracle.bfs(A, 0)
    assert np.array_equal(got, want) and st["levels"] == it


@pytest.mark.parametrize("direction", ["top_down", "optimizing"])
def test_bfs_rmat_s20_ef16_config(direction):
    """BFS workload: symmetrized unweighted R-MAT s20 ef16, self-loops removed;
    4 of the 64 seeded roots (seed 4242) checked against the oracle bit for
    bit, with the workload's own top-down-only plan and with the per-level
    direction switch."""
    A = graph(20, 16)
    plan = cb.BfsPlan(A.to_device(), direction)
    for root in cb.bfs_roots(A, 64, seed=4242)[:4]:
        got, st = plan.run(int(root))
        want, it = oracle.bfs(A, int(root))
        assert st["levels"] == it and np.array_equal(got, want)
        if direction == "top_down":
            assert st["bottom_up_levels"] == 0
    plan.release()


@pytest.mark.parametrize("semiring", ["plus_times", "min_plus", "select_max"])
@pytest.mark.parametrize("dtype", ["int64", "float64", "float32", "bool"])
def test_spmv_dense_matches_oracle(semiring, dtype):
    """dcsc_gespmv on the device: integer semirings bit-exact, FP within 1e-12
    (FP64) / 1e-5 (FP32) relative (atomic accumulation order)."""
    rng = np.random.default_rng(8)
    A = cb.gen_rmat(14, 8, seed=21)
    xd = {"int64": np.int64, "float64": np.float64, "float32": np.float32, "bool": np.int64}[dtype]
    if dtype == "bool":
        A = A.astype("bool")
    elif dtype == "int64":
        A.val = rng.integers(-9, 10, A.nnz).astype(np.int64)
    else:
        A = A.astype(dtype)
        A.val = rng.uniform(0.5, 2.0, A.nnz).astype(xd)
    x = (rng.uniform(0.5, 2.0, A.n) if xd != np.int64 else rng.integers(-20, 20, A.n)).astype(xd)
    y = (rng.uniform(0.5, 2.0, A.m) if xd != np.int64 else rng.integers(-5, 5, A.m)).astype(xd)
    got = cb.spmv_dense(A.to_device(), x, y, semiring)
    want = oracle.spmv_dense(A, x, y, semiring)
    if xd == np.int64 or semiring != "plus_times":
        assert np.array_equal(got, want)
    else:
        tol = 1e-12 if xd == np.float64 else 1e-5
        assert np.allclose(got, want, rtol=tol, atol=0)


def test_spmv_dense_device_accumulates_in_place():
    """Device-resident vectors: two calls accumulate twice (y += A*x); vectors
    of the wrong length or mixed dtypes are refused before any launch."""
    rng = np.random.default_rng(9)
    A = cb.gen_rmat(12, 8, seed=5)
    A.val = rng.integers(-9, 10, A.nnz).astype(np.int64)
    dA = A.to_device()
    x = rng.integers(-20, 20, A.n).astype(np.int64)
    dx = cb.DeviceArray.from_host(x)
    dy = cb.DeviceArray.from_host(np.zeros(A.m, np.int64))
    try:
        cb.spmv_dense_device(dA, dx, dy)
        cb.spmv_dense_device(dA, dx, dy)
        want = oracle.spmv_dense(A, x, np.zeros(A.m, np.int64), "plus_times")
        assert np.array_equal(dy.to_host(), 2 * want)
        short = cb.DeviceArray.from_host(np.zeros(A.m - 1, np.int64))
        with pytest.raises(ValueError):
            cb.spmv_dense_device(dA, dx, short)
        short.release()
        fy = cb.DeviceArray.from_host(np.zeros(A.m, np.float64))
        with pytest.raises(ValueError):
            cb.spmv_dense_device(dA, dx, fy)
        fy.release()
    finally:
        dx.release()
        dy.release()


@pytest.mark.parametrize("shape", ["rmat_s16", "tall"])
def test_spmv_dense_bucketed_and_direct_forms(shape):
    """Both device forms of dcsc_gespmv: the bucketed one (R-MAT s16: ~1 M
    entries in 16 row blocks, the hot block split over several accumulating
    workgroups) and the direct atomic one (m above 2^24 rows)."""
    rng = np.random.default_rng(21)
    if shape == "rmat_s16":
        A = cb.gen_rmat(16, 16, seed=3)
    else:
        m, n, k = (1 << 24) + 4096, 5000, 200_000
        key = np.unique(rng.integers(0, n, k) * m + rng.integers(0, m, k))
        A = cb.Dcsc.from_coo(m, n, key % m, key // m, np.ones(key.size, np.int64), "int64")
    A.val = rng.integers(-9, 10, A.nnz).astype(np.int64)
    dA = A.to_device()
    x = rng.integers(-20, 20, A.n).astype(np.int64)
    for sr in ("plus_times", "min_plus", "select_max"):
        y = rng.integers(-50, 50, A.m).astype(np.int64)
        assert np.array_equal(cb.spmv_dense(dA, x, y, sr), oracle.spmv_dense(A, x, y, sr))
    Af = A.astype("float64")
    Af.val = rng.uniform(0.5, 2.0, A.nnz)
    xf, yf = rng.uniform(0.5, 2.0, A.n), rng.uniform(0.5, 2.0, A.m)
    got = cb.spmv_dense(Af.to_device(), xf, yf, "plus_times")
    assert np.allclose(got, oracle.spmv_dense(Af, xf, yf, "plus_times"), rtol=1e-12, atol=0)


@pytest.mark.parametrize("shape", ["rmat_s16", "tall"])
@pytest.mark.parametrize("fault", ["row_out_of_range", "rows_unsorted", "column_out_of_range", "cp_descends"])
def test_spmv_dense_rejects_malformed_matrix(shape, fault):
    """The structure check of dcsc_gespmv's operand runs with the product (the
    bucketed form: its per-column half beside the launches, its per-entry half
    in the counting pass, and the kernels that index through the structure
    wait for the verdict; the direct form checks first).  A malformed matrix
    is reported as CBX_ERR_INVALID and y is left as it was."""
    rng = np.random.default_rng(23)
    if shape == "rmat_s16":
        A = cb.gen_rmat(16, 16, seed=3)
    else:
        m, n, k = (1 << 24) + 4096, 5000, 50_000
        key = np.unique(rng.integers(0, n, k) * m + rng.integers(0, m, k))
        A = cb.Dcsc.from_coo(m, n, key % m, key // m, np.ones(key.size, np.int64), "int64")
    bad = cb.Dcsc(A.m, A.n, A.cp.copy(), A.jc.copy(), A.ir.copy(), A.val.copy())
    q = int(np.argmax(np.diff(bad.cp) >= 2))  # a column with two entries or more
    if fault == "row_out_of_range":
        bad.ir[bad.cp[q + 1] - 1] = A.m  # the column's last row, one past the end
    elif fault == "rows_unsorted":
        b0 = bad.cp[q]
        bad.ir[b0], bad.ir[b0 + 1] = bad.ir[b0 + 1], bad.ir[b0]
    elif fault == "column_out_of_range":
        bad.jc[-1] = A.n
    else:
        bad.cp[q + 1] = bad.cp[q]  # an empty stored column
    dA = bad.to_device()
    x = rng.integers(-20, 20, A.n).astype(np.int64)
    y0 = rng.integers(-50, 50, A.m).astype(np.int64)
    dx, dy = cb.DeviceArray.from_host(x), cb.DeviceArray.from_host(y0)
    try:
        with pytest.raises(cb.CbxError) as ei:
            cb.spmv_dense_device(dA, dx, dy, "plus_times")
        assert ei.value.status == cb.ERR_INVALID
        assert np.array_equal(dy.to_host(), y0)
        good = A.to_device()
        cb.spmv_dense_device(good, dx, dy, "plus_times")  # the library is usable afterwards
        assert np.array_equal(dy.to_host(), oracle.spmv_dense(A, x, y0, "plus_times"))
    finally:
        dx.release()
        dy.release()


@pytest.mark.parametrize("scale", [12, 16])
def test_bfs_directions_give_identical_parents(scale):
    """Top-down only, bottom-up only and the per-level switch give the same
    parents (the largest frontier neighbour) and level counts, all equal to
    the oracle; the switching plan runs some levels each way at s16."""
    A = graph(scale)
    dA = A.to_device()
    plans = {d: cb.BfsPlan(dA, d) for d in ("top_down", "bottom_up", "optimizing")}
    for root in cb.bfs_roots(A, 4, seed=100 + scale):
        want, it = oracle.bfs(A, int(root))
        for d, plan in plans.items():
            got, st = plan.run(int(root))
            assert st["levels"] == it and np.array_equal(got, want), d
            if d == "top_down":
                assert st["bottom_up_levels"] == 0
            elif d == "bottom_up":
                assert st["bottom_up_levels"] == st["levels"]
            elif scale == 16 and st["reached"] > A.n // 4:
                assert 0 < st["bottom_up_levels"] < st["levels"]
    for plan in plans.values():
        plan.release()
    with pytest.raises(KeyError):
        cb.BfsPlan(dA, "sideways")


def test_bfs_directed_graph_runs_top_down():
    """A directed pattern (R-MAT s14 without symmetrisation): the plan finds it
    asymmetric, runs every level top-down even when asked to optimise, and
    refuses bottom-up levels; parents equal the oracle's SpMV-based BFS."""
    A = cb.gen_rmat(14, 16, no_self_loops=True, pattern=True).astype("bool")
    assert not np.array_equal(A.to_dense(), A.to_dense().T)
    dA = A.to_device()
    plan = cb.BfsPlan(dA)
    for root in cb.bfs_roots(A, 4, seed=3):
        got, st = plan.run(int(root))
        want, it = oracle.bfs(A, int(root))
        assert st["levels"] == it and st["bottom_up_levels"] == 0
        assert np.array_equal(got, want)
    with pytest.raises(cb.CbxError) as ei:
        plan.set_direction("bottom_up")
    assert ei.value.status == cb.ERR_INVALID
    plan.release()
