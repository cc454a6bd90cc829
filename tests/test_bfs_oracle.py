"""Pins the oracle's SpMSpV and BFS (oracle/cbx_oracle.cpp) against the genuine
reference kernels (oracle/_ref: generic_gespmv / SpMXSpV and the TopDownBFS
loop around them) and against a brute-force BFS that applies the definition
(parent = largest-index frontier neighbour, root re-reached at level 2)."""
from __future__ import annotations

import numpy as np
import pytest

import combblas_amd as cb
import oracle

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (no reference tree)")


def graph(scale, ef=8, seed=None):
    return cb.gen_rmat(scale, ef, seed=seed, symmetrize=True, no_self_loops=True, pattern=True).astype("bool")


def brute_bfs(D: np.ndarray, root: int):
    n = D.shape[0]
    parents = -np.ones(n, np.int64)
    frontier = [root]
    it = 0
    while frontier:
        cand = {}
        for k in frontier:
            for i in np.nonzero(D[:, k])[0]:
                cand[i] = max(cand.get(i, -1), k)
        frontier = sorted(i for i in cand if parents[i] == -1)
        for i in frontier:
            parents[i] = cand[i]
        it += 1
    return parents, it


@pytest.mark.parametrize("seed", range(4))
def test_oracle_bfs_matches_bruteforce(seed):
    A = graph(6, 4, seed=seed + 11)
    D = A.to_dense().astype(bool)
    for root in cb.bfs_roots(A, 3, seed=seed):
        p1, i1 = oracle.bfs(A, int(root))
        p2, i2 = brute_bfs(D, int(root))
        assert i1 == i2 and np.array_equal(p1, p2)


def test_oracle_bfs_isolated_root():
    A = cb.Dcsc.from_coo(5, 5, [0, 1], [1, 0], [1, 1], "bool")
    p, it = oracle.bfs(A, 3)
    assert it == 1 and np.all(p == -1)


@needs_ref
@pytest.mark.parametrize("scale", [10, 13])
def test_oracle_bfs_matches_reference(scale):
    A = graph(scale, 16)
    for root in cb.bfs_roots(A, 4, seed=scale):
        p1, i1 = oracle.bfs(A, int(root))
        p2, i2 = oracle.ref_bfs(A, int(root))
        assert i1 == i2 and np.array_equal(p1, p2)


@needs_ref
@pytest.mark.parametrize("semiring,dtype", [("select_max", "bool"), ("plus_times", "int64"),
                                            ("min_plus", "int64"), ("plus_times", "float64")])
def test_oracle_spmspv_matches_reference(semiring, dtype):
    """The oracle's SpMXSpV against the genuine generic_gespmv, bit for bit
    (the FP64 sums too: the oracle keeps the reference's heap order); min-plus
    with operands at numeric_limits<int64>::max() exercises inf_plus."""
    rng = np.random.default_rng(3)
    G = graph(11, 8)
    A = G.astype(dtype)
    if dtype == "int64":
        A.val = rng.integers(-5, 9, A.nnz).astype(np.int64)
        if semiring == "min_plus":
            A.val[rng.random(A.nnz) < 0.1] = np.iinfo(np.int64).max
    elif dtype == "float64":
        A.val = rng.uniform(-2.0, 2.0, A.nnz)
    for frac in (0.001, 0.05, 0.5):
        x = np.sort(rng.choice(A.n, max(1, int(frac * A.n)), replace=False)).astype(np.int32)
        if dtype == "float64":
            xv = rng.uniform(-1.0, 1.0, x.size)
        else:
            xv = rng.integers(-100, 100, x.size).astype(np.int64)
            if semiring == "min_plus":
                xv[rng.random(x.size) < 0.1] = np.iinfo(np.int64).max
        y1 = oracle.spmspv(A, x, xv, semiring)
        y2 = oracle.ref_spmspv(A, x, xv, semiring)
        assert np.array_equal(y1[0], y2[0])
        assert np.array_equal(y1[1].view(np.uint8), y2[1].view(np.uint8))


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (no reference tree)")
@pytest.mark.parametrize("splits", [1, 4, 13])
def test_reference_threaded_bfs_matches_oracle(splits):
    """The CPU baseline's BFS path (RowSplit + generic_gespmv_threaded, the
    multithreaded TopDownBFS) gives the oracle's parents and level counts
    whatever the number of row splits."""
    A = cb.gen_rmat(13, 16, symmetrize=True, no_self_loops=True, pattern=True).astype("bool")
    for root in cb.bfs_roots(A, 3, seed=splits):
        got, levels, _ = oracle.ref_bfs_threaded(A, [int(root)], splits)
        want, it = oracle.bfs(A, int(root))
        assert levels == it and np.array_equal(got, want)
