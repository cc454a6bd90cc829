"""Shared checks for the parity tests (test infrastructure)."""
from __future__ import annotations

import numpy as np

from combblas_amd import Dcsc


def same_pattern(X: Dcsc, Y: Dcsc) -> None:
    assert (X.m, X.n) == (Y.m, Y.n), "shape"
    assert X.nzc == Y.nzc and np.array_equal(X.jc, Y.jc), "stored columns differ"
    assert np.array_equal(X.cp, Y.cp), "column pointers differ"
    assert np.array_equal(X.ir, Y.ir), "row indices differ"


def bit_exact(X: Dcsc, Y: Dcsc) -> None:
    same_pattern(X, Y)
    assert X.val.dtype == Y.val.dtype
    assert np.array_equal(X.val.view(np.uint8), Y.val.view(np.uint8)), "values differ"


def sort_columns(X: Dcsc) -> Dcsc:
    """Rows ascending within each column (for unsorted-mode comparisons)."""
    cols = X.col_ids()
    order = np.lexsort((X.ir, cols))
    return Dcsc(X.m, X.n, X.cp, X.jc, X.ir[order], X.val[order], X.dtype)


def random_dcsc(rng, m, n, density, dtype="int64", lo=1, hi=9) -> Dcsc:
    mask = rng.random((m, n)) < density
    r, c = np.nonzero(mask)
    if dtype.startswith("float"):
        v = rng.uniform(lo, hi, r.size)
    else:
        v = rng.integers(lo, hi + 1, r.size)
    return Dcsc.from_coo(m, n, r, c, v, dtype)


def mcl_columns(seed: int, ncols: int = 240, m: int = 6000, dtype: str = "float32") -> Dcsc:
    """Columns shaped like an MCL expansion output, with the cases the prune /
    recover / select epilogue branches on: empty columns, short and long
    columns, sums above and below the recovery fraction, values exactly at the
    threshold, and runs of equal values around the selection rank."""
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    grid = np.array([1e-5, 5e-5, 1e-4, 2e-4, 5e-4, 1e-3, 4e-3], np.float64)
    for j in range(ncols):
        kind = j % 6
        if kind == 0 and j % 12 == 0:
            continue  # empty column
        length = int(rng.choice([rng.integers(1, 60), rng.integers(60, 1500), rng.integers(1500, 4000)]))
        r = rng.choice(m, size=min(length, m), replace=False)
        w = rng.lognormal(0.0, 1.0, r.size)
        if kind in (1, 4):  # values from a small grid: ties at every rank, hits on the threshold
            v = rng.choice(grid, r.size)
        else:
            v = w / w.sum() * (rng.uniform(0.5, 1.0) if kind == 2 else 1.0)
            if kind == 3:  # mass concentrated in a few entries
                v[: max(1, r.size // 50)] += 0.3
        rows.append(r)
        cols.append(np.full(r.size, j))
        vals.append(v)
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    v = np.concatenate(vals)
    if dtype == "float32":
        v = v.astype(np.float32)
    return Dcsc.from_coo(m, ncols, r, c, v, dtype)


# (hard threshold, select, recover, recover fraction) of the MCL epilogue:
# the expansion-step workload's (prune 1e-4, top-1000), HipMCL's defaults
# (MCL.cpp:140-160: 1e-4, 1100, 1400, 0.9), and harsher settings that reach
# every branch on small columns.
MCL_PARAMS = [(1e-4, 1000, 0, 0.9), (1e-4, 1100, 1400, 0.9), (1e-3, 50, 100, 0.5), (0.0, 10, 0, 0.9),
              (1e-4, 0, 300, 0.95), (2e-4, 200, 0, 0.0)]



def stage_slices(A, B, nstages: int):
    """(A(:, K_s), B(K_s, :)) for nstages contiguous ranges K_s of the inner
    dimension, at the full shape: the operands of SUMMA's stages on one process."""
    import numpy as _np
    cuts = _np.linspace(0, A.n, nstages + 1).astype(_np.int64)
    return [(A.column_range(int(cuts[s]), int(cuts[s + 1])), B.row_range(int(cuts[s]), int(cuts[s + 1])))
            for s in range(nstages)]
