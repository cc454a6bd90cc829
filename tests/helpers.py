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
