"""Closed-form checks of device results that need no CPU product.

The device reduces every stored column of a matrix X to the linear checksum
sum_i u(i) X(i,j) with u(i) = mix64(i ^ salt) in wrapping uint64 arithmetic
(cbx_dcsc_column_checksums).  For an integer plus-times product C = A*B that
checksum has the closed form sum_k B(k,j) w_k, w_k = sum_i u(i) A(i,k), which
costs O(nnz(A) + nnz(B)) on the host, so a product of any size (including
one whose output does not fit in host memory) is checked entry for entry up
to a 2^-64 collision chance.
"""
from __future__ import annotations

import numpy as np

from . import Dcsc


def mix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser on uint64 arrays (wrapping arithmetic); the same
    function the device checksum kernel applies to row indices."""
    x = x + np.uint64(0x9E3779B97F4A7C15)
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def product_checksums(A: Dcsc, B: Dcsc, salt: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """Closed form of the per-column checksums of C = A*B (plus-times, int64
    values, wrapping uint64).  Returns (B.jc, checksum per stored column of
    B); a column whose product is empty gets 0."""
    with np.errstate(over="ignore"):
        u = mix64(A.ir.astype(np.uint64) ^ np.uint64(salt))
        contrib = u * A.val.astype(np.int64).view(np.uint64)
        w_col = np.add.reduceat(contrib, A.cp[:-1]) if A.nzc else np.zeros(0, np.uint64)
        w = np.zeros(A.n, np.uint64)
        w[A.jc] = w_col
        t = w[B.ir] * B.val.astype(np.int64).view(np.uint64)
        cs = np.add.reduceat(t, B.cp[:-1]) if B.nzc else np.zeros(0, np.uint64)
    return B.jc.copy(), cs


def product_matches(C, A: Dcsc, B: Dcsc, salt: int = 0x5EED, col_nnz: np.ndarray | None = None) -> bool:
    """Device product C (a DeviceDcsc) against the closed form: a valid DCSC
    whose stored columns and per-column checksums are those of A*B.  With
    col_nnz (the exact entry count of C(:,j) for every stored column of B,
    e.g. from cbx_spgemm_symbolic) the per-column entry counts must match
    too, which a zero-valued spurious or dropped entry would change while
    leaving the checksum alone."""
    C.check()
    got = C.column_checksums(salt=salt)
    cpC, jcC = C.structure()
    jcB, want = product_checksums(A, B, salt=salt)
    pos = np.searchsorted(jcB, jcC)
    if pos.size and (pos.max() >= jcB.size or not np.array_equal(jcB[pos], jcC)):
        return False
    dropped = np.ones(jcB.size, bool)
    dropped[pos] = False
    ok = bool(np.array_equal(got, want[pos]) and not np.any(want[dropped]))
    if col_nnz is not None:
        col_nnz = np.asarray(col_nnz, np.int64)
        ok = ok and col_nnz.size == jcB.size and bool(np.array_equal(np.diff(cpC), col_nnz[pos])) \
            and not np.any(col_nnz[dropped])
    return ok


def stencil27_coarse_rowsums(g: int) -> np.ndarray:
    """Row sums of R*A*R^T for the 27-point stencil on a g^3 grid under 2x2x2
    aggregation: sum over the fine vertices i of aggregate I of A's row sum,
    26 - (number of grid neighbours of i)."""
    x = np.arange(g)
    nb = 1 + (x > 0) + (x < g - 1)  # neighbours-or-self along one axis
    rs = 26 - (nb[:, None, None] * nb[None, :, None] * nb[None, None, :] - 1)  # indexed [z, y, x]
    gc = (g + 1) // 2
    agg = (x // 2)
    out = np.zeros((gc, gc, gc), np.float64)
    np.add.at(out, (agg[:, None, None], agg[None, :, None], agg[None, None, :]), rs)
    return out.reshape(-1)
