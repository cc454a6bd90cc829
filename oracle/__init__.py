"""ORACLE -- test infrastructure only (results are compared against it; never the product).

Python binding of the CPU restatement in ``oracle/cbx_oracle.cpp``
(``oracle/liboracle.so``) and, when it was built in this checkout, of the
genuine CombBLAS kernels compiled from the reference tree
(``oracle/_ref/libcbref.so``).  Only tests/, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU-baseline leg use this package.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from combblas_amd import Dcsc, cbx_dcsc, SEMIRINGS

HERE = Path(__file__).resolve().parent
ORACLE_LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libcbref.so"
REF_TREE = Path("/root/reference")

MODES = {"hybrid": 0, "hash": 1, "hash_unsorted": 2}

_orc = None
_ref = None


def build(ref: bool | None = None) -> None:
    """Compile liboracle.so, and _ref/libcbref.so when the reference tree is present."""
    targets = ["all"]
    if ref is None:
        ref = (REF_TREE / "include" / "CombBLAS" / "CombBLAS.h").exists()
    if ref:
        targets.append("ref")
    r = subprocess.run(["make", "-s", "-C", str(HERE), *targets], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")


def _load_oracle() -> C.CDLL:
    global _orc
    if _orc is None:
        if not ORACLE_LIB.exists():
            build(ref=False)
        L = C.CDLL(str(ORACLE_LIB))
        P = C.POINTER
        L.orc_spgemm.argtypes = [P(cbx_dcsc), P(cbx_dcsc), C.c_int32, C.c_int32, C.c_int64, C.c_int64,
                                 P(cbx_dcsc), P(C.c_int64)]
        L.orc_spgemm.restype = C.c_int32
        L.orc_symbolic.argtypes = [P(cbx_dcsc), P(cbx_dcsc), C.c_void_p, C.c_void_p]
        L.orc_symbolic.restype = C.c_int32
        L.orc_free.argtypes = [P(cbx_dcsc)]
        L.orc_free.restype = None
        L.orc_time_spgemm.argtypes = [P(cbx_dcsc), P(cbx_dcsc), C.c_int32, C.c_int32, C.c_int64, C.c_int64,
                                      P(C.c_double), P(C.c_int64), P(C.c_int64)]
        L.orc_time_spgemm.restype = C.c_int32
        L.orc_spmspv.argtypes = [P(cbx_dcsc), C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p,
                                 C.c_void_p]
        L.orc_spmspv.restype = C.c_int32
        L.orc_bfs.argtypes = [P(cbx_dcsc), C.c_int64, C.c_void_p]
        L.orc_bfs.restype = C.c_int32
        L.orc_mcl_prune.argtypes = [P(cbx_dcsc), C.c_double, C.c_int64, C.c_int64, C.c_double, P(cbx_dcsc),
                                    C.c_void_p]
        L.orc_mcl_prune.restype = C.c_int32
        L.orc_spmv_dense.argtypes = [P(cbx_dcsc), C.c_void_p, C.c_void_p, C.c_int32, C.c_int32]
        L.orc_spmv_dense.restype = C.c_int32
        L.orc_galerkin_direct.argtypes = [C.c_int32, P(cbx_dcsc)]
        L.orc_galerkin_direct.restype = C.c_int32
        L.orc_estimate_sampling.argtypes = [P(cbx_dcsc), P(cbx_dcsc), C.c_int32, C.c_void_p, P(C.c_float),
                                             P(C.c_int64)]
        L.orc_estimate_sampling.restype = C.c_int32
        L.orc_exp_samples.argtypes = [C.c_int64, C.c_void_p]
        L.orc_exp_samples.restype = None
        L.orc_max_threads.restype = C.c_int32
        L.orc_set_threads.argtypes = [C.c_int32]
        _orc = L
    return _orc


def ref_available() -> bool:
    return REF_LIB.exists()


def _load_ref() -> C.CDLL:
    global _ref
    if _ref is None:
        if not REF_LIB.exists():
            raise RuntimeError("oracle/_ref/libcbref.so is not built (needs the reference tree)")
        L = C.CDLL(str(REF_LIB))
        P = C.POINTER
        L.ref_spgemm.argtypes = [P(cbx_dcsc), P(cbx_dcsc), C.c_int32, C.c_int32, P(cbx_dcsc)]
        L.ref_spgemm.restype = C.c_int32
        L.ref_symbolic.argtypes = [P(cbx_dcsc), P(cbx_dcsc), C.c_void_p, C.c_void_p]
        L.ref_symbolic.restype = C.c_int32
        L.ref_time_hybrid_i64.argtypes = [P(cbx_dcsc), P(cbx_dcsc), P(C.c_double), P(C.c_int64)]
        L.ref_time_hybrid_i64.restype = C.c_int32
        for nm in ("ref_spmspv_selectmax_bool", "ref_spmspv_plustimes_i64", "ref_spmspv_minplus_i64",
                   "ref_spmspv_plustimes_f64"):
            getattr(L, nm).argtypes = [P(cbx_dcsc), C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
            getattr(L, nm).restype = C.c_int32
        L.ref_bfs.argtypes = [P(cbx_dcsc), C.c_int64, C.c_void_p]
        L.ref_bfs.restype = C.c_int32
        L.ref_time_hybrid_dcsc_i64.argtypes = [P(cbx_dcsc), P(cbx_dcsc), P(C.c_double), P(C.c_double),
                                               P(C.c_int64)]
        L.ref_time_hybrid_dcsc_i64.restype = C.c_int32
        L.ref_bfs_threaded.argtypes = [P(cbx_dcsc), C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, P(C.c_double)]
        L.ref_bfs_threaded.restype = C.c_int32
        L.ref_estimate_sampling.argtypes = [P(cbx_dcsc), P(cbx_dcsc), C.c_int32]
        L.ref_estimate_sampling.restype = C.c_int64
        L.ref_spmv_dense.argtypes = [P(cbx_dcsc), C.c_void_p, C.c_void_p, C.c_int32]
        L.ref_spmv_dense.restype = C.c_int32
        L.ref_mcl_prune.argtypes = [P(cbx_dcsc), C.c_double, C.c_int64, C.c_int64, C.c_double, P(cbx_dcsc)]
        L.ref_mcl_prune.restype = C.c_int32
        L.ref_multiway_merge.argtypes = [C.c_int32, P(cbx_dcsc), C.c_int32, P(cbx_dcsc)]
        L.ref_multiway_merge.restype = C.c_int32
        L.ref_graph500_tuples.argtypes = [C.c_int64, C.c_int64, C.c_void_p, C.c_int64, C.c_int32, P(cbx_dcsc)]
        L.ref_graph500_tuples.restype = C.c_int32
        L.ref_free.argtypes = [P(cbx_dcsc)]
        L.ref_free.restype = None
        _ref = L
    return _ref


def spgemm(A: Dcsc, B: Dcsc, semiring: str = "plus_times", mode: str = "hybrid",
           cols: tuple[int, int] | None = None) -> tuple[Dcsc, int]:
    """Oracle product (restatement).  cols=(c0, c1) restricts to B's stored
    columns [c0, c1).  Returns (C, flops)."""
    L = _load_oracle()
    sa, ka = A._cstruct()
    sb, kb = B._cstruct()
    out = cbx_dcsc()
    fl = C.c_int64(0)
    c0, c1 = cols if cols is not None else (0, -1)
    st = L.orc_spgemm(C.byref(sa), C.byref(sb), SEMIRINGS[semiring], MODES[mode], c0, c1, C.byref(out),
                      C.byref(fl))
    if st != 0:
        raise RuntimeError(f"orc_spgemm failed with status {st}")
    return Dcsc.from_struct(out, L.orc_free), int(fl.value)


def symbolic(A: Dcsc, B: Dcsc) -> tuple[np.ndarray, np.ndarray]:
    """(flop_j, nnz_j) for every stored column of B (estimateFLOP, estimateNNZ_Hash)."""
    L = _load_oracle()
    sa, ka = A._cstruct()
    sb, kb = B._cstruct()
    flop = np.zeros(B.nzc, np.int64)
    nnz = np.zeros(B.nzc, np.int64)
    st = L.orc_symbolic(C.byref(sa), C.byref(sb), flop.ctypes.data, nnz.ctypes.data)
    if st != 0:
        raise RuntimeError(f"orc_symbolic failed with status {st}")
    return flop, nnz


def time_spgemm(A: Dcsc, B: Dcsc, semiring: str = "plus_times", mode: str = "hybrid",
                cols: tuple[int, int] | None = None) -> tuple[float, int, int]:
    """Seconds, nnz and flops of the restated kernel on B's stored columns cols."""
    L = _load_oracle()
    sa, ka = A._cstruct()
    sb, kb = B._cstruct()
    secs, nnz, fl = C.c_double(0), C.c_int64(0), C.c_int64(0)
    c0, c1 = cols if cols is not None else (0, -1)
    st = L.orc_time_spgemm(C.byref(sa), C.byref(sb), SEMIRINGS[semiring], MODES[mode], c0, c1, C.byref(secs),
                           C.byref(nnz), C.byref(fl))
    if st != 0:
        raise RuntimeError(f"orc_time_spgemm failed with status {st}")
    return secs.value, int(nnz.value), int(fl.value)


def ref_time_hybrid_i64(A: Dcsc, B: Dcsc) -> tuple[float, int]:
    """Seconds and nnz of the genuine LocalHybridSpGEMM<PlusTimes<int64>> (oracle/_ref)."""
    L = _load_ref()
    sa, ka = A._cstruct()
    sb, kb = B._cstruct()
    secs, nnz = C.c_double(0), C.c_int64(0)
    if L.ref_time_hybrid_i64(C.byref(sa), C.byref(sb), C.byref(secs), C.byref(nnz)) != 0:
        raise RuntimeError("ref_time_hybrid_i64 failed")
    return secs.value, int(nnz.value)


def ref_time_hybrid_dcsc_i64(A: Dcsc, B: Dcsc) -> tuple[float, float, int]:
    """(seconds of the genuine LocalHybridSpGEMM<PlusTimes<int64>>, seconds of it plus the
    SpDCCols(SpTuples) conversion, nnz) on B's columns (oracle/_ref)."""
    L = _load_ref()
    sa, ka = A._cstruct()
    sb, kb = B._cstruct()
    t1, t2, nnz = C.c_double(0), C.c_double(0), C.c_int64(0)
    if L.ref_time_hybrid_dcsc_i64(C.byref(sa), C.byref(sb), C.byref(t1), C.byref(t2), C.byref(nnz)) != 0:
        raise RuntimeError("ref_time_hybrid_dcsc_i64 failed")
    return t1.value, t2.value, int(nnz.value)


def ref_bfs_threaded(A: Dcsc, roots, splits: int) -> tuple[np.ndarray, int, float]:
    """The reference's multithreaded BFS level loop (RowSplit + generic_gespmv_threaded) over
    `roots`: (parents of the last root, total levels, seconds of the level loops)."""
    L = _load_ref()
    sa, ka = A._cstruct()
    r = np.ascontiguousarray(roots, np.int64)
    par = np.zeros(A.n, np.int64)
    secs = C.c_double(0)
    lv = L.ref_bfs_threaded(C.byref(sa), r.ctypes.data, int(r.size), int(splits), par.ctypes.data, C.byref(secs))
    return par, int(lv), secs.value


def column_slice(B: Dcsc, c0: int, c1: int) -> Dcsc:
    """B restricted to its stored columns [c0, c1) (same shape), like
    SpDCCols::ColSplit (SpDCCols.cpp:1054) without renumbering."""
    b0, b1 = int(B.cp[c0]), int(B.cp[c1])
    return Dcsc(B.m, B.n, B.cp[c0:c1 + 1] - b0, B.jc[c0:c1], B.ir[b0:b1], B.val[b0:b1], B.dtype)


def set_threads(n: int) -> None:
    _load_oracle().orc_set_threads(n)


def max_threads() -> int:
    return int(_load_oracle().orc_max_threads())


def ref_spgemm(A: Dcsc, B: Dcsc, semiring: str = "plus_times", mode: str = "hybrid") -> Dcsc:
    """The genuine reference kernel (only where oracle/_ref was built)."""
    L = _load_ref()
    sa, ka = A._cstruct()
    sb, kb = B._cstruct()
    out = cbx_dcsc()
    st = L.ref_spgemm(C.byref(sa), C.byref(sb), SEMIRINGS[semiring], MODES[mode], C.byref(out))
    if st != 0:
        raise RuntimeError(f"ref_spgemm failed with status {st}")
    return Dcsc.from_struct(out, L.ref_free)


def ref_merge(parts: list[Dcsc], semiring: str = "plus_times") -> Dcsc:
    """The reference's MultiwayMerge<SR> (MultiwayMerge.h:428) over the parts'
    tuples, returned as DCSC (only where oracle/_ref was built)."""
    L = _load_ref()
    keep = [p._cstruct() for p in parts]
    arr = (cbx_dcsc * len(parts))(*[k[0] for k in keep])
    out = cbx_dcsc()
    st = L.ref_multiway_merge(len(parts), arr, SEMIRINGS[semiring], C.byref(out))
    if st != 0:
        raise RuntimeError(f"ref_multiway_merge failed with status {st}")
    return Dcsc.from_struct(out, L.ref_free)


def _combine(m: int, n: int, rows, cols, vals, dt: str, semiring: str) -> Dcsc:
    """Entries in column-major order with equal (row, col) combined by the
    semiring's add (stable: ties in input order), as DCSC."""
    rows = np.asarray(rows, np.int64)
    cols = np.asarray(cols, np.int64)
    order = np.lexsort((rows, cols))
    rows, cols, vals = rows[order], cols[order], vals[order]
    if rows.size == 0:
        return Dcsc(m, n, np.zeros(1, np.int64), np.zeros(0, np.int64), np.zeros(0, np.int64), vals, dt)
    head = np.ones(rows.size, bool)
    head[1:] = (rows[1:] != rows[:-1]) | (cols[1:] != cols[:-1])
    starts = np.nonzero(head)[0]
    if dt == "bool":
        v = np.logical_or.reduceat(vals.astype(bool), starts).astype(vals.dtype)
    elif semiring == "plus_times":
        if dt == "int64":  # two's-complement wrap, like the reference's int64 adds
            v = np.add.reduceat(vals.view(np.uint64), starts).view(np.int64)
        else:
            v = np.add.reduceat(vals, starts)
    elif semiring == "min_plus":
        v = np.minimum.reduceat(vals, starts)
    elif semiring == "select_max":
        v = np.maximum.reduceat(vals, starts)
    else:
        raise ValueError(semiring)
    r, c = rows[starts], cols[starts]
    jc, cnt = np.unique(c, return_counts=True)
    cp = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64)
    return Dcsc(m, n, cp, jc.astype(np.int64), r, v, dt)


def merge(parts: list[Dcsc], semiring: str = "plus_times") -> Dcsc:
    """Restatement of MultiwayMerge<SR> (MultiwayMerge.h:428-541, SerialMerge
    :193-250): the parts' entries in column-major order, equal (row, col)
    combined with SR::add, nothing dropped (explicit zeros stay).  Ties are
    added in part order; the reference pops them in heap order, which changes
    only floating-point rounding."""
    if not parts:
        raise ValueError("merge needs at least one part")
    m, n, dt = parts[0].m, parts[0].n, parts[0].dtype
    for p in parts:
        if (p.m, p.n, p.dtype) != (m, n, dt):
            raise ValueError("parts differ in shape or dtype")
    return _combine(m, n, np.concatenate([p.ir for p in parts]), np.concatenate([p.col_ids() for p in parts]),
                    np.concatenate([p.val for p in parts]), dt, semiring)


def from_coo(m: int, n: int, rows, cols, vals=None, dtype: str = "int64", semiring: str = "plus_times",
             remove_loops: bool = False) -> Dcsc:
    """Restatement of the Graph500 SpTuples constructor (SpTuples.cpp:70-110:
    values 1, duplicates summed, self loops dropped) generalised like
    SpTuples::RemoveDuplicates(BinOp) (:271): vals None means ones, equal
    (row, col) combined with the semiring's add, loops dropped on request."""
    from combblas_amd import NP_DTYPES
    rows = np.asarray(rows, np.int64)
    cols = np.asarray(cols, np.int64)
    vals = np.ones(rows.size, NP_DTYPES[dtype]) if vals is None else np.asarray(vals, NP_DTYPES[dtype])
    if np.any((rows < 0) | (rows >= m) | (cols < 0) | (cols >= n)):
        raise ValueError("entries outside the matrix")
    if remove_loops:
        keep = rows != cols
        rows, cols, vals = rows[keep], cols[keep], vals[keep]
    return _combine(m, n, rows, cols, vals, dtype, semiring)


def ref_from_coo_graph500(m: int, n: int, rows, cols, remove_loops: bool = True) -> Dcsc:
    """The reference's Graph500 constructor SpTuples(maxnnz, m, n, edges,
    removeloops) (SpTuples.cpp:70), as int64 DCSC (only where oracle/_ref was built)."""
    L = _load_ref()
    edges = np.empty(2 * len(rows), np.int64)
    edges[0::2] = rows
    edges[1::2] = cols
    out = cbx_dcsc()
    st = L.ref_graph500_tuples(int(m), int(n), edges.ctypes.data, len(rows), int(bool(remove_loops)), C.byref(out))
    if st != 0:
        raise RuntimeError(f"ref_graph500_tuples failed with status {st}")
    return Dcsc.from_struct(out, L.ref_free)


def ref_symbolic(A: Dcsc, B: Dcsc) -> tuple[np.ndarray, np.ndarray]:
    L = _load_ref()
    sa, ka = A._cstruct()
    sb, kb = B._cstruct()
    flop = np.zeros(B.nzc, np.int64)
    nnz = np.zeros(B.nzc, np.int64)
    L.ref_symbolic(C.byref(sa), C.byref(sb), flop.ctypes.data, nnz.ctypes.data)
    return flop, nnz


def estimate_sampling(A: Dcsc, B: Dcsc, nrounds: int = 5) -> tuple[np.ndarray, float, int]:
    """estimateNNZ_sampling (mtSpGEMM.h:936-1054) restated: (per stored column
    of B the min-hash estimate of nnz(C(:,j)), their float sum, the int64
    estimate the reference returns)."""
    L = _load_oracle()
    sa, ka = A._cstruct()
    sb, kb = B._cstruct()
    col = np.zeros(B.nzc, np.float32)
    tot, est = C.c_float(0), C.c_int64(0)
    st = L.orc_estimate_sampling(C.byref(sa), C.byref(sb), nrounds, col.ctypes.data, C.byref(tot), C.byref(est))
    if st != 0:
        raise RuntimeError(f"orc_estimate_sampling failed with status {st}")
    return col, float(tot.value), int(est.value)


def exp_samples(count: int) -> np.ndarray:
    """The first `count` exponential(1) draws of estimateNNZ_sampling's stream (float32)."""
    out = np.zeros(count, np.float32)
    _load_oracle().orc_exp_samples(count, out.ctypes.data)
    return out


def ref_estimate_sampling(A: Dcsc, B: Dcsc, nrounds: int = 5) -> int:
    """The genuine estimateNNZ_sampling's int64 estimate (only where oracle/_ref was built)."""
    L = _load_ref()
    sa, ka = A._cstruct()
    sb, kb = B._cstruct()
    return int(L.ref_estimate_sampling(C.byref(sa), C.byref(sb), nrounds))


def mcl_prune(A: Dcsc, thr: float = 1e-4, select: int = 1000, recover: int = 0, pct: float = 0.9,
              return_cut: bool = False):
    """MCLPruneRecoverySelect (ParFriends.h:185-345) restated on one process; A float32 or float64.
    With return_cut, also the per-stored-column threshold that the final PruneColumn applied."""
    L = _load_oracle()
    sa, ka = A._cstruct()
    out = cbx_dcsc()
    cut = np.zeros(A.nzc, np.float32 if A.dtype == "float32" else np.float64)
    st = L.orc_mcl_prune(C.byref(sa), thr, select, recover, pct, C.byref(out), cut.ctypes.data)
    if st != 0:
        raise RuntimeError(f"orc_mcl_prune failed with status {st}")
    X = Dcsc.from_struct(out, L.orc_free)
    return (X, cut) if return_cut else X


_VDT = {np.dtype(np.int64): 0, np.dtype(np.float64): 1, np.dtype(np.float32): 2}


def spmv_dense(A: Dcsc, x: np.ndarray, y: np.ndarray, semiring: str = "plus_times") -> np.ndarray:
    """dcsc_gespmv restated: returns y accumulated with A*x (y is copied, not modified)."""
    L = _load_oracle()
    sa, ka = A._cstruct()
    x = np.ascontiguousarray(x)
    out = np.array(y, dtype=x.dtype, copy=True)
    st = L.orc_spmv_dense(C.byref(sa), x.ctypes.data, out.ctypes.data, SEMIRINGS[semiring], _VDT[x.dtype])
    if st != 0:
        raise RuntimeError(f"orc_spmv_dense failed with status {st}")
    return out


def ref_spmv_dense(A: Dcsc, x: np.ndarray, y: np.ndarray, semiring: str = "plus_times") -> np.ndarray:
    """The genuine dcsc_gespmv (only where oracle/_ref was built)."""
    L = _load_ref()
    sa, ka = A._cstruct()
    x = np.ascontiguousarray(x)
    out = np.array(y, dtype=x.dtype, copy=True)
    st = L.ref_spmv_dense(C.byref(sa), x.ctypes.data, out.ctypes.data, SEMIRINGS[semiring])
    if st != 0:
        raise RuntimeError(f"ref_spmv_dense failed with status {st}")
    return out


def galerkin_direct(g: int) -> Dcsc:
    """R*A*R^T of the 27-point stencil under 2x2x2 aggregation, summed directly (no SpGEMM)."""
    L = _load_oracle()
    out = cbx_dcsc()
    st = L.orc_galerkin_direct(g, C.byref(out))
    if st != 0:
        raise RuntimeError(f"orc_galerkin_direct failed with status {st}")
    return Dcsc.from_struct(out, L.orc_free)


def transpose(X: Dcsc) -> Dcsc:
    """Host transpose (rows sorted within columns)."""
    cols = X.col_ids()
    return Dcsc.from_coo(X.n, X.m, cols, X.ir, X.val, X.dtype)


def ref_mcl_prune(A: Dcsc, thr: float = 1e-4, select: int = 1000, recover: int = 0, pct: float = 0.9) -> Dcsc:
    """The genuine MCLPruneRecoverySelect on a 1x1 process grid (only where oracle/_ref was built)."""
    L = _load_ref()
    sa, ka = A._cstruct()
    out = cbx_dcsc()
    st = L.ref_mcl_prune(C.byref(sa), thr, select, recover, pct, C.byref(out))
    if st != 0:
        raise RuntimeError(f"ref_mcl_prune failed with status {st}")
    return Dcsc.from_struct(out, L.ref_free)


# ------------------------------------------------------------------ checks
from combblas_amd.checks import mix64 as _mix, product_checksums  # noqa: E402,F401  (closed forms)


def column_digests(X: Dcsc) -> np.ndarray:
    """Order-sensitive digest of every stored column: sum over the column's
    entries t = 0, 1, ... of (t+1) * mix(row, value bits), modulo 2^64.
    Returns uint64[nzc]."""
    with np.errstate(over="ignore"):
        if X.nnz == 0:
            return np.zeros(X.nzc, np.uint64)
        vb = X.val
        if vb.dtype.itemsize == 8:
            bits = vb.view(np.uint64)
        elif vb.dtype.itemsize == 4:
            bits = vb.view(np.uint32).astype(np.uint64)
        else:
            bits = vb.astype(np.uint64)
        h = _mix(X.ir.astype(np.uint64) ^ _mix(bits))
        pos = np.arange(X.nnz, dtype=np.int64) - np.repeat(X.cp[:-1], np.diff(X.cp))
        h = h * (pos.astype(np.uint64) + np.uint64(1))
        return np.add.reduceat(h, X.cp[:-1]) if X.nzc else np.zeros(0, np.uint64)


_XDT = {np.dtype(np.int64): 0, np.dtype(np.float64): 1, np.dtype(np.float32): 2}


def spmspv(A: Dcsc, xind: np.ndarray, xval: np.ndarray, semiring: str = "select_max") -> tuple[np.ndarray, np.ndarray]:
    """Oracle y = A x (restated SpMXSpV); boolean A uses the boolean form."""
    L = _load_oracle()
    sa, ka = A._cstruct()
    xind = np.ascontiguousarray(xind, np.int32)
    xval = np.ascontiguousarray(xval)
    yi = np.zeros(A.m, np.int32)
    yv = np.zeros(A.m, xval.dtype)
    ny = L.orc_spmspv(C.byref(sa), xind.ctypes.data, xval.ctypes.data, xind.size, SEMIRINGS[semiring],
                      _XDT[xval.dtype], yi.ctypes.data, yv.ctypes.data)
    return yi[:ny], yv[:ny]


def bfs(A: Dcsc, root: int) -> tuple[np.ndarray, int]:
    """Oracle top-down BFS (TopDownBFS.cpp:442-448): (parents, iterations)."""
    L = _load_oracle()
    sa, ka = A._cstruct()
    parents = np.zeros(A.n, np.int64)
    it = L.orc_bfs(C.byref(sa), int(root), parents.ctypes.data)
    return parents, int(it)


def ref_spmspv(A: Dcsc, xind: np.ndarray, xval: np.ndarray, semiring: str) -> tuple[np.ndarray, np.ndarray]:
    """Genuine generic_gespmv: SelectMax on a boolean A, or on an A of the
    vector's type PlusTimes<int64>, MinPlus<int64> or PlusTimes<double>."""
    L = _load_ref()
    sa, ka = A._cstruct()
    xind = np.ascontiguousarray(xind, np.int32)
    fdt = A.dtype == "float64"
    xval = np.ascontiguousarray(xval, np.float64 if fdt else np.int64)
    yi = np.zeros(A.m, np.int32)
    yv = np.zeros(A.m, xval.dtype)
    if semiring == "select_max":
        fn = L.ref_spmspv_selectmax_bool
    elif semiring == "min_plus":
        fn = L.ref_spmspv_minplus_i64
    else:
        fn = L.ref_spmspv_plustimes_f64 if fdt else L.ref_spmspv_plustimes_i64
    ny = fn(C.byref(sa), xind.ctypes.data, xval.ctypes.data, xind.size, yi.ctypes.data, yv.ctypes.data)
    return yi[:ny], yv[:ny]


def ref_bfs(A: Dcsc, root: int) -> tuple[np.ndarray, int]:
    L = _load_ref()
    sa, ka = A._cstruct()
    parents = np.zeros(A.n, np.int64)
    it = L.ref_bfs(C.byref(sa), int(root), parents.ctypes.data)
    return parents, int(it)

