"""Host-side mirror of CombBLAS's node-local kernel interface over libcbx.so.

The product is the C ABI in ``include/cbx.h`` (built into
``combblas_amd/libcbx.so`` from ``combblas_amd/csrc``).  This module binds it
with ctypes so Python tests, the benchmark and users can call it with numpy
arrays:

* :class:`Dcsc` -- a host DCSC matrix (``cp, jc, ir, val`` as numpy arrays),
  the layout of ``combblas::Dcsc<int64_t, NT>`` (include/CombBLAS/dcsc.h:125-131).
* :class:`DeviceDcsc` -- the same arrays resident in device memory.
* :func:`LocalHybridSpGEMM` / :func:`LocalSpGEMMHash` -- the reference's
  local products (include/CombBLAS/mtSpGEMM.h:213-217, :463-467) on the GPU.
* :func:`spgemm` -- the device-resident form used by the benchmark.

There is no CPU fallback: every compute call goes to the HIP library and
raises :class:`CbxError` when it is missing or when no GPU is present.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libcbx.so"

# ---------------------------------------------------------------- constants
OK, ERR_INVALID, ERR_HIP, ERR_NOMEM, ERR_UNSUPPORTED, ERR_LIMIT = range(6)

SEMIRINGS = {"plus_times": 0, "min_plus": 1, "select_max": 2}
DTYPES = {"int64": 0, "float64": 1, "float32": 2, "bool": 3}
NP_DTYPES = {"int64": np.int64, "float64": np.float64, "float32": np.float32, "bool": np.uint8}
MODES = {"two_phase": 0, "unsorted": 1, "learned": 2}
HOST, DEVICE = 0, 1
GEN_SYMMETRIZE, GEN_NO_SELF_LOOPS, GEN_PATTERN = 1, 2, 4


class CbxError(RuntimeError):
    """Raised when a cbx call returns a non-zero status."""

    def __init__(self, status: int, msg: str):
        super().__init__(f"cbx status {status}: {msg}")
        self.status = status


class cbx_dcsc(C.Structure):
    _fields_ = [
        ("m", C.c_int64), ("n", C.c_int64), ("nnz", C.c_int64), ("nzc", C.c_int64),
        ("cp", C.c_void_p), ("jc", C.c_void_p), ("ir", C.c_void_p), ("val", C.c_void_p),
        ("dtype", C.c_int32), ("location", C.c_int32),
    ]


class cbx_spgemm_stats(C.Structure):
    _fields_ = [
        ("flops", C.c_int64), ("nnz", C.c_int64), ("nzc", C.c_int64),
        ("overflow_cols", C.c_int64), ("predicted_nnz", C.c_int64), ("actual_nnz_predicted", C.c_int64),
        ("learned_fallback", C.c_int64),
        ("units_light", C.c_int64), ("units_heavy", C.c_int64), ("nnz_main", C.c_int64),
        ("ms_setup", C.c_double), ("ms_symbolic", C.c_double), ("ms_alloc", C.c_double),
        ("ms_numeric", C.c_double), ("ms_main_kernel", C.c_double), ("ms_total", C.c_double),
        ("nnz_before_filter", C.c_int64),
    ]


class cbx_mlp_train_opts(C.Structure):
    _fields_ = [("steps", C.c_int32), ("batch", C.c_int32), ("lr", C.c_float), ("pad", C.c_int32),
                ("seed", C.c_uint64)]


MLP_FEATURES, MLP_HIDDEN = 6, 64
MLP_PARAMS = 64 * 6 + 64 + 64 * 64 + 64 + 64 + 1 + 12


class cbx_bfs_stats(C.Structure):
    _fields_ = [("levels", C.c_int32), ("bottom_up_levels", C.c_int32), ("reached", C.c_int64), ("ms", C.c_double)]


class cbx_mcl_stats(C.Structure):
    _fields_ = [(k, C.c_int64) for k in ("flops", "nnz_product", "nnz", "phases", "recovered_cols", "selected_cols",
                                         "reselected_cols", "overflow_cols", "learned_fallback")] + \
               [(k, C.c_double) for k in ("ms_product", "ms_prune", "ms_total")]


class cbx_spgemm_config(C.Structure):
    _fields_ = [("hash_max_flop", C.c_int64), ("hash_mid_flop", C.c_int64), ("tiny_flop", C.c_int64),
                ("heavy_nnz", C.c_int64), ("unit_split_nnz", C.c_int64), ("learned_target", C.c_float),
                ("capacity_margin", C.c_float), ("row_map_share", C.c_float), ("tiny_wg_waves", C.c_int32),
                ("unit_order", C.c_int32), ("tiny_algo", C.c_int32), ("sym_band", C.c_int32),
                ("mcl_fuse", C.c_int32), ("unit_lists", C.c_int32), ("unit_xcd_group", C.c_int32)]


_lib_handle = None


def lib() -> C.CDLL:
    """Load libcbx.so (built in-tree); raises if it is missing."""
    global _lib_handle
    if _lib_handle is None:
        if not LIB_PATH.exists():
            raise CbxError(ERR_UNSUPPORTED, f"{LIB_PATH} is not built; run `python -m combblas_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        P = C.POINTER
        L.cbx_last_error.restype = C.c_char_p
        L.cbx_device_count.restype = C.c_int32
        L.cbx_device_synchronize.restype = C.c_int32
        L.cbx_device_name.argtypes = [C.c_char_p, C.c_size_t]
        L.cbx_device_name.restype = C.c_int32
        L.cbx_release_cached_memory.restype = None
        L.cbx_dcsc_free.argtypes = [P(cbx_dcsc)]
        L.cbx_dcsc_free.restype = None
        L.cbx_dcsc_to_device.argtypes = [P(cbx_dcsc), P(cbx_dcsc)]
        L.cbx_dcsc_to_host.argtypes = [P(cbx_dcsc), P(cbx_dcsc)]
        L.cbx_gen_rmat.argtypes = [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                   C.c_uint32, P(cbx_dcsc)]
        L.cbx_gen_powerlaw_stochastic.argtypes = [C.c_int64, C.c_double, C.c_double, C.c_int64, C.c_uint64,
                                                  P(cbx_dcsc)]
        L.cbx_gen_powerlaw_stochastic.restype = C.c_int32
        L.cbx_spmv_dense.argtypes = [P(cbx_dcsc), C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]
        L.cbx_spmv_dense.restype = C.c_int32
        L.cbx_spgemm_get_config.argtypes = [P(cbx_spgemm_config)]
        L.cbx_spgemm_get_config.restype = C.c_int32
        L.cbx_spgemm_set_config.argtypes = [P(cbx_spgemm_config)]
        L.cbx_spgemm_set_config.restype = C.c_int32
        L.cbx_gen_stencil27.argtypes = [C.c_int32, P(cbx_dcsc)]
        L.cbx_gen_stencil27.restype = C.c_int32
        L.cbx_gen_aggregation.argtypes = [C.c_int32, C.c_int32, P(cbx_dcsc)]
        L.cbx_gen_aggregation.restype = C.c_int32
        L.cbx_dcsc_transpose.argtypes = [P(cbx_dcsc), P(cbx_dcsc), C.c_void_p]
        L.cbx_dcsc_merge.argtypes = [C.c_int32, P(cbx_dcsc), C.c_int32, P(cbx_dcsc), C.c_void_p]
        L.cbx_dcsc_merge.restype = C.c_int32
        L.cbx_dcsc_from_coo.argtypes = [C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int32,
                                        C.c_int32, C.c_int32, P(cbx_dcsc), C.c_void_p]
        L.cbx_dcsc_from_coo.restype = C.c_int32
        L.cbx_dcsc_transpose.restype = C.c_int32
        L.cbx_spgemm_batched.argtypes = [C.c_int64, P(cbx_dcsc), P(cbx_dcsc), C.c_int32, C.c_int32, P(cbx_dcsc),
                                         P(cbx_spgemm_stats), C.c_void_p]
        L.cbx_spgemm_batched.restype = C.c_int32
        L.cbx_galerkin_batched.argtypes = [C.c_int64, P(cbx_dcsc), P(cbx_dcsc), P(cbx_dcsc), P(cbx_spgemm_stats),
                                           C.c_void_p]
        L.cbx_galerkin_batched.restype = C.c_int32
        L.cbx_mcl_prune.argtypes = [P(cbx_dcsc), C.c_double, C.c_int64, C.c_int64, C.c_double, P(cbx_dcsc),
                                    P(cbx_mcl_stats), C.c_void_p]
        L.cbx_mcl_prune.restype = C.c_int32
        L.cbx_mcl_expand.argtypes = [P(cbx_dcsc), P(cbx_dcsc), C.c_double, C.c_int64, C.c_int64, C.c_double,
                                     C.c_int64, C.c_int32, P(cbx_dcsc), P(cbx_mcl_stats), C.c_void_p]
        L.cbx_mcl_expand.restype = C.c_int32
        L.cbx_spgemm.argtypes = [P(cbx_dcsc), P(cbx_dcsc), C.c_int32, C.c_int32, P(cbx_dcsc),
                                 P(cbx_spgemm_stats), C.c_void_p]
        L.cbx_dcsc_check.argtypes = [P(cbx_dcsc)]
        L.cbx_dcsc_structure_to_host.argtypes = [P(cbx_dcsc), C.c_void_p, C.c_void_p]
        L.cbx_dcsc_structure_to_host.restype = C.c_int32
        L.cbx_dcsc_column_checksums.argtypes = [P(cbx_dcsc), C.c_uint64, C.c_void_p]
        L.cbx_malloc.argtypes = [P(C.c_void_p), C.c_size_t]
        L.cbx_free.argtypes = [C.c_void_p]
        L.cbx_free.restype = None
        L.cbx_memcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int32]
        L.cbx_spmspv.argtypes = [P(cbx_dcsc), C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p,
                                 C.c_void_p, P(C.c_int32), C.c_void_p]
        L.cbx_spmspv_noalloc.argtypes = L.cbx_spmspv.argtypes
        L.cbx_spmspv_noalloc.restype = C.c_int32
        L.cbx_bfs_plan_create.argtypes = [P(cbx_dcsc), P(C.c_void_p)]
        L.cbx_bfs_run.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, P(cbx_bfs_stats)]
        L.cbx_bfs_plan_destroy.argtypes = [C.c_void_p]
        L.cbx_bfs_plan_destroy.restype = None
        L.cbx_bfs_plan_set_direction.argtypes = [C.c_void_p, C.c_int32]
        L.cbx_bfs_plan_set_direction.restype = C.c_int32
        L.cbx_spgemm_symbolic.argtypes = [P(cbx_dcsc), P(cbx_dcsc), C.c_void_p, C.c_void_p]
        L.cbx_spgemm_set_predictor.argtypes = [C.c_void_p]
        L.cbx_spgemm_set_predictor.restype = C.c_int32
        L.cbx_spgemm_features.argtypes = [P(cbx_dcsc), P(cbx_dcsc), C.c_void_p, C.c_void_p]
        L.cbx_mlp_forward.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
        L.cbx_mlp_gradient.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, P(C.c_float)]
        L.cbx_mlp_train.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, P(cbx_mlp_train_opts),
                                    C.c_void_p]
        L.cbx_spgemm_estimate_sampling.argtypes = [P(cbx_dcsc), P(cbx_dcsc), C.c_int32, C.c_void_p, P(C.c_double),
                                                   C.c_void_p]
        L.cbx_spgemm_estimate_sampling.restype = C.c_int32
        L.cbx_spgemm_set_sampling_predictor.argtypes = [C.c_int32]
        L.cbx_spgemm_set_sampling_predictor.restype = C.c_int32
        for name in ("cbx_spgemm_symbolic", "cbx_spgemm_features", "cbx_mlp_forward", "cbx_mlp_gradient",
                     "cbx_mlp_train"):
            getattr(L, name).restype = C.c_int32
        for name in ("cbx_dcsc_to_device", "cbx_dcsc_to_host", "cbx_gen_rmat", "cbx_spgemm", "cbx_dcsc_check",
                     "cbx_dcsc_column_checksums", "cbx_malloc", "cbx_memcpy", "cbx_spmspv", "cbx_bfs_plan_create",
                     "cbx_bfs_run"):
            getattr(L, name).restype = C.c_int32
        _lib_handle = L
    return _lib_handle


def _check(status: int) -> None:
    if status != OK:
        raise CbxError(status, lib().cbx_last_error().decode(errors="replace"))


def device_count() -> int:
    return int(lib().cbx_device_count())


def device_name() -> str:
    """Name and architecture of the current device (from the HIP runtime libcbx.so links)."""
    buf = C.create_string_buffer(256)
    _check(lib().cbx_device_name(buf, 256))
    return buf.value.decode(errors="replace")


def synchronize() -> None:
    _check(lib().cbx_device_synchronize())


def release_cached_memory() -> None:
    lib().cbx_release_cached_memory()


# ------------------------------------------------------------------ matrices
@dataclass
class Dcsc:
    """Host DCSC matrix: the arrays of combblas::Dcsc (dcsc.h:125-131)."""

    m: int
    n: int
    cp: np.ndarray  # int64[nzc+1]
    jc: np.ndarray  # int64[nzc]
    ir: np.ndarray  # int64[nnz]
    val: np.ndarray  # dtype[nnz]
    dtype: str = "int64"

    @property
    def nnz(self) -> int:
        return int(self.ir.shape[0])

    @property
    def nzc(self) -> int:
        return int(self.jc.shape[0])

    def astype(self, dtype: str) -> "Dcsc":
        v = self.val
        if dtype == "bool":
            v = (v != 0).astype(np.uint8)
        else:
            v = v.astype(NP_DTYPES[dtype])
        return Dcsc(self.m, self.n, self.cp, self.jc, self.ir, v, dtype)

    def _cstruct(self) -> tuple[cbx_dcsc, list]:
        arrs = [np.ascontiguousarray(self.cp, np.int64), np.ascontiguousarray(self.jc, np.int64),
                np.ascontiguousarray(self.ir, np.int64),
                np.ascontiguousarray(self.val, NP_DTYPES[self.dtype])]
        s = cbx_dcsc(self.m, self.n, self.nnz, self.nzc, *(a.ctypes.data for a in arrs),
                     DTYPES[self.dtype], HOST)
        return s, arrs

    def to_device(self) -> "DeviceDcsc":
        s, keep = self._cstruct()
        out = cbx_dcsc()
        _check(lib().cbx_dcsc_to_device(C.byref(s), C.byref(out)))
        del keep
        return DeviceDcsc(out)

    @staticmethod
    def from_struct(s: cbx_dcsc, free_fn) -> "Dcsc":
        """Copy a host cbx_dcsc into numpy arrays, then release it with free_fn."""
        dt = {v: k for k, v in DTYPES.items()}[s.dtype]
        nzc, nnz = s.nzc, s.nnz

        def arr(ptr, n, t):
            if n == 0 or not ptr:
                return np.zeros(0, t)
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(np.ctypeslib.as_ctypes_type(t))), (n,)).copy()

        cp = arr(s.cp, nzc + 1, np.int64) if s.cp else np.zeros(1, np.int64)
        out = Dcsc(int(s.m), int(s.n), cp, arr(s.jc, nzc, np.int64), arr(s.ir, nnz, np.int64),
                   arr(s.val, nnz, NP_DTYPES[dt]), dt)
        free_fn(C.byref(s))
        return out

    # ---- conversions used by tests
    def col_ids(self) -> np.ndarray:
        """Column index of every stored entry."""
        return np.repeat(self.jc, np.diff(self.cp))

    def column_range(self, lo: int, hi: int) -> "Dcsc":
        """The columns lo..hi-1 at the full shape (others empty): A(:, K) of a SUMMA stage."""
        keep = (self.jc >= lo) & (self.jc < hi)
        cnt = np.diff(self.cp)[keep]
        first = self.cp[:-1][keep]
        idx = (np.repeat(first - np.concatenate([[0], np.cumsum(cnt)[:-1]]), cnt) + np.arange(cnt.sum())
               if cnt.size else np.zeros(0, np.int64))
        return Dcsc(self.m, self.n, np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64), self.jc[keep].copy(),
                    self.ir[idx].copy(), self.val[idx].copy(), self.dtype)

    def row_range(self, lo: int, hi: int) -> "Dcsc":
        """The rows lo..hi-1 at the full shape (others empty): B(K, :) of a SUMMA stage."""
        sel = (self.ir >= lo) & (self.ir < hi)
        cols = self.col_ids()[sel]
        jc, cnt = np.unique(cols, return_counts=True)
        return Dcsc(self.m, self.n, np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64), jc.astype(np.int64),
                    self.ir[sel].copy(), self.val[sel].copy(), self.dtype)

    def to_dense(self) -> np.ndarray:
        d = np.zeros((self.m, self.n), dtype=NP_DTYPES[self.dtype])
        d[self.ir, self.col_ids()] = self.val
        return d

    @staticmethod
    def from_coo(m: int, n: int, rows, cols, vals, dtype: str = "int64") -> "Dcsc":
        """Build from unique (row, col) entries (sorted here by column, row)."""
        rows = np.asarray(rows, np.int64)
        cols = np.asarray(cols, np.int64)
        vals = np.asarray(vals)
        order = np.lexsort((rows, cols))
        rows, cols, vals = rows[order], cols[order], vals[order]
        if rows.size:
            key = cols * max(m, 1) + rows
            if np.any(np.diff(key) == 0):
                raise ValueError("duplicate entries")
        jc, starts = np.unique(cols, return_index=True)
        cp = np.append(starts, rows.size).astype(np.int64)
        if dtype == "bool":
            v = (vals != 0).astype(np.uint8)
        else:
            v = vals.astype(NP_DTYPES[dtype])
        return Dcsc(m, n, cp, jc.astype(np.int64), rows, v, dtype)

    @staticmethod
    def from_dense(d: np.ndarray, dtype: str = "int64", mask: np.ndarray | None = None) -> "Dcsc":
        """Entries where mask (default d != 0) holds; explicit zeros allowed via mask."""
        if mask is None:
            mask = d != 0
        r, c = np.nonzero(mask)
        return Dcsc.from_coo(d.shape[0], d.shape[1], r, c, d[r, c], dtype)


class DeviceDcsc:
    """A DCSC whose arrays live in device memory (owned, freed on release)."""

    def __init__(self, s: cbx_dcsc):
        self.s = s

    @property
    def m(self) -> int:
        return int(self.s.m)

    @property
    def n(self) -> int:
        return int(self.s.n)

    @property
    def nnz(self) -> int:
        return int(self.s.nnz)

    @property
    def nzc(self) -> int:
        return int(self.s.nzc)

    def to_host(self) -> Dcsc:
        out = cbx_dcsc()
        _check(lib().cbx_dcsc_to_host(C.byref(self.s), C.byref(out)))
        return Dcsc.from_struct(out, lib().cbx_dcsc_free)

    def check(self) -> None:
        """Raise CbxError unless the arrays form a valid DCSC (sorted rows)."""
        _check(lib().cbx_dcsc_check(C.byref(self.s)))

    def structure(self) -> tuple[np.ndarray, np.ndarray]:
        """(cp, jc) copied to the host without the row/value arrays."""
        cp = np.zeros(self.nzc + 1, np.int64)
        jc = np.zeros(self.nzc, np.int64)
        _check(lib().cbx_dcsc_structure_to_host(C.byref(self.s), cp.ctypes.data, jc.ctypes.data))
        return cp, jc

    def column_checksums(self, salt: int = 0) -> np.ndarray:
        """uint64[nzc]: sum_i mix64(i ^ salt) * bits(X(i,j)) per stored column."""
        out = np.zeros(self.nzc, np.uint64)
        _check(lib().cbx_dcsc_column_checksums(C.byref(self.s), salt, out.ctypes.data))
        return out

    def release(self) -> None:
        if self.s is not None and _lib_handle is not None:
            _lib_handle.cbx_dcsc_free(C.byref(self.s))
        self.s = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


# ---------------------------------------------------------------- generators
def rmat_seed(scale: int) -> int:
    """Seed of the R-MAT product workloads: 0xCB1A5 xor scale."""
    return 0xCB1A5 ^ scale


def gen_rmat(scale: int, edgefactor: int, a: float = 0.57, b: float = 0.19, c: float = 0.19,
             seed: int | None = None, symmetrize: bool = False, no_self_loops: bool = False,
             pattern: bool = False) -> Dcsc:
    """R-MAT matrix (values uniform in [1,9], duplicates summed) as host DCSC int64."""
    flags = (GEN_SYMMETRIZE if symmetrize else 0) | (GEN_NO_SELF_LOOPS if no_self_loops else 0) | \
        (GEN_PATTERN if pattern else 0)
    s = cbx_dcsc()
    _check(lib().cbx_gen_rmat(scale, edgefactor, a, b, c, rmat_seed(scale) if seed is None else seed, flags,
                              C.byref(s)))
    return Dcsc.from_struct(s, lib().cbx_dcsc_free)


def gen_powerlaw_stochastic(n: int, mean: float = 50.0, alpha: float = 2.1, dmax: int | None = None,
                            seed: int = 0x3C1) -> Dcsc:
    """Column-stochastic power-law matrix (configuration model, log-normal weights) as host DCSC float32.

    dmax defaults to the structural cut-off sqrt(n * mean) of an uncorrelated configuration model."""
    if dmax is None:
        dmax = max(int(mean) + 1, int((n * mean) ** 0.5))
    s = cbx_dcsc()
    _check(lib().cbx_gen_powerlaw_stochastic(n, mean, alpha, dmax, seed, C.byref(s)))
    return Dcsc.from_struct(s, lib().cbx_dcsc_free)


# -------------------------------------------------------------------- SpGEMM
def spgemm(A: DeviceDcsc, B: DeviceDcsc, semiring: str = "plus_times", mode: str = "two_phase",
           stream: int | None = None) -> tuple[DeviceDcsc, dict]:
    """C = A*B on device-resident DCSC; returns (C on device, stats)."""
    out = cbx_dcsc()
    st = cbx_spgemm_stats()
    _check(lib().cbx_spgemm(C.byref(A.s), C.byref(B.s), SEMIRINGS[semiring], MODES[mode], C.byref(out),
                            C.byref(st), stream))
    stats = {k: getattr(st, k) for k, _ in cbx_spgemm_stats._fields_}
    return DeviceDcsc(out), stats


_VEC_DTYPES = {np.dtype(np.int64): DTYPES["int64"], np.dtype(np.float64): DTYPES["float64"],
               np.dtype(np.float32): DTYPES["float32"]}


def spmv_dense_device(A: DeviceDcsc, x: "DeviceArray", y: "DeviceArray", semiring: str = "plus_times") -> None:
    """dcsc_gespmv on device-resident vectors: y += A*x in place (x: A.n entries, y: A.m, one dtype)."""
    if x.dtype != y.dtype or x.dtype not in _VEC_DTYPES:
        raise ValueError("x and y must share an int64, float64 or float32 dtype")
    if x.n != A.n or y.n != A.m:
        raise ValueError(f"vector lengths {x.n}, {y.n} do not match A ({A.m} x {A.n})")
    _check(lib().cbx_spmv_dense(C.byref(A.s), x.ptr, y.ptr, SEMIRINGS[semiring], _VEC_DTYPES[x.dtype], None))


def spmv_dense(A: DeviceDcsc, x: np.ndarray, y: np.ndarray, semiring: str = "plus_times") -> np.ndarray:
    """dcsc_gespmv on the device: returns y accumulated with A*x (host arrays in and out)."""
    x = np.ascontiguousarray(x)
    dx = DeviceArray.from_host(x)
    dy = DeviceArray.from_host(np.ascontiguousarray(y, dtype=x.dtype))
    try:
        spmv_dense_device(A, dx, dy, semiring)
        return dy.to_host()
    finally:
        dx.release()
        dy.release()


def get_config() -> dict:
    """The process-wide SpGEMM tunables (bin thresholds, hash-table classes, unit sizes)."""
    c = cbx_spgemm_config()
    _check(lib().cbx_spgemm_get_config(C.byref(c)))
    return {k: getattr(c, k) for k, _ in cbx_spgemm_config._fields_}


def set_config(**kw) -> dict:
    """Updates some tunables (validated by the library); returns the full new configuration."""
    cur = get_config()
    unknown = set(kw) - set(cur)
    if unknown:
        raise ValueError(f"unknown tunables: {sorted(unknown)}")
    cur.update(kw)
    _check(lib().cbx_spgemm_set_config(C.byref(cbx_spgemm_config(**cur))))
    return cur


def gen_stencil27(g: int) -> Dcsc:
    """27-point Poisson stencil on a g^3 grid (FP64), as in GalerkinNew.cpp's grid3d input."""
    s = cbx_dcsc()
    _check(lib().cbx_gen_stencil27(g, C.byref(s)))
    return Dcsc.from_struct(s, lib().cbx_dcsc_free)


def gen_aggregation(g: int, transposed: bool = True) -> Dcsc:
    """2x2x2 aggregation of a g^3 grid: T = R^T (fine x coarse) or R (coarse x fine), FP64."""
    s = cbx_dcsc()
    _check(lib().cbx_gen_aggregation(g, 1 if transposed else 0, C.byref(s)))
    return Dcsc.from_struct(s, lib().cbx_dcsc_free)


def transpose(X: DeviceDcsc, stream: int | None = None) -> DeviceDcsc:
    """SpDCCols::Transpose on the device."""
    out = cbx_dcsc()
    _check(lib().cbx_dcsc_transpose(C.byref(X.s), C.byref(out), stream))
    return DeviceDcsc(out)


def merge(parts: list[DeviceDcsc], semiring: str = "plus_times", stream: int | None = None) -> DeviceDcsc:
    """MultiwayMerge<SR> on the device: the semiring sum of same-shape parts (SUMMA's stage merge)."""
    if not parts:
        raise ValueError("merge needs at least one part")
    out = cbx_dcsc()
    _check(lib().cbx_dcsc_merge(len(parts), _desc_array(parts), SEMIRINGS[semiring], C.byref(out), stream))
    return DeviceDcsc(out)


def from_coo_device(m: int, n: int, rows: "DeviceArray", cols: "DeviceArray", vals: "DeviceArray | None" = None,
                    dtype: str = "int64", semiring: str = "plus_times", remove_loops: bool = False,
                    stream: int | None = None) -> DeviceDcsc:
    """Edge list to DCSC on the device (SpTuples' Graph500 constructor / RemoveDuplicates):
    duplicates combined with the semiring's add, optional self-loop removal, vals None = ones."""
    if rows.n != cols.n or (vals is not None and vals.n != rows.n):
        raise ValueError("rows, cols and vals must have the same length")
    out = cbx_dcsc()
    _check(lib().cbx_dcsc_from_coo(int(m), int(n), rows.ptr, cols.ptr, vals.ptr if vals is not None else None,
                                   rows.n, DTYPES[dtype], SEMIRINGS[semiring], int(bool(remove_loops)),
                                   C.byref(out), stream))
    return DeviceDcsc(out)


def _desc_array(mats) -> C.Array:
    arr = (cbx_dcsc * len(mats))()
    for i, M in enumerate(mats):
        arr[i] = M.s
    return arr


def _stats(st: cbx_spgemm_stats) -> dict:
    return {k: getattr(st, k) for k, _ in cbx_spgemm_stats._fields_}


def spgemm_batched(As, Bs, semiring: str = "plus_times", mode: str = "two_phase",
                   stream: int | None = None) -> tuple[list[DeviceDcsc], dict]:
    """C[i] = A[i]*B[i] for a batch of device matrices, computed as one packed product."""
    n = len(As)
    out = (cbx_dcsc * n)()
    st = cbx_spgemm_stats()
    _check(lib().cbx_spgemm_batched(n, _desc_array(As), _desc_array(Bs), SEMIRINGS[semiring], MODES[mode], out,
                                    C.byref(st), stream))
    return [DeviceDcsc(out[i]) for i in range(n)], _stats(st)


def galerkin_batched(As, Ts, stream: int | None = None) -> tuple[list[DeviceDcsc], dict]:
    """SAT[i] = T[i]^T * (A[i] * T[i]) for a batch (GalerkinNew.cpp:105-106)."""
    n = len(As)
    out = (cbx_dcsc * n)()
    st = cbx_spgemm_stats()
    _check(lib().cbx_galerkin_batched(n, _desc_array(As), _desc_array(Ts), out, C.byref(st), stream))
    return [DeviceDcsc(out[i]) for i in range(n)], _stats(st)


def mcl_prune(C_: DeviceDcsc, thr: float = 1e-4, select: int = 1000, recover: int = 0, pct: float = 0.9,
              stream: int | None = None) -> tuple[DeviceDcsc, dict]:
    """MCLPruneRecoverySelect (ParFriends.h:185-345) on a device matrix; returns (pruned copy, stats)."""
    out = cbx_dcsc()
    st = cbx_mcl_stats()
    _check(lib().cbx_mcl_prune(C.byref(C_.s), thr, select, recover, pct, C.byref(out), C.byref(st), stream))
    return DeviceDcsc(out), {k: getattr(st, k) for k, _ in cbx_mcl_stats._fields_}


def mcl_expand(A: DeviceDcsc, B: DeviceDcsc | None = None, thr: float = 1e-4, select: int = 1000, recover: int = 0,
               pct: float = 0.9, phases: int = 0, mode: str = "two_phase",
               stream: int | None = None) -> tuple[DeviceDcsc, dict]:
    """One MCL expansion: A*B (B defaults to A) in column phases, each pruned as cbx_mcl_prune."""
    out = cbx_dcsc()
    st = cbx_mcl_stats()
    B = A if B is None else B
    _check(lib().cbx_mcl_expand(C.byref(A.s), C.byref(B.s), thr, select, recover, pct, phases, MODES[mode],
                                C.byref(out), C.byref(st), stream))
    return DeviceDcsc(out), {k: getattr(st, k) for k, _ in cbx_mcl_stats._fields_}


def LocalHybridSpGEMM(A: Dcsc, B: Dcsc, semiring: str = "plus_times") -> Dcsc:
    """GPU counterpart of combblas::LocalHybridSpGEMM<SR,NT>(A, B, false, false)
    (mtSpGEMM.h:213-460): the product with rows sorted within each column,
    returned in DCSC like SpDCCols(SpTuples) (SpDCCols.cpp:109-184)."""
    dA, dB = A.to_device(), B.to_device()
    dC, _ = spgemm(dA, dB, semiring, "two_phase")
    return dC.to_host()


def LocalSpGEMMHash(A: Dcsc, B: Dcsc, semiring: str = "plus_times", sort: bool = True) -> Dcsc:
    """GPU counterpart of combblas::LocalSpGEMMHash (mtSpGEMM.h:463-656);
    with sort=False the rows of a column come in no particular order."""
    dA, dB = A.to_device(), B.to_device()
    dC, _ = spgemm(dA, dB, semiring, "two_phase" if sort else "unsorted")
    return dC.to_host()


# ------------------------------------------------------------------ vectors
class DeviceArray:
    """A device buffer from the library's cache holding a 1-D numpy-typed array."""

    def __init__(self, n: int, dtype):
        self.n, self.dtype = int(n), np.dtype(dtype)
        p = C.c_void_p()
        _check(lib().cbx_malloc(C.byref(p), max(self.n * self.dtype.itemsize, 8)))
        self.ptr = p.value

    @staticmethod
    def from_host(a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = DeviceArray(a.size, a.dtype)
        if a.size:
            _check(lib().cbx_memcpy(d.ptr, a.ctypes.data, a.nbytes, 0))
        return d

    def to_host(self, n: int | None = None) -> np.ndarray:
        n = self.n if n is None else n
        out = np.empty(n, self.dtype)
        if n:
            _check(lib().cbx_memcpy(out.ctypes.data, self.ptr, n * self.dtype.itemsize, 1))
        return out

    def release(self) -> None:
        if self.ptr and _lib_handle is not None:
            _lib_handle.cbx_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


_XDT = {np.dtype(np.int64): "int64", np.dtype(np.float64): "float64", np.dtype(np.float32): "float32"}


def spmspv(A: DeviceDcsc, xind: np.ndarray, xval: np.ndarray, semiring: str = "select_max",
           noalloc: bool = False) -> tuple[np.ndarray, np.ndarray]:
    """y = A (x) x for a sparse x (SpMXSpV, SpImpl.cpp:55-213); returns (yind, yval) on the host.
    noalloc=True calls cbx_spmspv_noalloc (scratch in the library's static
    per-GPU array instead of allocations; same result)."""
    xind = np.ascontiguousarray(xind, np.int32)
    xval = np.ascontiguousarray(xval)
    dt = _XDT[xval.dtype]
    dx, dv = DeviceArray.from_host(xind), DeviceArray.from_host(xval)
    dyi, dyv = DeviceArray(A.m, np.int32), DeviceArray(A.m, xval.dtype)
    ny = C.c_int32(0)
    fn = lib().cbx_spmspv_noalloc if noalloc else lib().cbx_spmspv
    _check(fn(C.byref(A.s), dx.ptr, dv.ptr, xind.size, SEMIRINGS[semiring], DTYPES[dt], dyi.ptr, dyv.ptr,
              C.byref(ny), None))
    return dyi.to_host(ny.value), dyv.to_host(ny.value)


BFS_DIRECTIONS = {"optimizing": 0, "top_down": 1, "bottom_up": 2}


class BfsPlan:
    """BFS plan over a device adjacency matrix (square, symmetric): TopDownBFS's
    parents, with each level run top-down or bottom-up (same result)."""

    def __init__(self, A: DeviceDcsc, direction: str = "optimizing"):
        h = C.c_void_p()
        _check(lib().cbx_bfs_plan_create(C.byref(A.s), C.byref(h)))
        self.h = h.value
        self.n = A.n
        self.parents = DeviceArray(A.n, np.int64)
        self.set_direction(direction)

    def set_direction(self, direction: str) -> None:
        """"optimizing" (per level, the default), "top_down" or "bottom_up"."""
        _check(lib().cbx_bfs_plan_set_direction(self.h, BFS_DIRECTIONS[direction]))

    def run(self, root: int, copy: bool = True) -> tuple[np.ndarray | None, dict]:
        st = cbx_bfs_stats()
        _check(lib().cbx_bfs_run(self.h, int(root), self.parents.ptr, C.byref(st)))
        stats = {"levels": st.levels, "bottom_up_levels": st.bottom_up_levels, "reached": st.reached, "ms": st.ms}
        return (self.parents.to_host() if copy else None), stats

    def release(self) -> None:
        if self.h and _lib_handle is not None:
            _lib_handle.cbx_bfs_plan_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def bfs_roots(A: Dcsc, count: int, seed: int = 4242) -> np.ndarray:
    """`count` distinct BFS roots drawn uniformly (seeded) among the vertices
    of degree > 0 (the reference draws its roots among non-isolated vertices,
    TopDownBFS.cpp:380-411)."""
    deg = np.zeros(A.n, np.int64)
    deg[A.jc] = np.diff(A.cp)
    cand = np.nonzero(deg > 0)[0]
    rng = np.random.default_rng(seed)
    return rng.choice(cand, size=min(count, cand.size), replace=False)


# ------------------------------------------------------- learned predictor
def spgemm_symbolic(A: DeviceDcsc, B: DeviceDcsc) -> np.ndarray:
    """Exact nnz of C(:,j) for every stored column j of B (the training labels)."""
    d = DeviceArray(max(B.nzc, 1), np.int64)
    _check(lib().cbx_spgemm_symbolic(C.byref(A.s), C.byref(B.s), d.ptr, None))
    return d.to_host(B.nzc)


def spgemm_estimate_sampling(A: DeviceDcsc, B: DeviceDcsc, nrounds: int = 5) -> tuple[np.ndarray, float]:
    """estimateNNZ_sampling per column on the device: (float32[nzc_B] estimates, their sum)."""
    d = DeviceArray(max(B.nzc, 1), np.float32)
    tot = C.c_double(0)
    _check(lib().cbx_spgemm_estimate_sampling(C.byref(A.s), C.byref(B.s), nrounds, d.ptr, C.byref(tot), None))
    return d.to_host(B.nzc), float(tot.value)


def set_sampling_predictor(nrounds: int) -> None:
    """Learned mode sized by estimateNNZ_sampling's per-column estimates (nrounds in [2, 8]) instead of
    the network; 0 goes back to the network."""
    _check(lib().cbx_spgemm_set_sampling_predictor(int(nrounds)))


def spgemm_features(A: DeviceDcsc, B: DeviceDcsc) -> np.ndarray:
    """float32[nzc_B, 6] predictor features (see cbx_spgemm_features)."""
    d = DeviceArray(max(B.nzc * MLP_FEATURES, 1), np.float32)
    _check(lib().cbx_spgemm_features(C.byref(A.s), C.byref(B.s), d.ptr, None))
    return d.to_host(B.nzc * MLP_FEATURES).reshape(B.nzc, MLP_FEATURES)


def mlp_forward(params: np.ndarray, X: np.ndarray) -> np.ndarray:
    """Predicted log2(1 + nnz) for feature rows X (on the device)."""
    X = np.ascontiguousarray(X, np.float32)
    dp, dx = DeviceArray.from_host(np.ascontiguousarray(params, np.float32)), DeviceArray.from_host(X)
    do = DeviceArray(max(len(X), 1), np.float32)
    _check(lib().cbx_mlp_forward(dp.ptr, dx.ptr, len(X), do.ptr, None))
    return do.to_host(len(X))


def mlp_gradient(params: np.ndarray, X: np.ndarray, y: np.ndarray) -> tuple[np.ndarray, float]:
    """Full-batch MSE gradient and loss from the device kernels."""
    dp = DeviceArray.from_host(np.ascontiguousarray(params, np.float32))
    dx = DeviceArray.from_host(np.ascontiguousarray(X, np.float32))
    dy = DeviceArray.from_host(np.ascontiguousarray(y, np.float32))
    dg = DeviceArray(MLP_PARAMS, np.float32)
    loss = C.c_float(0)
    _check(lib().cbx_mlp_gradient(dp.ptr, dx.ptr, dy.ptr, len(X), dg.ptr, C.byref(loss)))
    return dg.to_host(), float(loss.value)


def mlp_train(params: np.ndarray, X: np.ndarray, y: np.ndarray, steps: int, batch: int, lr: float,
              seed: int = 7) -> tuple[np.ndarray, np.ndarray]:
    """Adam training on the device; returns (trained params, per-step loss)."""
    dp = DeviceArray.from_host(np.ascontiguousarray(params, np.float32))
    dx = DeviceArray.from_host(np.ascontiguousarray(X, np.float32))
    dy = DeviceArray.from_host(np.ascontiguousarray(y, np.float32))
    curve = np.zeros(steps, np.float32)
    o = cbx_mlp_train_opts(steps, batch, lr, 0, seed)
    _check(lib().cbx_mlp_train(dp.ptr, dx.ptr, dy.ptr, len(X), C.byref(o), curve.ctypes.data))
    return dp.to_host(), curve


def set_predictor(params: np.ndarray | None) -> None:
    """Install (or remove, with None) the nnz predictor used by mode="learned"."""
    if params is None:
        _check(lib().cbx_spgemm_set_predictor(None))
        return
    p = np.ascontiguousarray(params, np.float32)
    if p.size != MLP_PARAMS:
        raise ValueError(f"expected {MLP_PARAMS} parameters, got {p.size}")
    _check(lib().cbx_spgemm_set_predictor(p.ctypes.data))
