"""The C ABI library loads, exports every symbol include/cbx.h declares, and
rejects bad arguments before touching a device (CPU-only checks)."""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import combblas_amd as cb

HEADER = Path(__file__).resolve().parents[1] / "include" / "cbx.h"


def declared_functions() -> list[str]:
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(cbx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = cb.lib()
    names = declared_functions()
    assert len(names) >= 8
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, f"missing exports: {missing}"


def test_device_count_never_fails():
    assert cb.device_count() >= 0


def test_null_and_host_arguments_rejected():
    L = cb.lib()
    out = cb.cbx_dcsc()
    assert L.cbx_spgemm(None, None, 0, 0, C.byref(out), None, None) == cb.ERR_INVALID
    A = cb.gen_rmat(6, 4)
    s, keep = A._cstruct()  # host-resident struct: cbx_spgemm wants device memory
    assert L.cbx_spgemm(C.byref(s), C.byref(s), 0, 0, C.byref(out), None, None) == cb.ERR_INVALID
    assert b"device" in L.cbx_last_error()


def test_generator_arguments_checked():
    with pytest.raises(cb.CbxError):
        cb.gen_rmat(0, 4)
    with pytest.raises(cb.CbxError):
        cb.gen_rmat(5, 4, a=0.9, b=0.2, c=0.2)


def test_generator_rmat_shape():
    A = cb.gen_rmat(12, 8)
    assert A.m == A.n == 4096
    assert np.all(np.diff(A.cp) > 0) and np.all(np.diff(A.jc) > 0)
    cols = A.col_ids()
    key = cols * A.m + A.ir
    assert np.all(np.diff(key) > 0), "rows strictly increasing within columns"
    assert A.val.min() >= 1


def test_tunables_validated_without_gpu():
    base = cb.get_config()
    try:
        assert base["unit_split_nnz"] == 131072 and base["heavy_nnz"] == 65536
        assert abs(base["row_map_share"] - 0.25) < 1e-6 and base["tiny_wg_waves"] == 4 and base["unit_order"] == 1
        assert base["unit_lists"] == 1 and base["mcl_fuse"] == 1 and base["sym_band"] == 0
        assert base["unit_xcd_group"] == 0
        got = cb.set_config(hash_max_flop=2048, heavy_nnz=131072, unit_split_nnz=262144)
        assert got["hash_max_flop"] == 2048 and cb.get_config()["heavy_nnz"] == 131072
        assert cb.get_config()["unit_split_nnz"] == 262144
        for bad in ({"tiny_flop": 500}, {"hash_mid_flop": 2000}, {"hash_max_flop": 8192}, {"heavy_nnz": 10},
                    {"unit_split_nnz": 100}, {"learned_target": 0.0}, {"learned_target": 5.0},
                    {"capacity_margin": -1.0}, {"row_map_share": 1.5}, {"row_map_share": -0.1},
                    {"tiny_wg_waves": 3}, {"tiny_wg_waves": 0}, {"tiny_wg_waves": 8}, {"unit_order": 2},
                    {"unit_order": -1}, {"tiny_algo": 2}, {"sym_band": 32},
                    {"mcl_fuse": 2}, {"unit_lists": 2}, {"unit_lists": -1}, {"unit_xcd_group": 3},
                    {"unit_xcd_group": 256}, {"unit_xcd_group": -8}):
            with pytest.raises(cb.CbxError):
                cb.set_config(**bad)
        with pytest.raises(ValueError):
            cb.set_config(no_such_knob=1)
    finally:
        cb.set_config(**base)
    assert cb.get_config() == base


def test_stored_tuning_table_is_current():
    """Every stored choice names exactly the library's tunables, passes its
    validation, is the fastest same-output point of its own sweep, and is
    what autotune.apply installs."""
    from combblas_amd import autotune
    table = autotune.load_table()
    assert table["version"] == autotune.TABLE_VERSION and table["entries"]
    base = cb.get_config()
    try:
        for name, e in table["entries"].items():
            assert name in autotune.WORKLOADS
            assert set(e["best"]) == set(base), name
            ok = [r for r in e["sweep"] if r["same_nnz"]]
            fastest = min(ok, key=lambda r: r["ms"])
            assert all(e["best"][k] == v for k, v in fastest["config"].items()), name
            assert abs(e["best_ms"] - fastest["ms"]) < 1e-9
            cb.set_config(**base)
            got = autotune.apply(name)
            assert got is not None
            for k, v in e["best"].items():
                assert abs(cb.get_config()[k] - v) <= 1e-6 * max(1.0, abs(v)), (name, k)
    finally:
        cb.set_config(**base)
