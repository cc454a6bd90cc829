"""GPU parity of cbx_spgemm (through the C ABI in libcbx.so) against the
oracle and the golden fixtures written from the genuine reference.

Integer, boolean, min-plus and select-max results must be bit-exact; FP64
and FP32 plus-times must lie within the a-priori bound
    |c_hat - c| <= gamma_k(u) * (|A| |B|)_ij,   gamma_k = k u / (1 - k u),
with k the column's flop count (the length of the longest possible sum),
u = 2^-53 for FP64 and 2^-24 for FP32 (FP32 compared against an FP64 oracle
of the same FP32 inputs).  For positive inputs this is within 1e-12 relative
per entry for FP64 at every k below 4.5e3.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import combblas_amd as cb
import oracle
from tests.helpers import bit_exact, random_dcsc, same_pattern, sort_columns

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"

SR_DTYPES = [("plus_times", "int64"), ("plus_times", "float64"), ("plus_times", "float32"),
             ("plus_times", "bool"), ("min_plus", "int64"), ("min_plus", "float64"), ("min_plus", "float32"),
             ("select_max", "int64"), ("select_max", "float64"), ("select_max", "float32")]


def gpu_spgemm(A, B, semiring="plus_times", mode="two_phase"):
    dA, dB = A.to_device(), B.to_device()
    dC, stats = cb.spgemm(dA, dB, semiring, mode)
    return dC.to_host(), stats


def check_fp(got: cb.Dcsc, A: cb.Dcsc, B: cb.Dcsc) -> None:
    """Pattern exact, values within gamma_k * (|A||B|) against an FP64 oracle."""
    A64, B64 = A.astype("float64"), B.astype("float64")
    want, _ = oracle.spgemm(A64, B64, "plus_times")
    same_pattern(got, want)
    absA = cb.Dcsc(A64.m, A64.n, A64.cp, A64.jc, A64.ir, np.abs(A64.val), "float64")
    absB = cb.Dcsc(B64.m, B64.n, B64.cp, B64.jc, B64.ir, np.abs(B64.val), "float64")
    mag, _ = oracle.spgemm(absA, absB, "plus_times")
    same_pattern(mag, want)
    k = np.repeat(_col_flops(A64, B64, want), np.diff(want.cp))
    u = 2.0 ** -53 if got.dtype == "float64" else 2.0 ** -24
    ku = (k + 1) * u
    bound = ku / (1 - ku) * mag.val + 1e-300
    err = np.abs(got.val.astype(np.float64) - want.val)
    assert np.all(err <= bound), f"max err/bound = {np.max(err / bound):.3g}"


def _col_flops(A, B, C):
    flop, _ = oracle.symbolic(A, B)
    # map B's stored columns to C's stored columns
    pos = np.searchsorted(B.jc, C.jc)
    return flop[pos]


def signed(X: cb.Dcsc, rng) -> cb.Dcsc:
    Y = X.astype(X.dtype)
    if X.dtype.startswith("float"):
        Y.val = rng.standard_normal(X.nnz).astype(Y.val.dtype)
    elif X.dtype == "int64":
        Y.val = rng.integers(-9, 10, X.nnz).astype(np.int64)
    return Y


def test_device_present():
    assert cb.device_count() >= 1, "no GPU visible to libcbx.so"


def test_c1_rmat_s14_ef8_matches_golden():
    """Config 1: R-MAT s14 ef8, PlusTimes<int64>, two-phase, bit-exact."""
    g = np.load(GOLDEN / "spgemm_rmat_s14_ef8_plus_times_int64.npz")
    A = cb.gen_rmat(14, 8)
    got, stats = gpu_spgemm(A, A)
    assert stats["flops"] == int(g["flops"])
    assert got.nnz == int(g["nnz_c"]) and stats["nnz"] == got.nnz
    assert np.array_equal(got.jc, g["jc"])
    assert np.array_equal(np.diff(got.cp), g["col_nnz"])
    assert np.array_equal(oracle.column_digests(got), g["col_digest"])
    want, _ = oracle.spgemm(A, A)
    bit_exact(got, want)


@pytest.mark.parametrize("semiring,dtype", SR_DTYPES)
def test_semirings_rmat(semiring, dtype):
    rng = np.random.default_rng(11)
    A = signed(cb.gen_rmat(12, 8).astype(dtype), rng)
    B = signed(cb.gen_rmat(12, 8, seed=5).astype(dtype), rng)
    got, _ = gpu_spgemm(A, B, semiring)
    if semiring == "plus_times" and dtype.startswith("float"):
        check_fp(got, A, B)
    else:
        want, _ = oracle.spgemm(A, B, semiring)
        bit_exact(got, want)


@pytest.mark.parametrize("semiring,dtype", SR_DTYPES)
@pytest.mark.parametrize("waves", [1, 4])
def test_tiny_tier_lds_table(semiring, dtype, waves):
    """The tiny tier's LDS-table form (tiny_algo 1: per-wave hash table,
    ranks by counting) in two workgroup shapes, on products whose columns
    hold up to 256 products: the staged two-phase numeric pass and, through
    the learned mode's exact small columns, the symbolic form, against the
    oracle (integer semirings bit-exact, FP plus-times within the bound)."""
    rng = np.random.default_rng(23)
    A = signed(cb.gen_rmat(12, 4).astype(dtype), rng)
    B = signed(cb.gen_rmat(12, 4, seed=9).astype(dtype), rng)
    want = None if semiring == "plus_times" and dtype.startswith("float") else oracle.spgemm(A, B, semiring)[0]
    base = cb.get_config()
    try:
        cb.set_config(tiny_algo=1, tiny_wg_waves=waves)
        got, st = gpu_spgemm(A, B, semiring)
        P = np.zeros(cb.MLP_PARAMS, np.float32)
        P[-6:] = 1.0  # feature scale: a constant prediction (learned mode's tiny columns are exact either way)
        cb.set_predictor(P)
        got_l, st_l = gpu_spgemm(A, B, semiring, "learned")
    finally:
        cb.set_predictor(None)
        cb.set_config(**base)
    for C in (got, got_l):
        if want is None:
            check_fp(C, A, B)
        else:
            bit_exact(C, want)


@pytest.mark.parametrize("name", sorted(p.name for p in GOLDEN.glob("spgemm_tiny_*.npz")))
def test_tiny_golden(name):
    g = np.load(GOLDEN / name)
    semiring, dtype, mode = str(g["semiring"]), str(g["dtype"]), str(g["mode"])
    A = cb.Dcsc(int(g["a_m"]), int(g["a_n"]), g["a_cp"], g["a_jc"], g["a_ir"], g["a_val"], dtype)
    B = cb.Dcsc(int(g["b_m"]), int(g["b_n"]), g["b_cp"], g["b_jc"], g["b_ir"], g["b_val"], dtype)
    want = cb.Dcsc(A.m, B.n, g["c_cp"], g["c_jc"], g["c_ir"], g["c_val"], dtype)
    got, _ = gpu_spgemm(A, B, semiring, "two_phase" if mode == "hybrid" else "unsorted")
    if mode != "hybrid":
        got, want = sort_columns(got), sort_columns(want)
    if semiring == "plus_times" and dtype.startswith("float"):
        same_pattern(got, want)
        check_fp(got, A, B)
    else:
        bit_exact(got, want)


def test_dense_tier_multiple_row_tiles():
    """s16: 65536 rows = 4 row tiles of the dense tier, heavy columns."""
    A = cb.gen_rmat(16, 8)
    got, stats = gpu_spgemm(A, A)
    want, fl = oracle.spgemm(A, A)
    assert stats["flops"] == fl
    bit_exact(got, want)


def test_tall_matrix_multiple_symbolic_passes():
    """3M rows: 184 numeric row tiles and 3 symbolic bitmap passes."""
    rng = np.random.default_rng(3)
    m, k, n = 3_000_000, 400, 150
    per_col = 160
    rows = np.concatenate([np.sort(rng.choice(m, per_col, replace=False)) for _ in range(k)])
    cols = np.repeat(np.arange(k), per_col)
    A = cb.Dcsc.from_coo(m, k, rows, cols, rng.integers(1, 10, rows.size))
    B = random_dcsc(rng, k, n, 0.12)
    got, _ = gpu_spgemm(A, B)
    want, _ = oracle.spgemm(A, B)
    bit_exact(got, want)


def test_unsorted_mode_same_entries():
    A = cb.gen_rmat(13, 8)
    got, _ = gpu_spgemm(A, A, mode="unsorted")
    want, _ = oracle.spgemm(A, A)
    bit_exact(sort_columns(got), want)


def test_empty_and_ragged():
    rng = np.random.default_rng(5)
    # A with empty columns; B touching only some of them -> some empty output columns
    A = cb.Dcsc.from_coo(50, 40, [1, 3, 7], [2, 2, 9], [1, 2, 3])
    B = random_dcsc(rng, 40, 30, 0.2)
    got, _ = gpu_spgemm(A, B)
    want, _ = oracle.spgemm(A, B)
    bit_exact(got, want)
    Z = cb.Dcsc.from_coo(50, 40, [], [], [])
    got, stats = gpu_spgemm(Z, B)
    assert got.nnz == 0 and got.nzc == 0 and stats["nnz"] == 0


def test_explicit_zeros_kept():
    A = cb.Dcsc.from_coo(2, 2, [0, 0], [0, 1], [1, -1])
    B = cb.Dcsc.from_coo(2, 1, [0, 1], [0, 0], [1, 1])
    got, _ = gpu_spgemm(A, B)
    assert got.nnz == 1 and got.ir[0] == 0 and got.val[0] == 0


def test_invalid_input_rejected():
    A = cb.gen_rmat(8, 4)
    bad = cb.Dcsc(A.m, A.n, A.cp.copy(), A.jc.copy(), A.ir.copy(), A.val.copy())
    c = int(np.argmax(np.diff(bad.cp) >= 2))
    b0 = bad.cp[c]
    bad.ir[b0], bad.ir[b0 + 1] = bad.ir[b0 + 1], bad.ir[b0]  # rows out of order
    with pytest.raises(cb.CbxError) as ei:
        gpu_spgemm(bad, A)
    assert ei.value.status == cb.ERR_INVALID
    with pytest.raises(cb.CbxError):
        gpu_spgemm(A, cb.gen_rmat(9, 4))  # inner dimensions differ


@pytest.mark.parametrize("split,share", [(4096, 0.25), (4096, 0.0), (131072, 0.25)])
def test_large_units_with_and_without_row_maps(split, share):
    """Columns above unit_split_nnz outputs take the 16384-output units: with
    the symbolic pass's row maps (share > 0) they are computed in one sweep,
    without them (share 0) in the counting + accumulating sweeps.  Both are
    bit-exact against the oracle; the low split makes large units at s16."""
    A = cb.gen_rmat(16, 16)
    base = cb.get_config()
    try:
        cb.set_config(unit_split_nnz=split, row_map_share=share)
        got, st = gpu_spgemm(A, A)
    finally:
        cb.set_config(**base)
    want, _ = oracle.spgemm(A, A)
    if split == 4096:
        assert st["units_heavy"] > 0
    assert np.array_equal(got.cp, want.cp) and np.array_equal(got.jc, want.jc)
    assert np.array_equal(got.ir, want.ir) and np.array_equal(got.val, want.val)


@pytest.mark.parametrize("share", [0.25, 0.0])
@pytest.mark.parametrize("dtype", ["float32", "bool"])
def test_large_units_with_4_byte_accumulators(share, dtype):
    """Large units of 4-byte accumulators (FP32, bool) have a 12288-output
    value window (16384 for 8-byte ones): split lowered to 4096 so most
    columns take them, with and without row maps.  FP32 values k/16, k in
    {1, 2}: products are multiples of 2^-8 and a column's at most 660,596 of
    them sum below 2^14, exact in 22 bits, so the product equals the
    oracle's bit for bit."""
    A = cb.gen_rmat(16, 16)
    rng = np.random.default_rng(5)
    A.val = (rng.integers(1, 3, A.nnz) / 16.0).astype(np.float32) if dtype == "float32" else np.ones(A.nnz, np.uint8)
    A = cb.Dcsc(A.m, A.n, A.cp, A.jc, A.ir, A.val, dtype)
    base = cb.get_config()
    try:
        cb.set_config(unit_split_nnz=4096, row_map_share=share)
        got, st = gpu_spgemm(A, A)
    finally:
        cb.set_config(**base)
    want, _ = oracle.spgemm(A, A)
    assert st["units_heavy"] > 0
    bit_exact(got, want)


@pytest.mark.parametrize("share", [0.25, 0.0])
def test_large_units_on_tall_matrices(share):
    """m above 2^20 rows: the symbolic pass runs several bitmap passes and
    writes the row map pass by pass; large units (split lowered to 4096) read
    slices of it.  Bit-exact against the oracle with and without row maps."""
    rng = np.random.default_rng(31)
    m, n, k = (1 << 21) + 5000, 2000, 300
    key = np.unique(rng.integers(0, n, 400_000) * m + rng.integers(0, m, 400_000))
    A = cb.Dcsc.from_coo(m, n, key % m, key // m, rng.integers(-4, 5, key.size).astype(np.int64), "int64")
    keyb = np.unique(rng.integers(0, k, 15_000) * n + rng.integers(0, n, 15_000))
    B = cb.Dcsc.from_coo(n, k, keyb % n, keyb // n, rng.integers(-4, 5, keyb.size).astype(np.int64), "int64")
    base = cb.get_config()
    try:
        cb.set_config(unit_split_nnz=4096, row_map_share=share)
        got, st = gpu_spgemm(A, B)
    finally:
        cb.set_config(**base)
    want, _ = oracle.spgemm(A, B)
    assert st["units_heavy"] > 0
    assert np.array_equal(got.cp, want.cp) and np.array_equal(got.jc, want.jc)
    assert np.array_equal(got.ir, want.ir) and np.array_equal(got.val, want.val)


@pytest.mark.parametrize("m", [300_007, (1 << 18) + 1])
def test_symbolic_bitmap_sub_passes(m):
    """One 256-block cutting pass whose bitmap is built in 64-block sub-passes,
    the last one partial (74 blocks: 64 + 10; 65 blocks: 64 + 1): unit counts,
    row maps and the product are bit-exact against the oracle."""
    rng = np.random.default_rng(37)
    n, k = 1500, 200
    key = np.unique(rng.integers(0, n, 300_000) * m + rng.integers(0, m, 300_000))
    A = cb.Dcsc.from_coo(m, n, key % m, key // m, rng.integers(-4, 5, key.size).astype(np.int64), "int64")
    keyb = np.unique(rng.integers(0, k, 12_000) * n + rng.integers(0, n, 12_000))
    B = cb.Dcsc.from_coo(n, k, keyb % n, keyb // n, rng.integers(-4, 5, keyb.size).astype(np.int64), "int64")
    base = cb.get_config()
    try:
        cb.set_config(unit_split_nnz=4096)
        got, st = gpu_spgemm(A, B)
    finally:
        cb.set_config(**base)
    want, _ = oracle.spgemm(A, B)
    assert st["units_light"] + st["units_heavy"] > 0
    bit_exact(got, want)


@pytest.mark.parametrize("band", [16, 64])
@pytest.mark.parametrize("m,share", [((1 << 21) + 5000, 0.25), (300_007, 0.0), (300_007, 0.25)])
def test_banded_symbolic_pass(band, m, share):
    """sym_band: the unit tier's symbolic pass in bands of 16 / 64 row blocks
    (one workgroup per column and band, then the per-column cut), on tall
    operands with a partial last band (300,007 rows = 74 blocks) and several
    256-block cutting passes (2^21 + 5000 rows), with and without row maps:
    bit-exact against the oracle, as many units as the per-column pass."""
    rng = np.random.default_rng(41)
    n, k = 1500, 200
    key = np.unique(rng.integers(0, n, 300_000) * m + rng.integers(0, m, 300_000))
    A = cb.Dcsc.from_coo(m, n, key % m, key // m, rng.integers(-4, 5, key.size).astype(np.int64), "int64")
    keyb = np.unique(rng.integers(0, k, 12_000) * n + rng.integers(0, n, 12_000))
    B = cb.Dcsc.from_coo(n, k, keyb % n, keyb // n, rng.integers(-4, 5, keyb.size).astype(np.int64), "int64")
    base = cb.get_config()
    try:
        # the per-column pass with the 16-block cut (no row lists), which the banded pass shares
        cb.set_config(unit_split_nnz=4096, row_map_share=share, sym_band=0, unit_lists=0)
        ref, st0 = gpu_spgemm(A, B)
        cb.set_config(unit_split_nnz=4096, row_map_share=share, sym_band=band)
        got, st = gpu_spgemm(A, B)
        P = np.zeros(cb.MLP_PARAMS, np.float32)
        P[-6:] = 1.0  # feature scale
        P[-13] = 20.0  # output bias: every column predicted far above heavy_nnz, so all take the exact path
        cb.set_predictor(P)
        cb.set_config(unit_split_nnz=4096, row_map_share=share, sym_band=band, heavy_nnz=4096)
        got_l, st_l = gpu_spgemm(A, B, "plus_times", "learned")  # its exact columns run the banded pass too
    finally:
        cb.set_predictor(None)
        cb.set_config(**base)
    want, _ = oracle.spgemm(A, B)
    assert st["units_light"] + st["units_heavy"] > 0
    assert (st["units_light"], st["units_heavy"]) == (st0["units_light"], st0["units_heavy"])
    for C in (ref, got, got_l):
        bit_exact(C, want)


def _spread(m, n, k, seed, dtype="int64"):
    """A (m x n) with rows spread uniformly over all of m, B (n x k): the
    product's columns hold ~10^4 outputs thinly over every 4096-row block.
    Values are nonzero integers in [-4, 4]: no product is a signed zero, whose
    sign under max would depend on the order of equal products."""
    rng = np.random.default_rng(seed)
    nz = np.array([-4, -3, -2, -1, 1, 2, 3, 4], np.int64)
    key = np.unique(rng.integers(0, n, 400_000) * m + rng.integers(0, m, 400_000))
    A = cb.Dcsc.from_coo(m, n, key % m, key // m, rng.choice(nz, key.size), "int64")
    keyb = np.unique(rng.integers(0, k, 15_000) * n + rng.integers(0, n, 15_000))
    B = cb.Dcsc.from_coo(n, k, keyb % n, keyb // n, rng.choice(nz, keyb.size), "int64")
    return A.astype(dtype), B.astype(dtype)


@pytest.mark.parametrize("m", [200_003, 300_007, (1 << 21) + 5000])
@pytest.mark.parametrize("semiring,dtype", [("plus_times", "int64"), ("min_plus", "float64"),
                                            ("select_max", "float32"), ("plus_times", "bool")])
def test_small_units_by_row_list(m, semiring, dtype):
    """unit_lists: the symbolic pass writes the sorted rows of the small-unit
    columns, cuts their units by output count over any number of row blocks,
    and those wider than 16 blocks take their ranks from the list
    (k_num_list).  Thin columns over every block of A's rows: 49 blocks (the
    symbolic pass's single-sub-pass form), one 74-block pass in two
    sub-passes, and three 256-block passes: fewer units than the 16-block
    cut, and the same product as without lists and as the oracle, bit for
    bit, in both modes."""
    A, B = _spread(m, 2000, 300, 43, dtype)
    base = cb.get_config()
    P = np.zeros(cb.MLP_PARAMS, np.float32)
    P[-6:] = 1.0  # feature scale
    P[-13] = 20.0  # output bias: predictions clamp to min(flop_j, m), above heavy_nnz: every column exact
    try:
        cb.set_config(unit_lists=0)
        plain, st0 = gpu_spgemm(A, B, semiring)
        cb.set_config(unit_lists=1)
        got, st = gpu_spgemm(A, B, semiring)
        cb.set_predictor(P)  # learned mode: its exact columns cut list units too
        cb.set_config(unit_lists=1, heavy_nnz=4096)
        got_l, st_l = gpu_spgemm(A, B, semiring, "learned")
    finally:
        cb.set_predictor(None)
        cb.set_config(**base)
    want, _ = oracle.spgemm(A, B, semiring)
    assert 0 < st["units_light"] < st0["units_light"]
    assert st["nnz_main"] == st0["nnz_main"] > 0
    assert st_l["units_heavy"] == st["units_light"] + st["units_heavy"]  # the exact columns' units
    bit_exact(plain, want)
    bit_exact(got, want)
    bit_exact(got_l, want)


@pytest.mark.parametrize("group", [8, 16, 32, 64, 128])
def test_xcd_grouped_unit_order(group):
    """unit_xcd_group: the unit launches hand runs of `group` consecutive
    units to each XCD (a permutation of workgroups to units, identity on the
    tail after the last full super-group of 8 * group): the same product as
    the round-robin order and as the oracle, in both modes (large units split
    at 4096 outputs so all three unit kernels run)."""
    A = cb.gen_rmat(16, 16)
    dA = A.to_device()
    base = cb.get_config()
    P = np.zeros(cb.MLP_PARAMS, np.float32)
    P[-6:] = 1.0
    P[-13] = 20.0  # learned mode: every unit-tier column exact, through the same unit kernels
    try:
        cb.set_config(unit_split_nnz=4096, unit_order=1)
        ref, st0 = cb.spgemm(dA, dA)
        cb.set_config(unit_split_nnz=4096, unit_order=1, unit_xcd_group=group)
        got, st = cb.spgemm(dA, dA)
        cb.set_predictor(P)
        cb.set_config(unit_split_nnz=4096, unit_order=1, unit_xcd_group=group, heavy_nnz=4096)
        got_l, _ = cb.spgemm(dA, dA, "plus_times", "learned")
    finally:
        cb.set_predictor(None)
        cb.set_config(**base)
    want, _ = oracle.spgemm(A, A)
    assert st["units_light"] + st["units_heavy"] > 8 * group
    for C in (ref, got, got_l):
        bit_exact(C.to_host(), want)


def test_row_list_units_with_large_units_and_row_maps():
    """Row-list units beside large units (split lowered to 4096) that read
    row maps, on several 256-block passes: per pass, columns over the split
    take large units, the others list or bitmap units."""
    A, B = _spread((1 << 21) + 5000, 2000, 300, 47)
    base = cb.get_config()
    try:
        cb.set_config(unit_split_nnz=4096, row_map_share=0.25)
        got, st = gpu_spgemm(A, B)
    finally:
        cb.set_config(**base)
    want, _ = oracle.spgemm(A, B)
    assert st["units_heavy"] > 0
    bit_exact(got, want)


def test_operand_validation_with_empty_columns():
    """The product's operand validation accepts explicitly empty columns in cp
    (a descent of rows where a column starts after an empty one is legal) and
    still rejects a descent inside a column."""
    m = 64
    # columns 3, 7 (empty), 9 of a 64 x 16 matrix; column 9 starts below column 3's last row
    cp = np.array([0, 3, 3, 6], np.int64)
    jc = np.array([3, 7, 9], np.int64)
    ir = np.array([10, 20, 30, 1, 2, 40], np.int64)
    val = np.array([1, 2, 3, 4, 5, 6], np.int64)
    A = cb.Dcsc(m, 16, cp, jc, ir, val)
    B = cb.gen_rmat(4, 4)  # 16 x 16
    got, _ = gpu_spgemm(A, B)
    keep = cb.Dcsc(m, 16, np.array([0, 3, 6], np.int64), np.array([3, 9], np.int64), ir, val)
    want, _ = oracle.spgemm(keep, B)
    assert np.array_equal(got.ir, want.ir) and np.array_equal(got.val, want.val)
    assert np.array_equal(got.cp, want.cp) and np.array_equal(got.jc, want.jc)
    bad_ir = ir.copy()
    bad_ir[4] = 0  # 1 then 0 inside column 9
    with pytest.raises(cb.CbxError) as ei:
        gpu_spgemm(cb.Dcsc(m, 16, cp, jc, bad_ir, val), B)
    assert ei.value.status == cb.ERR_INVALID


def test_device_structure_check_cases():
    """cbx_dcsc_check accepts a valid DCSC (rows descend only where a column
    starts) and rejects each malformation: rows out of order or repeated
    inside a column, a row out of range, jc not ascending, cp not starting at
    0, and an empty column in cp."""
    A = cb.gen_rmat(10, 8)
    A.to_device().check()
    c = int(np.argmax(np.diff(A.cp) >= 2))
    b0 = int(A.cp[c])

    def variant(edit):
        X = cb.Dcsc(A.m, A.n, A.cp.copy(), A.jc.copy(), A.ir.copy(), A.val.copy())
        edit(X)
        return X

    def swap_rows(X):
        X.ir[b0], X.ir[b0 + 1] = X.ir[b0 + 1], X.ir[b0]

    def repeat_row(X):
        X.ir[b0 + 1] = X.ir[b0]

    def row_out_of_range(X):
        X.ir[-1] = X.m

    def jc_unsorted(X):
        X.jc[0], X.jc[1] = X.jc[1], X.jc[0]

    def cp_offset(X):
        X.cp[0] = 1

    def empty_column(X):
        X.cp[1] = X.cp[0]

    for edit in (swap_rows, repeat_row, row_out_of_range, jc_unsorted, cp_offset, empty_column):
        dX = variant(edit).to_device()
        with pytest.raises(cb.CbxError) as ei:
            dX.check()
        assert ei.value.status == cb.ERR_INVALID, edit.__name__
        dX.release()


@pytest.mark.parametrize("scale,ef", [(18, 16)])
def test_full_size_linear_checksums(scale, ef):
    """Large R-MAT product checked without materialising it on the host:
    device structure check (rows sorted and in range) plus every column's
    weighted sum against its closed form sum_k (u^T A)_k A(k,j)."""
    A = cb.gen_rmat(scale, ef)
    dA = A.to_device()
    dC, stats = cb.spgemm(dA, dA)
    dC.check()
    got = dC.column_checksums(salt=12345)
    _, jcC = dC.structure()
    jcB, want = oracle.product_checksums(A, A, salt=12345)
    pos = np.searchsorted(jcB, jcC)
    assert np.array_equal(jcB[pos], jcC)
    assert np.array_equal(got, want[pos])
    dropped = np.ones(jcB.size, bool)
    dropped[pos] = False
    assert not np.any(want[dropped])
    assert stats["units_light"] + stats["units_heavy"] > 0


def heavy_column_sample(A: cb.Dcsc, nheavy: int, nrand: int, seed: int) -> cb.Dcsc:
    """B = the nheavy columns of A with the most products in A*A plus nrand
    random ones, at A's shape: full-length columns of A*A at a fraction of
    its cost."""
    lens = np.zeros(A.n, np.int64)
    lens[A.jc] = np.diff(A.cp)
    flop = np.add.reduceat(lens[A.ir], A.cp[:-1])
    rng = np.random.default_rng(seed)
    qs = np.unique(np.concatenate([np.argsort(flop)[-nheavy:], rng.choice(A.nzc, nrand, replace=False)]))
    idx = np.concatenate([np.arange(A.cp[q], A.cp[q + 1]) for q in qs])
    return cb.Dcsc.from_coo(A.m, A.n, A.ir[idx], np.repeat(A.jc[qs], np.diff(A.cp)[qs]), A.val[idx], A.dtype)


@pytest.mark.parametrize("scale", [16, 18])
def test_fp64_plus_times_long_columns_rtol_1e_12(scale):
    """FP64 plus-times with positive values on R-MAT columns of 10^5 and more
    products (the heaviest columns of A*A plus random ones): every entry
    within 1e-12 relative of the FP64 oracle (sums of positive terms in
    another order differ by about sqrt(k) * 2^-53 relatively)."""
    rng = np.random.default_rng(scale)
    A = cb.gen_rmat(scale, 16).astype("float64")
    A.val = rng.uniform(0.5, 2.0, A.nnz)
    B = heavy_column_sample(A, 16, 48, seed=scale)
    got, st = gpu_spgemm(A, B)
    want, flops = oracle.spgemm(A, B)
    same_pattern(got, want)
    colflop, _ = oracle.symbolic(A, B)
    assert colflop.max() >= 100_000
    rel = np.abs(got.val - want.val) / want.val
    assert rel.max() <= 1e-12, f"max relative error {rel.max():.3g}"
