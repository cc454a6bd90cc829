"""Writes the golden fixtures of tests/golden/ from the genuine CombBLAS
kernels (oracle/_ref/libcbref.so, built from the reference tree by
oracle/Makefile).  Run in a checkout that has the reference tree:

    make -C oracle all ref && python tests/golden/make_golden.py

Fixtures are data only: generated inputs (or the generator's parameters) and
the reference's outputs.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import combblas_amd as cb  # noqa: E402
import oracle  # noqa: E402

OUT = Path(__file__).resolve().parent


def tiny_inputs(dtype: str, seed: int):
    rng = np.random.default_rng(seed)
    m, k, n = 48, 40, 36

    def mk(rows, cols, dens):
        mask = rng.random((rows, cols)) < dens
        r, c = np.nonzero(mask)
        if dtype.startswith("float"):
            v = rng.standard_normal(r.size)
        elif dtype == "bool":
            v = rng.random(r.size) < 0.8
        else:
            v = rng.integers(-9, 10, r.size)
        return cb.Dcsc.from_coo(rows, cols, r, c, v, dtype)

    return mk(m, k, 0.1), mk(k, n, 0.12)


def main() -> None:
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libcbref.so missing: run `make -C oracle ref` with the reference tree present")
    # C1: R-MAT s14 ef8, PlusTimes<int64>, LocalHybridSpGEMM
    A = cb.gen_rmat(14, 8)
    R = oracle.ref_spgemm(A, A, "plus_times", "hybrid")
    flop, nnz = oracle.ref_symbolic(A, A)
    np.savez_compressed(OUT / "spgemm_rmat_s14_ef8_plus_times_int64.npz",
                        nnz_a=A.nnz, flops=int(flop.sum()), nnz_c=R.nnz, jc=R.jc, col_nnz=np.diff(R.cp),
                        col_digest=oracle.column_digests(R))
    combos = [("plus_times", "int64"), ("plus_times", "float64"), ("plus_times", "float32"),
              ("plus_times", "bool"), ("min_plus", "int64"), ("min_plus", "float64"),
              ("select_max", "int64"), ("select_max", "float32")]
    for i, (sr, dt) in enumerate(combos):
        A, B = tiny_inputs(dt, 1000 + i)
        for mode in ("hybrid", "hash_unsorted"):
            C = oracle.ref_spgemm(A, B, sr, mode)
            np.savez_compressed(
                OUT / f"spgemm_tiny_{sr}_{dt}_{mode}.npz", semiring=sr, dtype=dt, mode=mode,
                a_m=A.m, a_n=A.n, a_cp=A.cp, a_jc=A.jc, a_ir=A.ir, a_val=A.val,
                b_m=B.m, b_n=B.n, b_cp=B.cp, b_jc=B.jc, b_ir=B.ir, b_val=B.val,
                c_cp=C.cp, c_jc=C.jc, c_ir=C.ir, c_val=C.val)
    # MCLPruneRecoverySelect on small MCL-shaped inputs (ParFriends.h:185-345)
    sys.path.insert(0, str(ROOT / "tests"))
    from helpers import MCL_PARAMS, mcl_columns
    for i, params in enumerate(MCL_PARAMS[:3]):
        dt = "float64" if i == 2 else "float32"
        A = mcl_columns(100 + i, ncols=40, m=2500, dtype=dt)
        R = oracle.ref_mcl_prune(A, *params)
        thr, select, recover, pct = params
        np.savez_compressed(OUT / f"mcl_prune_{i}_{dt}.npz", dtype=dt, m=A.m, n=A.n, thr=thr, select=select,
                            recover=recover, pct=pct, a_cp=A.cp, a_jc=A.jc, a_ir=A.ir.astype(np.int32),
                            a_val=A.val, c_cp=R.cp, c_jc=R.jc, c_ir=R.ir.astype(np.int32), c_val=R.val)
    print("fixtures written to", OUT)


if __name__ == "__main__":
    main()
