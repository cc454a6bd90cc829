"""Autotuner for the SpGEMM tunables (cbx_spgemm_set_config).

The reference hard-codes its bin switch and hash-table sizing
(mtSpGEMM.h:311, :366-367).  Here the flop bound of the LDS hash tier, the
flop bounds of its table sizes (the LDS table sizes), the tiny tier's bound
and workgroup shape, the output size that sends a column to the large units
(the other unit workgroup shape) and the learned pass's unit target and
exact-column bound are measured per workload: every candidate of a small
grid runs the workload on the device (after a warm-up), the fastest
configuration with an unchanged output size wins, and the result is stored
in a versioned table keyed by workload name and device.  Workloads: the
R-MAT products, the 256 batched Galerkin products and the MCL expansion.

    python -m combblas_amd.autotune rmat_s20_ef16 galerkin_256   # on a GPU machine
    combblas_amd.autotune.apply("galerkin_256")                  # install the stored choice
"""
from __future__ import annotations

import itertools
import json
import sys
import time
from pathlib import Path

import combblas_amd as cb

TABLE_PATH = Path(__file__).resolve().parent / "tuning" / "table.json"
TABLE_VERSION = 1

# Galerkin and MCL workloads of the benchmark (the same inputs)
GALERKIN_GRIDS = dict(count=256, lo=16, hi=96, seed=0xC5)
MCL_PARAMS = (1e-4, 1000, 0, 0.9)  # hard threshold, select, recover number, recover fraction


def _spgemm_workload(scale: int, ef: int, mode: str):
    def setup():
        dA = cb.gen_rmat(scale, ef).to_device()

        def run():
            C, st = cb.spgemm(dA, dA, "plus_times", mode)
            C.release()
            return st["nnz"]
        return run, dA.release
    return setup


def _galerkin_workload():
    import numpy as np
    rng = np.random.default_rng(GALERKIN_GRIDS["seed"])
    grids = [int(g) for g in rng.integers(GALERKIN_GRIDS["lo"], GALERKIN_GRIDS["hi"] + 1, GALERKIN_GRIDS["count"])]

    def setup():
        As = [cb.gen_stencil27(g).to_device() for g in grids]
        Ts = [cb.gen_aggregation(g, True).to_device() for g in grids]

        def run():
            S, st = cb.galerkin_batched(As, Ts)
            for X in S:
                X.release()
            return st["nnz"]

        def done():
            for X in As + Ts:
                X.release()
        return run, done
    return setup


def _merge_workload(scale: int, ef: int, nstages: int):
    """SUMMA's stage merge: the local products of nstages stages of R-MAT
    A*A (computed once, on the configuration in force at setup), merged."""
    def setup():
        import numpy as np
        A = cb.gen_rmat(scale, ef)
        cuts = np.linspace(0, A.n, nstages + 1).astype(np.int64)
        parts = []
        for s in range(nstages):
            As = A.column_range(int(cuts[s]), int(cuts[s + 1])).to_device()
            Bs = A.row_range(int(cuts[s]), int(cuts[s + 1])).to_device()
            C, _ = cb.spgemm(As, Bs)
            parts.append(C)
            As.release()
            Bs.release()

        def run():
            M = cb.merge(parts)
            nnz = M.nnz
            M.release()
            return nnz

        def done():
            for p in parts:
                p.release()
        return run, done
    return setup


def _mcl_workload(n: int, mode: str):
    def setup():
        dA = cb.gen_powerlaw_stochastic(n).to_device()

        def run():
            P, st = cb.mcl_expand(dA, dA, *MCL_PARAMS, mode=mode)
            P.release()
            return st["nnz_product"]
        return run, dA.release
    return setup


GRID = {
    "hash_max_flop": [1024, 2048],
    "hash_mid_flop": [1024],
    "tiny_flop": [256],
    "unit_split_nnz": [131072, 262144],
    "unit_order": [1],
    "tiny_algo": [0, 1],
    "sym_band": [0, 16],
    "unit_lists": [0, 1],
    "unit_xcd_group": [0, 32],
}
# the learned single pass adds its own knobs; the hash tiers are swept more
# coarsely there to keep the product of the grids small
LEARNED_GRID = {
    "hash_max_flop": [1024, 2048],
    "hash_mid_flop": [1024],
    "tiny_flop": [256],
    "unit_split_nnz": [131072],
    "learned_target": [0.75, 0.9],
    "heavy_nnz": [4096, 8192, 16384],
    "sym_band": [0, 16],
    "unit_lists": [0, 1],
}
# Galerkin columns hold 216 (A*T) and 64 (S*(A*T)) products: the tiny tier's
# bound, form and workgroup shape decide where they run
GALERKIN_GRID = {
    "tiny_flop": [128, 256],
    "hash_mid_flop": [256, 1024],
    "tiny_wg_waves": [1, 2, 4],
    "tiny_algo": [0, 1],
}
MCL_GRID = {
    "hash_max_flop": [1024, 2048],
    "unit_split_nnz": [262144],
    "sym_band": [0, 16],
    "mcl_fuse": [0, 1],
    "unit_lists": [0, 1],
}
# the merge's product: each output column combines the stages' columns, so
# its flops are the stages' column lengths summed
MERGE_GRID = {
    "hash_max_flop": [1024, 2048, 4096],
    "hash_mid_flop": [1024],
    "tiny_algo": [0, 1],
    "unit_lists": [0, 1],
}
MCL_LEARNED_GRID = {
    "hash_max_flop": [2048, 4096],
    "learned_target": [0.75, 1.0],
    "heavy_nnz": [8192, 65536],
}

# name -> (setup, description, mode, grid); setup() returns (run, release):
# run() executes the workload once and returns its output size
WORKLOADS = {
    "rmat_s20_ef16": (_spgemm_workload(20, 16, "two_phase"), "R-MAT s20 ef16 A*A", "two_phase", GRID),
    "rmat_s20_ef16_learned": (_spgemm_workload(20, 16, "learned"), "R-MAT s20 ef16 A*A", "learned", LEARNED_GRID),
    "rmat_s16_ef16": (_spgemm_workload(16, 16, "two_phase"), "R-MAT s16 ef16 A*A", "two_phase", GRID),
    "galerkin_256": (_galerkin_workload(), "256 batched Galerkin products S*(A*T), grid edges 16..96, FP64",
                     "two_phase", GALERKIN_GRID),
    "merge_s17": (_merge_workload(17, 16, 4), "stage merge of 4 SUMMA stage products of R-MAT s17 ef16 A*A",
                  "two_phase", MERGE_GRID),
    "mcl_n1m": (_mcl_workload(1_000_000, "two_phase"), "MCL expansion A*A + prune, power-law n=10^6, FP32",
                "two_phase", MCL_GRID),
    "mcl_n1m_learned": (_mcl_workload(1_000_000, "learned"), "MCL expansion A*A + prune, power-law n=10^6, FP32",
                        "learned", MCL_LEARNED_GRID),
}


def device_name() -> str:
    """Label of the device a sweep ran on (through libcbx.so's own HIP runtime)."""
    try:
        return cb.device_name()
    except cb.CbxError:
        return "unknown"


def _time_config(run, reps: int) -> tuple[float, int]:
    size = run()  # warm-up (also fills the allocation cache)
    cb.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        run()
    cb.synchronize()
    return (time.perf_counter() - t) * 1e3 / reps, size


def tune(name: str, reps: int = 2, log=print) -> dict:
    setup, desc, mode, grid = WORKLOADS[name]
    if mode == "learned":  # the MCL expansion's input family has its own model
        from combblas_amd import predictor
        cb.set_predictor(predictor.load("powerlaw" if name.startswith("mcl") else "rmat"))
    run, release = setup()
    base = cb.get_config()
    keys = list(grid)
    results = []
    ref_size = None
    for combo in itertools.product(*(grid[k] for k in keys)):
        cfg = dict(base, **dict(zip(keys, combo)))
        if cfg["hash_mid_flop"] > cfg["hash_max_flop"] or cfg["tiny_flop"] > cfg["hash_mid_flop"]:
            continue
        cb.set_config(**cfg)
        ms, size = _time_config(run, reps)
        ref_size = size if ref_size is None else ref_size
        point = dict(zip(keys, combo))
        results.append({"config": point, "ms": round(ms, 3), "same_nnz": size == ref_size})
        log(f"{name}: {point} -> {ms:.2f} ms")
    cb.set_config(**base)
    release()
    best = min((r for r in results if r["same_nnz"]), key=lambda r: r["ms"])
    entry = {"workload": name, "description": desc, "device": device_name(), "mode": mode,
             "best": dict(base, **best["config"]), "best_ms": best["ms"],
             "default_ms": next((r["ms"] for r in results if all(base[k] == v for k, v in r["config"].items())), None),
             "sweep": results, "date": time.strftime("%Y-%m-%d")}
    return entry


def load_table(path: Path = TABLE_PATH) -> dict:
    if path.exists():
        t = json.loads(path.read_text())
        if t.get("version") == TABLE_VERSION:
            return t
    return {"version": TABLE_VERSION, "entries": {}}


def apply(name: str, path: Path = TABLE_PATH) -> dict | None:
    """Installs the stored configuration for `name` (None if the table has none
    or it was written for other tunables)."""
    e = load_table(path)["entries"].get(name)
    if e is None or set(e["best"]) != set(cb.get_config()):
        return None
    return cb.set_config(**e["best"])


def main(argv: list[str]) -> int:
    names = argv or ["rmat_s20_ef16"]
    out = Path(argv[argv.index("--out") + 1]) if "--out" in argv else TABLE_PATH
    names = [n for n in names if n in WORKLOADS]
    table = load_table(out)
    out.parent.mkdir(parents=True, exist_ok=True)
    for n in names:  # the table is written after every workload, so a cut-off sweep keeps the finished ones
        table["entries"][n] = tune(n)
        out.write_text(json.dumps(table, indent=1) + "\n")
    print(json.dumps({n: {k: table["entries"][n][k] for k in ("best", "best_ms", "default_ms")} for n in names},
                     indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
