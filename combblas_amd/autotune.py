"""Autotuner for the SpGEMM tunables (cbx_spgemm_set_config).

The reference hard-codes its bin switch and hash-table sizing
(mtSpGEMM.h:311, :366-367).  Here the flop bound of the LDS hash tier, the
flop bounds of its three table sizes, the output size that sends a column to
the large units and the learned pass's unit target are measured per
workload: every candidate of a small grid runs the workload's product on the
device (after a warm-up), the fastest exact configuration wins, and the
result is stored in a versioned table keyed by workload name and device.

    python -m combblas_amd.autotune rmat_s20_ef16      # on a GPU machine
    combblas_amd.autotune.apply("rmat_s20_ef16")       # install the stored choice
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

# name -> (generator, semiring, mode)
WORKLOADS = {
    "rmat_s20_ef16": (lambda: cb.gen_rmat(20, 16), "plus_times", "two_phase"),
    "rmat_s20_ef16_learned": (lambda: cb.gen_rmat(20, 16), "plus_times", "learned"),
    "rmat_s16_ef16": (lambda: cb.gen_rmat(16, 16), "plus_times", "two_phase"),
}

GRID = {
    "hash_max_flop": [1024, 2048, 4096],
    "hash_mid_flop": [512, 1024],
    "tiny_flop": [128, 256],
    "unit_split_nnz": [65536, 131072, 262144, 524288],
}
# the learned single pass adds its own knobs; the hash tiers are swept more
# coarsely there to keep the product of the grids small
LEARNED_GRID = {
    "hash_max_flop": [1024, 2048],
    "hash_mid_flop": [1024],
    "tiny_flop": [256],
    "unit_split_nnz": [131072, 262144],
    "learned_target": [0.75, 0.9],
    "heavy_nnz": [8192, 16384, 32768, 65536],
}


def device_name() -> str:
    """Label of the device a sweep ran on (through libcbx.so's own HIP runtime)."""
    try:
        return cb.device_name()
    except cb.CbxError:
        return "unknown"


def _time_config(dA, semiring: str, mode: str, reps: int) -> tuple[float, int]:
    C, st = cb.spgemm(dA, dA, semiring, mode)  # warm-up (also fills the allocation cache)
    nnz = st["nnz"]
    C.release()
    cb.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        C, st = cb.spgemm(dA, dA, semiring, mode)
        C.release()
    cb.synchronize()
    return (time.perf_counter() - t) * 1e3 / reps, nnz


def tune(name: str, reps: int = 2, log=print) -> dict:
    gen, semiring, mode = WORKLOADS[name]
    A = gen()
    dA = A.to_device()
    if mode == "learned":
        from combblas_amd import predictor
        cb.set_predictor(predictor.load())
    base = cb.get_config()
    grid = LEARNED_GRID if mode == "learned" else GRID
    keys = list(grid)
    results = []
    ref_nnz = None
    for combo in itertools.product(*(grid[k] for k in keys)):
        cfg = dict(zip(keys, combo))
        if cfg["hash_mid_flop"] > cfg["hash_max_flop"] or cfg["tiny_flop"] > cfg["hash_mid_flop"]:
            continue
        cb.set_config(**dict(base, **cfg))
        ms, nnz = _time_config(dA, semiring, mode, reps)
        ref_nnz = nnz if ref_nnz is None else ref_nnz
        results.append({"config": cfg, "ms": round(ms, 3), "same_nnz": nnz == ref_nnz})
        log(f"{name}: {cfg} -> {ms:.2f} ms")
    cb.set_config(**base)
    best = min((r for r in results if r["same_nnz"]), key=lambda r: r["ms"])
    entry = {"workload": name, "device": device_name(), "semiring": semiring, "mode": mode,
             "best": dict(base, **best["config"]), "best_ms": best["ms"],
             "default_ms": next((r["ms"] for r in results if all(base[k] == v for k, v in r["config"].items())), None),
             "sweep": results, "date": time.strftime("%Y-%m-%d")}
    dA.release()
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
    for n in names:
        table["entries"][n] = tune(n)
    out.parent.mkdir(parents=True, exist_ok=True)
    out.write_text(json.dumps(table, indent=1) + "\n")
    print(json.dumps({n: {k: table["entries"][n][k] for k in ("best", "best_ms", "default_ms")} for n in names},
                     indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
