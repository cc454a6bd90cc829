#!/usr/bin/env python3
"""Benchmark of the node-local sparse kernels on one MI355X.

Headline workload: C = A*A on PlusTimes<int64>, A = R-MAT scale 20, edge
factor 16 (a=.57, b=c=.19, d=.05, duplicates summed, values uniform in
[1,9]), int64 DCSC in and out; A is resident in HBM before the timed region.
A "step" is one complete local product, output arrays allocated from the
library's cache and returned after the step.  It is timed for the plain
two-phase path (symbolic + numeric) and for the learned single pass (the
trained nnz predictor replaces the symbolic pass of the large columns);
`value` is the GFLOPS (2 x multiply-adds / s) of the faster of the two, both
outputs are checked bit-exact at full size (device structure check +
per-column linear checksums against their closed form), and the ratio of
their times is `learned_vs_two_phase_speedup`.

Beside it, in the same line: the roofline of the dominant kernel (HIP
events on the library's stream), the CPU baseline (the reference's
LocalHybridSpGEMM on this host's cores over a bounded column sample), and
the other workloads of the same kernels: top-down BFS through SpMSpV on the
symmetrised s20 graph (64 roots), one MCL expansion step in FP32 (n = 10^6,
power-law degrees, prune 1e-4, top-1000), and 256 batched Galerkin products
R*A*R^T in FP64.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "SpGEMM GFLOPS + HBM GB/s (R-MAT s20 ef16 A·A); learned-vs-2-phase speedup"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def dcsc_bytes(nnz: int, nzc: int, vbytes: int = 8) -> int:
    """Algorithmic bytes of a DCSC with int64 indices: cp+jc (16 per column) + 8 + ir+val per entry."""
    return 16 * nzc + 8 + (8 + vbytes) * nnz


def host_cpu() -> dict:
    """CPU model, logical CPUs this process may run on, and the physical cores among them."""
    model, cores = "unknown", set()
    avail = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(range(os.cpu_count() or 1))
    try:
        cur = {}
        for line in open("/proc/cpuinfo").read().splitlines() + [""]:
            if not line.strip():
                if cur.get("processor") is not None and int(cur["processor"]) in avail:
                    cores.add((cur.get("physical id", "0"), cur.get("core id", cur["processor"])))
                cur = {}
                continue
            k, _, v = line.partition(":")
            cur[k.strip()] = v.strip()
            if k.strip() == "model name":
                model = v.strip()
    except OSError:
        pass
    return {"model": model, "logical_cpus": len(avail), "physical_cores": len(cores) or None}


def cpu_baseline(scale: int, ef: int, budget_s: float, bfs_roots=None, bfs_edges: int | None = None) -> dict | None:
    """The reference's kernels on this host's cores, in a child process (it
    never touches the GPU): LocalHybridSpGEMM<PlusTimes<int64>> alone and with
    the SpDCCols(SpTuples) conversion on B-column slices of the same product
    (one warm-up, then the median of three runs per slice), and the
    multithreaded BFS level loop (RowSplit + generic_gespmv_threaded) over the
    benchmark's roots."""
    roots = [int(r) for r in bfs_roots] if bfs_roots is not None else []
    code = f"""
import json, sys
sys.path.insert(0, {ROOT!r})
import numpy as np
import combblas_amd as cb
import oracle
A = cb.gen_rmat({scale}, {ef})
lens = np.zeros(A.n, np.int64); lens[A.jc] = np.diff(A.cp)
fl = np.add.reduceat(lens[A.ir], A.cp[:-1]); cum = np.cumsum(fl); total = int(cum[-1])
kind = "reference"
try:
    oracle._load_ref()
except Exception:
    kind = "port"
threads = int(oracle.max_threads())
spent = 0.0; flops = 0; t_prod = 0.0; t_dcsc = 0.0; slices = []
target = max(total // 2000, 1)          # first slice ~0.05% of the work
def timed(c0, c1):
    B = oracle.column_slice(A, c0, c1)
    if kind == "reference":
        return oracle.ref_time_hybrid_dcsc_i64(A, B)[:2]
    t = oracle.time_spgemm(A, A, cols=(c0, c1))[0]
    return t, None
for s in range(16):
    if spent >= {budget_s}:
        break
    mid = int(np.searchsorted(cum, (s + 0.5) / 16 * total))  # 16 disjoint positions across the work
    c0 = mid; c1 = int(np.searchsorted(cum, cum[c0] + min(target, total // 32))) + 1
    c1 = min(max(c1, c0 + 1), A.nzc)
    runs = [timed(c0, c1) for _ in range(4)]          # one warm-up, then three timed
    spent += sum(r[1] if r[1] is not None else r[0] for r in runs)
    t_prod += float(np.median([r[0] for r in runs[1:]]))
    if runs[0][1] is not None:
        t_dcsc += float(np.median([r[1] for r in runs[1:]]))
    flops += int(fl[c0:c1].sum()); slices.append([c0, c1])
    if runs[1][0] < {budget_s} / 64:
        target *= 2
out = dict(kind=kind, threads=threads, t_prod=t_prod, t_dcsc=t_dcsc, flops=flops, slices=slices, total=total)
roots = {roots!r}
if roots:
    G = cb.gen_rmat({scale}, {ef}, symmetrize=True, no_self_loops=True, pattern=True).astype("bool")
    if kind == "reference":
        oracle.ref_bfs_threaded(G, roots[:1], threads)     # warm-up
        _, levels, secs = oracle.ref_bfs_threaded(G, roots, threads)
    else:
        import time
        t0 = time.perf_counter(); levels = sum(oracle.bfs(G, r)[1] for r in roots); secs = time.perf_counter() - t0
    out["bfs"] = dict(secs=secs, levels=levels, roots=len(roots))
print(json.dumps(out))
"""
    try:
        r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=budget_s * 8 + 300)
        line = [l for l in r.stdout.splitlines() if l.startswith("{")][-1]
        d = json.loads(line)
    except Exception as e:  # noqa: BLE001 - a failed baseline must not hide the GPU number
        log(f"cpu baseline failed: {e}")
        return None
    cpu = host_cpu()
    frac = d["flops"] / d["total"]
    out = {
        "value": 2.0 * d["flops"] / d["t_prod"] / 1e9 if d["t_prod"] > 0 else None, "unit": "GFLOPS",
        "cores": d["threads"], "kind": d["kind"], "cpu": cpu,
        "gflops_with_dcsc_conversion": (2.0 * d["flops"] / d["t_dcsc"] / 1e9 if d["t_dcsc"] > 0 else None),
        "sample": (f"LocalHybridSpGEMM PlusTimes<int64> on {len(d['slices'])} B-column slices of the same "
                   f"R-MAT s{scale} ef{ef} A*A ({frac:.2%} of its {d['total']} multiply-adds), one warm-up then "
                   f"the median of 3 runs per slice (summed: {d['t_prod']:.2f} s alone, {d['t_dcsc']:.2f} s with "
                   f"SpDCCols(SpTuples)); OpenMP threads={d['threads']} on {cpu['model']}"),
    }
    if "bfs" in d:
        b = d["bfs"]
        ms = 1e3 * b["secs"] / b["roots"]
        out["bfs"] = {"ms_per_bfs_mean": round(ms, 3), "levels_mean": round(b["levels"] / b["roots"], 2),
                      "gteps": (round(bfs_edges / b["secs"] / 1e9, 4) if bfs_edges and b["secs"] > 0 else None),
                      "cores": d["threads"], "kind": d["kind"],
                      "sample": (f"TopDownBFS level loop, RowSplit({d['threads']}) + generic_gespmv_threaded"
                                 f"<SelectMaxSRing<bool,int64>>, {b['roots']} roots of the same graph")}
    return out


def run_spgemm(cb, A, dA, mode: str, steps: int, warmup: int, col_nnz=None) -> dict:
    from combblas_amd.checks import product_matches
    for _ in range(warmup):
        C, st = cb.spgemm(dA, dA, "plus_times", mode)
        C.release()
    cb.synchronize()
    stats = []
    t0 = time.perf_counter()
    for k in range(steps):
        C, st = cb.spgemm(dA, dA, "plus_times", mode)
        stats.append(st)
        if k + 1 < steps:
            C.release()
    cb.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    exact = product_matches(C, A, A, col_nnz=col_nnz)  # full size, outside the timed region
    C.release()
    return {"ms": ms, "stats": stats, "exact": exact}


def bench_bfs(cb, scale: int, ef: int, nroots: int = 64) -> dict:
    """TopDownBFS's parents from 64 seeded roots, each level top-down (SpMSpV)
    or bottom-up as the plan chooses; the same roots top-down only for
    comparison.  Both give the same parents (the tests check it)."""
    G = cb.gen_rmat(scale, ef, symmetrize=True, no_self_loops=True, pattern=True).astype("bool")
    deg = np.zeros(G.n, np.int64)
    deg[G.jc] = np.diff(G.cp)
    plan = cb.BfsPlan(G.to_device())
    roots = cb.bfs_roots(G, nroots)
    out = {"workload": f"BFS (TopDownBFS parents), symmetrised R-MAT s{scale} ef{ef} without self-loops, "
                       f"{len(roots)} roots (seed 4242); direction-optimising levels one captured graph each, "
                       f"top-down-only levels eight to a captured graph, hubs expanded in 1024-edge pieces"}
    total_edges = 0
    for direction in ("optimizing", "top_down"):
        plan.set_direction(direction)
        plan.run(int(roots[0]), copy=False)
        ms, teps, levels, bu = [], [], [], []
        for r in roots:
            par, st = plan.run(int(r), copy=True)
            edges = int(deg[par >= 0].sum()) // 2  # undirected edges in the reached component
            if direction == "top_down":
                total_edges += edges
            ms.append(st["ms"])
            levels.append(st["levels"])
            bu.append(st["bottom_up_levels"])
            teps.append(edges / (st["ms"] * 1e-3))
        hm = len(teps) / sum(1.0 / t for t in teps)
        res = {"gteps_harmonic_mean": round(hm / 1e9, 3), "ms_per_bfs_mean": round(float(np.mean(ms)), 3),
               "levels_mean": round(float(np.mean(levels)), 2)}
        if direction == "optimizing":
            res["bottom_up_levels_mean"] = round(float(np.mean(bu)), 2)
            out.update(res)
        else:
            out["top_down_only"] = res
    plan.release()
    out["_roots"] = [int(r) for r in roots]
    out["_edges_total"] = total_edges
    return out


def bench_mcl(cb, n: int = 1_000_000, reps: int = 2) -> dict:
    """The MCL expansion with exact (two-phase) sizing and with learned nnz
    sizing, each with its stored tunables; the faster one is the headline."""
    from combblas_amd import autotune, predictor
    A = cb.gen_powerlaw_stochastic(n)
    dA = A.to_device()
    params = autotune.MCL_PARAMS
    base = cb.get_config()
    per_mode = {}
    for mode in ("two_phase", "learned"):
        if mode == "learned":  # the network trained on this input family (power-law matrices)
            if not predictor.model_path("powerlaw").exists():
                continue
            cb.set_predictor(predictor.load("powerlaw"))
        cb.set_config(**base)
        tuned = autotune.apply("mcl_n1m" + ("_learned" if mode == "learned" else ""))
        P, st = cb.mcl_expand(dA, dA, *params, mode=mode)
        P.release()
        runs = []
        for _ in range(reps):
            t = time.perf_counter()
            P, st = cb.mcl_expand(dA, dA, *params, mode=mode)
            cb.synchronize()
            runs.append(((time.perf_counter() - t) * 1e3, st))
            P.release()
        per_mode[mode] = (float(np.mean([r[0] for r in runs])), runs[-1][1], tuned is not None)
    cb.set_config(**base)
    mode = min(per_mode, key=lambda k: per_mode[k][0])
    ms, st, _ = per_mode[mode]
    # FP32 rounding on a sample of 64 columns against the same product in FP64
    rng = np.random.default_rng(7)
    qs = np.sort(rng.choice(A.nzc, 64, replace=False))
    idx = np.concatenate([np.arange(A.cp[q], A.cp[q + 1]) for q in qs])
    Bs = cb.Dcsc.from_coo(A.m, A.n, A.ir[idx], np.repeat(A.jc[qs], np.diff(A.cp)[qs]), A.val[idx], "float32")
    C32, s32 = cb.spgemm(dA, Bs.to_device(), "plus_times")
    A64 = A.astype("float64")
    C64, _ = cb.spgemm(A64.to_device(), Bs.astype("float64").to_device(), "plus_times")
    h32, h64 = C32.to_host(), C64.to_host()
    same = np.array_equal(h32.cp, h64.cp) and np.array_equal(h32.ir, h64.ir)
    rel = np.abs(h32.val.astype(np.float64) - h64.val) / np.maximum(np.abs(h64.val), 1e-300)
    for X in (C32, C64, dA):
        X.release()
    return {"workload": f"MCL expansion A*A + MCLPruneRecoverySelect(prune 1e-4, select 1000), FP32, "
                        f"column-stochastic power-law n={n} (alpha 2.1, mean degree 50, configuration model, "
                        f"log-normal weights), column phases sized to free memory, exact or learned nnz sizing",
            "ms_per_expansion": round(ms, 2), "gflops": round(2.0 * st["flops"] / (ms * 1e-3) / 1e9, 2),
            "multiply_adds": st["flops"], "nnz_a": A.nnz, "nnz_product": st["nnz_product"], "nnz_kept": st["nnz"],
            "phases": st["phases"], "selected_cols": st["selected_cols"],
            "ms_product": round(st["ms_product"], 2), "ms_prune": round(st["ms_prune"], 2), "mode": mode,
            "modes": {k: {"ms_per_expansion": round(v[0], 2), "nnz_kept": v[1]["nnz"], "phases": v[1]["phases"],
                          "overflow_units": v[1]["overflow_cols"], "learned_fallback": v[1]["learned_fallback"],
                          "tuned": v[2]}
                      for k, v in per_mode.items()},
            "fp32_vs_fp64_sample": {"columns": 64, "same_pattern": bool(same),
                                    "max_rel_err": float(rel.max()) if rel.size else 0.0,
                                    "mean_rel_err": float(rel.mean()) if rel.size else 0.0}}


def bench_galerkin(cb, count: int = 256, lo: int = 16, hi: int = 96, reps: int = 2) -> dict:
    from combblas_amd.checks import stencil27_coarse_rowsums
    rng = np.random.default_rng(0xC5)
    grids = [int(g) for g in rng.integers(lo, hi + 1, count)]
    As, Ts = [], []
    nnz_a = 0
    for g in grids:
        A = cb.gen_stencil27(g)
        nnz_a += A.nnz
        As.append(A.to_device())
        Ts.append(cb.gen_aggregation(g, True).to_device())
        del A
    from combblas_amd import autotune
    base = cb.get_config()
    tuned = autotune.apply("galerkin_256")
    S, st = cb.galerkin_batched(As, Ts)
    for X in S:
        X.release()
    runs = []
    for i in range(reps):
        t = time.perf_counter()
        S, st = cb.galerkin_batched(As, Ts)
        cb.synchronize()
        runs.append(((time.perf_counter() - t) * 1e3, st))
        if i + 1 < reps:
            for X in S:
                X.release()
    cb.set_config(**base)
    ms = float(np.mean([r[0] for r in runs]))
    st = runs[-1][1]
    rowsums_ok = True
    for g, X in zip(grids, S):  # closed-form row sums and symmetry of every coarse operator
        h = X.to_host()
        rows = np.bincount(h.ir, weights=h.val, minlength=h.m)
        cols = h.col_ids()
        key = np.sort(h.ir * h.n + cols)
        keyt = np.sort(cols * h.m + h.ir)
        rowsums_ok &= bool(np.array_equal(rows, stencil27_coarse_rowsums(g)) and np.array_equal(key, keyt))
        X.release()
    for X in As + Ts:
        X.release()
    return {"workload": f"{count} batched Galerkin products S*(A*T), T = R^T 2x2x2 aggregation, A 27-point "
                        f"Poisson stencil, grid edge uniform in [{lo},{hi}] (seed 0xC5), FP64, one packed batch",
            "ms_per_batch": round(ms, 2), "gflops": round(2.0 * st["flops"] / (ms * 1e-3) / 1e9, 2),
            "multiply_adds": st["flops"], "nnz_a_total": nnz_a, "nnz_out_total": st["nnz"],
            "rowsums_and_symmetric_pattern_all": rowsums_ok, "tunables": tuned}


def bench_spmv(cb, scale: int, ef: int, reps: int = 20) -> dict:
    """dcsc_gespmv (y += A*x, dense x) on the SpGEMM workload's matrix, PlusTimes<int64>, x and y
    resident on the device; checked against numpy outside the timed region."""
    A = cb.gen_rmat(scale, ef)
    dA = A.to_device()
    rng = np.random.default_rng(7)
    x = rng.integers(-9, 10, A.n).astype(np.int64)
    dx = cb.DeviceArray.from_host(x)
    dy = cb.DeviceArray.from_host(np.zeros(A.m, np.int64))
    cb.spmv_dense_device(dA, dx, dy)  # warm-up
    cb.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        cb.spmv_dense_device(dA, dx, dy)
    cb.synchronize()
    ms = (time.perf_counter() - t) * 1e3 / reps
    got = dy.to_host()
    want = np.zeros(A.m, np.int64)
    np.add.at(want, A.ir, A.val * x[A.col_ids()])
    exact = bool(np.array_equal(got, want * (reps + 1)))
    alg = 16 * A.jc.size + 8 + 16 * A.nnz + 8 * A.jc.size + 16 * A.m  # A, the x entries used, y read+written
    for d in (dx, dy, dA):
        d.release()
    return {"workload": f"dense SpMV y += A*x (dcsc_gespmv), R-MAT s{scale} ef{ef}, PlusTimes<int64>, x and y in HBM",
            "ms_per_spmv": round(ms, 3), "nnz_a": A.nnz, "alg_bytes": alg,
            "gbps_algorithmic": round(alg / (ms * 1e-3) / 1e9, 1), "exact_after_all_reps": exact}


def bench_merge(cb, scale: int = 17, ef: int = 16, nstages: int = 4, reps: int = 3) -> dict:
    """SUMMA's stage merge (MultiwayMerge<PlusTimes<int64>>) on one GPU: the
    local products of nstages stages of R-MAT A*A, A(:, K_s) * B(K_s, :),
    merged on the device; the merged product is checked against A*A with the
    full-size column checksums outside the timed region."""
    from combblas_amd.checks import product_matches
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
    from combblas_amd import autotune
    base = cb.get_config()
    tuned = autotune.apply("merge_s17") is not None if (scale, ef, nstages) == (17, 16, 4) else False
    try:
        M = cb.merge(parts)  # warm-up
        M.release()
        cb.synchronize()
        t = time.perf_counter()
        for i in range(reps):
            M = cb.merge(parts)
            if i + 1 < reps:
                M.release()
        cb.synchronize()
    finally:
        cb.set_config(**base)
    ms = (time.perf_counter() - t) * 1e3 / reps
    exact = product_matches(M, A, A)
    nnz_in = sum(p.nnz for p in parts)
    alg = sum(16 * p.nnz + 16 * p.nzc + 8 for p in parts) + 16 * M.nnz + 16 * M.nzc + 8
    out = {"workload": f"stage merge of {nstages} SUMMA stage products of R-MAT s{scale} ef{ef} A*A, "
                       "PlusTimes<int64>, DCSC in HBM",
           "ms_per_merge": round(ms, 3), "nnz_in": nnz_in, "nnz_out": M.nnz, "alg_bytes": alg,
           "gbps_algorithmic": round(alg / (ms * 1e-3) / 1e9, 1), "merged_equals_product": exact, "tuned": tuned}
    M.release()
    for p in parts:
        p.release()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--edgefactor", type=int, default=16)
    ap.add_argument("--mode", default="both", choices=["both", "two_phase", "learned"])
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="CPU-baseline budget (0 disables)")
    ap.add_argument("--no-extra", action="store_true", help="skip the BFS, MCL and Galerkin workloads")
    ap.add_argument("--only", choices=["bfs", "mcl", "galerkin", "spmv", "merge"], default=None,
                    help="run only this workload and print its JSON (for profiling)")
    ap.add_argument("--copy-gib", type=float, default=4.0,
                    help="size of a device-to-device copy timed before the workload: the achievable HBM rate, and "
                         "the known-size kernel that scales the FETCH_SIZE / WRITE_SIZE counters (0 disables)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU's block of the local product; --gpus {args.gpus} is not supported "
              "(the node-local kernel runs on one GPU; across GPUs each process of SUMMA's grid runs its own "
              "local products and stage merge).", file=sys.stderr)
        return 2

    import combblas_amd as cb
    from combblas_amd import predictor

    if cb.device_count() < 1:
        raise SystemExit("no GPU visible")
    if args.only:
        fn = {"bfs": lambda: bench_bfs(cb, args.scale, args.edgefactor), "mcl": lambda: bench_mcl(cb),
              "galerkin": lambda: bench_galerkin(cb),
              "spmv": lambda: bench_spmv(cb, args.scale, args.edgefactor),
              "merge": lambda: bench_merge(cb)}[args.only]
        print(json.dumps({args.only: fn()}))
        return 0
    copy = None
    if args.copy_gib > 0:
        nb = int(args.copy_gib * 2 ** 30)
        src, dst = cb.DeviceArray(nb, np.uint8), cb.DeviceArray(nb, np.uint8)
        cb.lib().cbx_memcpy(dst.ptr, src.ptr, nb, 2)
        cb.synchronize()
        t = time.perf_counter()
        cb.lib().cbx_memcpy(dst.ptr, src.ptr, nb, 2)
        cb.synchronize()
        dt = time.perf_counter() - t
        copy = {"bytes": nb, "gbps_read_plus_write": round(2 * nb / dt / 1e9, 1)}
        src.release()
        dst.release()
    A = cb.gen_rmat(args.scale, args.edgefactor)
    dA = cb.Dcsc.to_device(A)
    from combblas_amd import autotune
    default_cfg = cb.get_config()
    tuned = {}

    # exact entry count of every output column (the two-phase symbolic pass,
    # outside any timed region): with the checksums it pins the full-size output
    col_nnz = cb.spgemm_symbolic(dA, dA)
    modes = ["two_phase", "learned"] if args.mode == "both" else [args.mode]
    if "learned" in modes:
        if predictor.MODEL_PATH.exists():
            cb.set_predictor(predictor.load())
        else:
            log("no trained predictor: learned mode skipped")
            modes.remove("learned")
    res = {}
    for mode in modes:
        # the autotuner's stored choice for this workload and mode, if any
        name = f"rmat_s{args.scale}_ef{args.edgefactor}" + ("_learned" if mode == "learned" else "")
        cb.set_config(**default_cfg)
        tuned[mode] = autotune.apply(name)
        log(f"spgemm {mode}: {args.warmup} warmup + {args.steps} steps, tunables {cb.get_config()}")
        res[mode] = run_spgemm(cb, A, dA, mode, args.steps, args.warmup, col_nnz)
        log(f"spgemm {mode}: {res[mode]['ms']:.2f} ms/step, exact={res[mode]['exact']}")
    baseline = None
    if "learned" in res:  # the same single pass sized without the network: the reference's estimator
        cb.set_predictor(None)
        cb.set_sampling_predictor(5)
        rb = run_spgemm(cb, A, dA, "learned", 1, 1, col_nnz)
        sb = rb["stats"][-1]
        baseline = {"predictor": "estimateNNZ_sampling min-hash estimates per column (5 rounds, mtSpGEMM.h:936-1054)",
                    "ms_per_step": round(rb["ms"], 3), "output_bit_exact": rb["exact"],
                    "overflow_units": sb["overflow_cols"], "units": sb["units_light"],
                    "overflow_pct": round(100.0 * sb["overflow_cols"] / max(sb["units_light"], 1), 3),
                    "fallback": int(sb["learned_fallback"]),
                    "predicted_vs_actual_nnz": (round(sb["predicted_nnz"] / sb["actual_nnz_predicted"], 4)
                                                if sb["actual_nnz_predicted"] else None)}
        cb.set_sampling_predictor(0)
        cb.set_predictor(predictor.load())
    # `value` is the faster of the measured exact paths; both are reported
    exact_modes = [m for m in res if res[m]["exact"]] or list(res)
    primary = min(exact_modes, key=lambda m: res[m]["ms"])
    r = res[primary]
    st_last = r["stats"][-1]
    nnz, flops, nzc = st_last["nnz"], st_last["flops"], st_last["nzc"]
    ms = r["ms"]
    gflops = 2.0 * flops / (ms * 1e-3) / 1e9
    bytes_step = dcsc_bytes(A.nnz, A.nzc) * 2 + dcsc_bytes(nnz, nzc)
    if primary == "two_phase":  # the symbolic pass re-reads the structures and writes counts
        bytes_step += 16 * A.nzc + 8 * A.nnz + 16 * A.nzc + 8 * A.nnz + 8 * A.nzc
    gbps = bytes_step / (ms * 1e-3) / 1e9
    main_ms = float(np.mean([s["ms_main_kernel"] for s in r["stats"]]))
    main_nnz = st_last["nnz_main"]
    main_bytes = 16 * main_nnz + dcsc_bytes(A.nnz, A.nzc) * 2
    achieved = main_bytes / (main_ms * 1e-3) / 1e9 if main_ms > 0 else None
    kernel = "k_num_learned" if primary == "learned" else "unit kernels (k_num_units, k_num_list, k_num_units_rm)"
    traffic = None  # HBM bytes per launch from the committed rocprofv3 --pmc summary of this kernel
    tpath = os.path.join(ROOT, "profiles", f"traffic_rmat_s{args.scale}_ef{args.edgefactor}.json")
    if os.path.exists(tpath):
        t = json.load(open(tpath))
        k = t["kernels"].get(primary)
        if k:
            traffic = int((k["fetch_kib"] * t["fetch_scale"] + k["write_kib"] * t["write_scale"]) * 1024)

    per_mode = {}
    for mode, rr in res.items():
        s = rr["stats"][-1]
        per_mode[mode] = {
            "ms_per_step": round(rr["ms"], 3), "gflops": round(2.0 * s["flops"] / (rr["ms"] * 1e-3) / 1e9, 3),
            "output_bit_exact": rr["exact"],
            "phase_ms": {k: round(float(np.mean([x[k] for x in rr["stats"]])), 3)
                         for k in ("ms_setup", "ms_symbolic", "ms_alloc", "ms_numeric", "ms_main_kernel",
                                   "ms_total")}}
    speedup = None
    mispredict = None
    if "learned" in res and "two_phase" in res:
        speedup = round(res["two_phase"]["ms"] / res["learned"]["ms"], 4)
        ls = res["learned"]["stats"]
        units = ls[-1]["units_light"]
        mispredict = {
            "units": units, "overflow_units": ls[-1]["overflow_cols"],
            "overflow_pct": round(100.0 * ls[-1]["overflow_cols"] / max(units, 1), 3),
            "fallback_steps": int(sum(x["learned_fallback"] for x in ls)),
            "predicted_vs_actual_nnz": round(ls[-1]["predicted_nnz"] / max(ls[-1]["actual_nnz_predicted"], 1), 4),
            "non_learned_baseline": baseline}

    out = {
        "metric": METRIC, "value": round(gflops, 3), "unit": "GFLOPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int64", "data": "synthetic",
        "config": {"workload": f"R-MAT s{args.scale} ef{args.edgefactor} A*A, PlusTimes<int64>, DCSC int64 in/out",
                   "mode": primary, "nnz_a": A.nnz, "nnz_c": nnz, "multiply_adds": flops},
        "hbm_gbps_alg": round(gbps, 1), "hbm_roofline_frac_step": round(gbps / HBM_PEAK_GBS, 4),
        "bytes_alg_step": bytes_step,
        "learned_vs_two_phase_speedup": speedup, "learned_mispredict": mispredict, "modes": per_mode,
        "output_bit_exact": all(rr["exact"] for rr in res.values()),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                     "traffic": traffic, "kernel": kernel, "alg_bytes_per_launch": main_bytes,
                     "kernel_ms": round(main_ms, 3)},
        "cpu_baseline": None,
        "device_copy": copy, "tunables": {m: (t or default_cfg) for m, t in tuned.items()},
    }
    dA.release()
    cb.release_cached_memory()
    cb.set_config(**default_cfg)  # the other workloads apply their own stored tunables
    if not args.no_extra:
        for name, fn in (("bfs", lambda: bench_bfs(cb, args.scale, args.edgefactor)),
                         ("mcl", lambda: bench_mcl(cb)), ("galerkin", lambda: bench_galerkin(cb)),
                         ("spmv", lambda: bench_spmv(cb, args.scale, args.edgefactor)),
                         ("merge", lambda: bench_merge(cb))):
            log(f"{name} ...")
            try:
                out[name] = fn()
            except Exception as e:  # noqa: BLE001 - one workload failing must not hide the others
                out[name] = {"error": f"{type(e).__name__}: {e}"}
            log(f"{name}: {json.dumps(out[name])}")
            cb.release_cached_memory()
    if args.cpu_seconds > 0:
        bfs = out.get("bfs") or {}
        cb_out = cpu_baseline(args.scale, args.edgefactor, args.cpu_seconds, bfs.get("_roots"),
                              bfs.get("_edges_total"))
        if cb_out and "bfs" in cb_out:
            bfs["cpu_baseline"] = cb_out.pop("bfs")
        out["cpu_baseline"] = cb_out
    if isinstance(out.get("bfs"), dict):
        out["bfs"].pop("_roots", None)
        out["bfs"].pop("_edges_total", None)
    print(json.dumps(out))
    return 0 if out["output_bit_exact"] else 1


if __name__ == "__main__":
    sys.exit(main())
