"""Per-column nnz predictor for the learned single-pass SpGEMM.

Training data are (features, exact nnz) pairs that the two-phase path's own
symbolic pass produces on generated products A*A.  Two models, one per input
family:

* "rmat": 32 R-MAT matrices, 8 at each scale 16..19, edge factor 16, seeds
  1000+i (held out: the s20 ef16 benchmark matrix);
* "powerlaw": the MCL expansion's products at its own size: three
  column-stochastic power-law matrices of its generator (n = 10^6, alpha 2.1,
  mean degree 50, seeds 2000..2002), each A times six column slices of
  itself, the way cbx_mcl_expand splits B into phases (held out: the
  expansion matrix itself, two of its slices).

The network (6 -> 64 -> 64 -> 1, ReLU, FP32) is trained on the GPU by the
library's own kernels (cbx_mlp_train) and predicts log2(1 + nnz_j).  Trained
parameters are kept under combblas_amd/models/ with a JSON record of how
they were made.

    python -m combblas_amd.predictor train [rmat|powerlaw]    # on a GPU machine
"""
from __future__ import annotations

import json
import sys
import time
from pathlib import Path

import numpy as np

import combblas_amd as cb

MODEL_DIR = Path(__file__).resolve().parent / "models"
MODEL_PATH = MODEL_DIR / "nnz_mlp_rmat.f32"
META_PATH = MODEL_DIR / "nnz_mlp_rmat.json"
KINDS = ("rmat", "powerlaw")

TRAIN_SCALES = (16, 17, 18, 19)
PER_SCALE = 8
EDGE_FACTOR = 16
POWERLAW_N = 1_000_000
POWERLAW_SEEDS = (2000, 2001, 2002)
POWERLAW_SLICES = 6


def model_path(kind: str = "rmat") -> Path:
    return MODEL_DIR / f"nnz_mlp_{kind}.f32"


def training_matrix(i: int) -> tuple[int, int]:
    """(scale, seed) of R-MAT training matrix i (0..31)."""
    return TRAIN_SCALES[i // PER_SCALE], 1000 + i


def training_set(kind: str) -> list[tuple[dict, "callable"]]:
    """(description, generator of the (A, B) pairs) of every training matrix of
    a model family."""
    if kind == "rmat":
        return [({"scale": sc, "seed": sd, "edge_factor": EDGE_FACTOR},
                 (lambda sc=sc, sd=sd: [(cb.gen_rmat(sc, EDGE_FACTOR, seed=sd),) * 2]))
                for sc, sd in map(training_matrix, range(len(TRAIN_SCALES) * PER_SCALE))]
    if kind == "powerlaw":
        return [({"n": POWERLAW_N, "seed": sd, "column_slices": POWERLAW_SLICES},
                 (lambda sd=sd: column_slices(cb.gen_powerlaw_stochastic(POWERLAW_N, seed=sd), POWERLAW_SLICES)))
                for sd in POWERLAW_SEEDS]
    raise KeyError(kind)


def column_slices(A: cb.Dcsc, k: int, which=None) -> list[tuple[cb.Dcsc, cb.Dcsc]]:
    """(A, A(:, J_i)) for k equal column ranges J_i (or those listed in which)."""
    edges = np.linspace(0, A.n, k + 1).astype(np.int64)
    return [(A, A.column_range(int(edges[i]), int(edges[i + 1]))) for i in (which or range(k))]


def heldout_pairs(kind: str) -> list[tuple[cb.Dcsc, cb.Dcsc]]:
    """The benchmark workload's own products, never trained on."""
    if kind == "rmat":
        A = cb.gen_rmat(20, EDGE_FACTOR)
        return [(A, A)]
    return column_slices(cb.gen_powerlaw_stochastic(POWERLAW_N), POWERLAW_SLICES, which=[0, 3])


def pairs_for(A: cb.Dcsc, B: cb.Dcsc | None = None) -> tuple[np.ndarray, np.ndarray]:
    """(features float32[n,6], log2(1+nnz) float32[n]) for the columns of A*B
    (B = A by default)."""
    dA = A.to_device()
    dB = dA if B is None or B is A else B.to_device()
    X = cb.spgemm_features(dA, dB)
    nnz = cb.spgemm_symbolic(dA, dB)
    keep = nnz > 0
    return X[keep], np.log2(1.0 + nnz[keep]).astype(np.float32)


def collect_pairs(pairs, log=print, label="") -> tuple[np.ndarray, np.ndarray]:
    Xs, ys = [], []
    for A, B in pairs:
        X, y = pairs_for(A, B)
        Xs.append(X)
        ys.append(y)
        log(f"{label}: {len(y)} columns")
    return np.concatenate(Xs), np.concatenate(ys)


def collect(kind: str = "rmat", log=print) -> tuple[np.ndarray, np.ndarray]:
    Xs, ys = [], []
    for i, (desc, gen) in enumerate(training_set(kind)):
        X, y = collect_pairs(gen(), log, f"{kind} matrix {i:2d}: {desc}")
        Xs.append(X)
        ys.append(y)
    return np.concatenate(Xs), np.concatenate(ys)


def init_params(X: np.ndarray, seed: int = 1) -> np.ndarray:
    """He-initialised weights, zero biases, feature mean / std from X."""
    F, H = cb.MLP_FEATURES, cb.MLP_HIDDEN
    rng = np.random.default_rng(seed)
    sd = X.std(0)
    sd[sd < 1e-6] = 1.0
    parts = [rng.normal(0, np.sqrt(2 / F), (H, F)), np.zeros(H), rng.normal(0, np.sqrt(2 / H), (H, H)), np.zeros(H),
             rng.normal(0, np.sqrt(1 / H), H), np.zeros(1), X.mean(0), sd]
    return np.concatenate([np.asarray(p, np.float64).reshape(-1) for p in parts]).astype(np.float32)


def relative_error(params: np.ndarray, X: np.ndarray, y: np.ndarray) -> dict:
    """Prediction quality on (X, y): errors of nnz = 2^y - 1."""
    pred = cb.mlp_forward(params, X)
    true_nnz = np.exp2(y.astype(np.float64)) - 1
    pred_nnz = np.exp2(pred.astype(np.float64)) - 1
    rel = (pred_nnz - true_nnz) / np.maximum(true_nnz, 1)
    return {"mean_abs_rel": float(np.mean(np.abs(rel))), "median_abs_rel": float(np.median(np.abs(rel))),
            "p95_abs_rel": float(np.quantile(np.abs(rel), 0.95)),
            "total_rel": float((pred_nnz.sum() - true_nnz.sum()) / true_nnz.sum()),
            "under_predicted_frac": float(np.mean(pred_nnz < true_nnz))}


def train(kind: str = "rmat", steps: int = 4000, batch: int = 4096, lr: float = 2e-3, seed: int = 7, log=print,
          out_dir: Path | None = None) -> dict:
    t0 = time.time()
    X, y = collect(kind, log=log)
    t1 = time.time()
    P0 = init_params(X)
    P, curve = cb.mlp_train(P0, X, y, steps=steps, batch=batch, lr=lr, seed=seed)
    t2 = time.time()
    q_train = relative_error(P, X, y)
    Xh, yh = collect_pairs(heldout_pairs(kind), log, f"{kind} held out")
    q_held = relative_error(P, Xh, yh)
    out_dir = Path(out_dir) if out_dir else MODEL_DIR
    out_dir.mkdir(parents=True, exist_ok=True)
    P.astype(np.float32).tofile(out_dir / model_path(kind).name)
    meta = {"architecture": "6-64-64-1 ReLU, FP32, target log2(1+nnz_j), MSE, Adam", "family": kind,
            "training_matrices": [d for d, _ in training_set(kind)],
            "samples": int(len(y)), "steps": steps, "batch": batch, "lr": lr,
            "seed": seed, "loss_first": float(curve[:50].mean()), "loss_last": float(curve[-50:].mean()),
            "collect_s": round(t1 - t0, 2), "train_s": round(t2 - t1, 2), "train_quality": q_train,
            ("heldout_rmat_s20_ef16" if kind == "rmat" else "heldout_powerlaw_n1m"): q_held}
    (out_dir / f"nnz_mlp_{kind}.json").write_text(json.dumps(meta, indent=2) + "\n")
    log(json.dumps(meta, indent=2))
    return meta


def ratio_params(ratio: float) -> np.ndarray:
    """A non-learned predictor in the network's parameter format:
    log2(1 + nnz_j) = log2(1 + flop_j) + log2(ratio), i.e. every column's size
    is its flop count times one global compression ratio."""
    F, H = cb.MLP_FEATURES, cb.MLP_HIDDEN
    P = np.zeros(cb.MLP_PARAMS, np.float32)
    oW1, oB1 = 0, H * F
    oW2 = oB1 + H
    oB2 = oW2 + H * H
    oW3 = oB2 + H
    oB3 = oW3 + H
    oMU = oB3 + 1
    oSD = oMU + F
    P[oW1 + 0 * F + 1] = 1.0  # hidden unit 0 = log2(1 + flop) (>= 0, passes the ReLU)
    P[oW2 + 0 * H + 0] = 1.0
    P[oW3 + 0] = 1.0
    P[oB3] = np.log2(ratio)
    P[oSD:oSD + F] = 1.0
    return P


def load(kind: str = "rmat") -> np.ndarray:
    path = model_path(kind)
    if not path.exists():
        raise FileNotFoundError(f"{path} missing: train it with `python -m combblas_amd.predictor train {kind}`")
    P = np.fromfile(path, np.float32)
    if P.size != cb.MLP_PARAMS:
        raise ValueError(f"{path}: expected {cb.MLP_PARAMS} floats, found {P.size}")
    return P


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "train":
        kind = sys.argv[2] if len(sys.argv) > 2 else "rmat"
        train(kind, out_dir=Path(sys.argv[3]) if len(sys.argv) > 3 else None)
    else:
        print(__doc__)
