"""ORACLE -- test infrastructure only.

FP64 numpy restatement of the nnz predictor (6 -> 64 -> 64 -> 1 ReLU MLP on
standardised features, mean squared error) used to check the device
forward/backward kernels in combblas_amd/csrc/cbx_mlp.hip.  The backward
pass is checked here independently by central finite differences.
"""
from __future__ import annotations

import numpy as np

F, H = 6, 64
SIZES = [("W1", (H, F)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H,)), ("b3", (1,)), ("mu", (F,)),
         ("sd", (F,))]


def unpack(P: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(P[o:o + n], np.float64).reshape(shape)
        o += n
    return out


def forward(P: np.ndarray, X: np.ndarray) -> tuple[np.ndarray, dict]:
    w = unpack(P)
    x = (np.asarray(X, np.float64) - w["mu"]) / w["sd"]
    a1 = x @ w["W1"].T + w["b1"]
    h1 = np.maximum(a1, 0)
    a2 = h1 @ w["W2"].T + w["b2"]
    h2 = np.maximum(a2, 0)
    out = h2 @ w["W3"] + w["b3"][0]
    return out, dict(x=x, a1=a1, h1=h1, a2=a2, h2=h2, w=w)


def loss_and_grad(P: np.ndarray, X: np.ndarray, y: np.ndarray) -> tuple[float, np.ndarray]:
    out, c = forward(P, X)
    B = len(y)
    e = out - np.asarray(y, np.float64)
    loss = float(np.mean(e * e))
    d3 = 2 * e / B
    w = c["w"]
    d2 = (d3[:, None] * w["W3"][None, :]) * (c["a2"] > 0)
    d1 = (d2 @ w["W2"]) * (c["a1"] > 0)
    g = dict(W1=d1.T @ c["x"], b1=d1.sum(0), W2=d2.T @ c["h1"], b2=d2.sum(0), W3=c["h2"].T @ d3, b3=np.array([d3.sum()]),
             mu=np.zeros(F), sd=np.zeros(F))
    return loss, np.concatenate([g[name].reshape(-1) for name, _ in SIZES])


def init_params(X: np.ndarray, seed: int = 1) -> np.ndarray:
    """He-initialised weights, zero biases, feature mean/std from X."""
    rng = np.random.default_rng(seed)
    mu = X.mean(0)
    sd = X.std(0)
    sd[sd < 1e-6] = 1.0
    parts = dict(W1=rng.normal(0, np.sqrt(2 / F), (H, F)), b1=np.zeros(H), W2=rng.normal(0, np.sqrt(2 / H), (H, H)),
                 b2=np.zeros(H), W3=rng.normal(0, np.sqrt(1 / H), H), b3=np.array([0.0]), mu=mu, sd=sd)
    return np.concatenate([np.asarray(parts[n], np.float64).reshape(-1) for n, _ in SIZES]).astype(np.float32)
