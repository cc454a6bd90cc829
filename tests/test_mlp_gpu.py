"""The predictor's device kernels against the FP64 numpy oracle (oracle/mlp.py,
whose backward pass is itself checked by finite differences in
tests/test_mlp_oracle.py): forward outputs and full-batch gradients agree
within FP32 accumulation error; a short training run lowers the loss."""
from __future__ import annotations

import numpy as np
import pytest

import combblas_amd as cb
from combblas_amd import predictor
from oracle import mlp

pytestmark = pytest.mark.gpu


def sample_data(n=3000, seed=0):
    rng = np.random.default_rng(seed)
    X = np.column_stack([rng.uniform(0, 10, n), rng.uniform(0, 20, n), rng.uniform(0, 12, n), rng.uniform(0, 8, n),
                         rng.uniform(0, 1, n), rng.uniform(0, 18, n)]).astype(np.float32)
    y = (0.8 * X[:, 1] - 0.1 * X[:, 0] + np.sin(X[:, 4] * 3)).astype(np.float32)
    return X, y


def test_forward_matches_oracle():
    X, y = sample_data()
    P = predictor.init_params(X, seed=3)
    got = cb.mlp_forward(P, X)
    want, _ = mlp.forward(P, X)
    assert np.allclose(got, want, rtol=1e-4, atol=1e-4)


def test_gradient_matches_oracle():
    X, y = sample_data(2000, 1)
    P = predictor.init_params(X, seed=5)
    g_dev, loss_dev = cb.mlp_gradient(P, X, y)
    loss, g = mlp.loss_and_grad(P, X, y)
    assert abs(loss_dev - loss) <= 1e-4 * max(1.0, loss)
    scale = np.abs(g).max()
    assert np.max(np.abs(g_dev - g)) <= 1e-4 * scale, np.max(np.abs(g_dev - g)) / scale
    assert np.all(g_dev[-12:] == 0)


def test_training_reduces_loss():
    X, y = sample_data(20000, 2)
    P = predictor.init_params(X, seed=1)
    P2, curve = cb.mlp_train(P, X, y, steps=300, batch=1024, lr=3e-3)
    assert curve[-20:].mean() < 0.2 * curve[:20].mean()
    l0, _ = mlp.loss_and_grad(P, X, y)
    l1, _ = mlp.loss_and_grad(P2, X, y)
    assert l1 < 0.2 * l0


def test_features_and_symbolic_labels():
    A = cb.gen_rmat(12, 8)
    dA = A.to_device()
    X = cb.spgemm_features(dA, dA)
    nnz = cb.spgemm_symbolic(dA, dA)
    import oracle
    flop, want = oracle.symbolic(A, A)
    assert np.array_equal(nnz, want)
    assert np.allclose(np.exp2(X[:, 1]) - 1, flop, rtol=1e-5)
    b = np.diff(A.cp)
    assert np.allclose(np.exp2(X[:, 0]) - 1, b, rtol=1e-5)


@pytest.mark.parametrize("dtype", ["float32", "bool", "float64"])
def test_symbolic_labels_any_value_type(dtype):
    """The symbolic pass alone reads no values, so it serves operands of every
    value type (the power-law model's labels come from float32 matrices; the
    symbolic-only call used to run the tiny tier's staged numeric form, which
    read a 1- or 4-byte value array as int64, past its end): per-column counts
    equal to the oracle's on the pattern, with tiny, hash and unit-tier
    columns present."""
    import oracle
    A = cb.gen_rmat(13, 8).astype(dtype)
    dA = A.to_device()
    nnz = cb.spgemm_symbolic(dA, dA)
    P = A.astype("int64")
    P.val = np.ones(P.nnz, np.int64)
    flop, want = oracle.symbolic(P, P)
    assert np.array_equal(nnz, want)
    assert (flop <= 256).any() and (flop > 4096).any()
