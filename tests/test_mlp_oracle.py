"""The FP64 MLP oracle's analytic gradient against central finite differences."""
import numpy as np

from oracle import mlp


def test_oracle_gradient_finite_differences():
    rng = np.random.default_rng(0)
    X = rng.normal(size=(40, 6))
    y = rng.normal(size=40)
    P = mlp.init_params(X).astype(np.float64)
    _, g = mlp.loss_and_grad(P, X, y)
    ntrain = 64 * 6 + 64 + 64 * 64 + 64 + 64 + 1
    for i in rng.choice(ntrain, 40, replace=False):
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += 1e-6
        Pm[i] -= 1e-6
        fd = (mlp.loss_and_grad(Pp, X, y)[0] - mlp.loss_and_grad(Pm, X, y)[0]) / 2e-6
        assert abs(fd - g[i]) <= 1e-4 * (abs(fd) + 1e-6)


def test_ratio_predictor_layout():
    """The non-learned baseline in the network's format computes log2(1+flop) + log2(ratio)."""
    from combblas_amd import predictor
    P = predictor.ratio_params(0.37).astype(np.float64)
    X = np.random.default_rng(0).uniform(0, 25, (50, 6))
    y, _ = mlp.forward(P, X)
    assert np.allclose(y, X[:, 1] + np.log2(0.37), atol=1e-6)
