"""CPU checks of the FP64 capacity-MLP oracle (oracle/mlp.py).

The analytic gradient is pinned by central finite differences of the loss for every one
of the 1377 parameters, and the forward pass by a per-sample evaluation written against
the flat layout (W1[j][k] at 8j + k etc.), so a transposed weight block or a dropped ReLU
gate fails one of them.
"""
import numpy as np

from oracle import mlp


def _sample(seed, B):
    rng = np.random.default_rng(seed)
    p = rng.normal(0.0, 0.4, mlp.N_PARAMS)
    X = rng.normal(0.0, 1.0, (B, mlp.F))
    t = rng.normal(5.0, 1.0, B)
    return p, X, t


def _forward_one(p, x):
    h1 = [max(0.0, p[mlp.B1 + j] + sum(p[mlp.W1 + j * mlp.F + k] * x[k] for k in range(mlp.F))) for j in range(mlp.H)]
    h2 = [max(0.0, p[mlp.B2 + j] + sum(p[mlp.W2 + j * mlp.H + k] * h1[k] for k in range(mlp.H))) for j in range(mlp.H)]
    return p[mlp.B3] + sum(p[mlp.W3 + k] * h2[k] for k in range(mlp.H))


def test_layout_and_size():
    assert mlp.N_PARAMS == 8 * 32 + 32 + 32 * 32 + 32 + 32 + 1 == 1377


def test_forward_matches_per_sample_evaluation():
    p, X, t = _sample(1, 7)
    y, _ = mlp.forward(p, X)
    for b in range(7):
        assert abs(y[b] - _forward_one(p, X[b])) < 1e-12 * max(1.0, abs(y[b]))
    loss, _ = mlp.loss_and_grad(p, X, t)
    assert abs(loss - np.mean((y - t) ** 2)) < 1e-12


def test_gradient_central_differences():
    p, X, t = _sample(2, 17)
    _, g = mlp.loss_and_grad(p, X, t)
    eps = 1e-6
    fd = np.empty(mlp.N_PARAMS)
    for j in range(mlp.N_PARAMS):
        a, b = p.copy(), p.copy()
        a[j] += eps
        b[j] -= eps
        fd[j] = (mlp.loss_and_grad(a, X, t)[0] - mlp.loss_and_grad(b, X, t)[0]) / (2 * eps)
    scale = np.abs(g).max()
    assert np.all(np.abs(fd - g) <= 1e-6 * np.abs(g) + 1e-7 * scale), np.abs(fd - g).max()
    # every block of parameters carries gradient signal on this sample
    for lo, hi in ((mlp.W1, mlp.B1), (mlp.B1, mlp.W2), (mlp.W2, mlp.B2), (mlp.B2, mlp.W3), (mlp.W3, mlp.B3)):
        assert np.abs(g[lo:hi]).max() > 1e-3
