"""FP64 restatement of the capacity MLP -- TEST INFRASTRUCTURE ONLY.

Same flat parameter layout as the device kernels (W1[32][8] b1[32] W2[32][32] b2[32]
W3[32] b3), forward y = W3 relu(W2 relu(W1 x + b1) + b2) + b3, loss mean((y - t)^2).
The analytic gradient is pinned by central finite differences (tests/test_oracle_mlp.py).
(No reference counterpart: the learned capacity is this project's addition.)
"""
from __future__ import annotations

import numpy as np

F, H = 8, 32
W1, B1 = 0, H * F
W2 = B1 + H
B2 = W2 + H * H
W3 = B2 + H
B3 = W3 + H
N_PARAMS = B3 + 1


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[W1:B1].reshape(H, F), p[B1:W2], p[W2:B2].reshape(H, H), p[B2:W3], p[W3:B3], p[B3])


def forward(p, X):
    w1, b1, w2, b2, w3, b3 = unpack(p)
    X = np.asarray(X, np.float64)
    z1 = X @ w1.T + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ w2.T + b2
    h2 = np.maximum(z2, 0.0)
    y = h2 @ w3 + b3
    return y, (X, z1, h1, z2, h2)


def loss_and_grad(p, X, t):
    w1, b1, w2, b2, w3, b3 = unpack(p)
    y, (X, z1, h1, z2, h2) = forward(p, X)
    t = np.asarray(t, np.float64)
    B = X.shape[0]
    r = y - t
    loss = float(np.mean(r * r))
    dy = 2.0 * r / B
    g = np.zeros(N_PARAMS)
    g[B3] = dy.sum()
    g[W3:B3] = h2.T @ dy
    dz2 = np.outer(dy, w3) * (z2 > 0)
    g[B2:W3] = dz2.sum(0)
    g[W2:B2] = (dz2.T @ h1).ravel()
    dz1 = (dz2 @ w2) * (z1 > 0)
    g[B1:W2] = dz1.sum(0)
    g[W1:B1] = (dz1.T @ X).ravel()
    return loss, g
