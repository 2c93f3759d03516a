"""TEST INFRASTRUCTURE ONLY: FP64 numpy restatement of the 6 -> 64 -> 64 -> K cost-model MLP.

No reference counterpart (the reference hard-codes its dispatch tiers, kernels/mod.rs:63-106).
This is the checker for the GPU forward, backward and Adam kernels (numrs_amd/csrc/selector.hip):

  h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), y = W3 h2 + b3          (W1: 64x6, W2: 64x64, W3: Kx64)
  loss = sum(mask * (y - t)^2) / sum(mask)                              (mask in {0, 1})
  Adam (Kingma & Ba 2015): m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2,
        p -= lr * (m / (1 - b1^step)) / (sqrt(v / (1 - b2^step)) + eps)

The parameter vector is laid out W1, b1, W2, b2, W3, b3 (row-major), as the device kernels read it.
The gradient is pinned against central finite differences in tests/test_oracle.py.
"""
from __future__ import annotations

import numpy as np

IN, H = 6, 64


def param_count(K: int) -> int:
    return H * IN + H + H * H + H + K * H + K


def unpack(p: np.ndarray, K: int):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + H * IN].reshape(H, IN); o += H * IN
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + K * H].reshape(K, H); o += K * H
    b3 = p[o:o + K]; o += K
    return W1, b1, W2, b2, W3, b3


def forward(p: np.ndarray, K: int, x: np.ndarray):
    """y (N, K) and the activations (z1, h1, z2, h2)."""
    W1, b1, W2, b2, W3, b3 = unpack(p, K)
    x = np.asarray(x, np.float64)
    z1 = x @ W1.T + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0.0)
    y = h2 @ W3.T + b3
    return y, (z1, h1, z2, h2)


def loss_and_grad(p: np.ndarray, K: int, x: np.ndarray, t: np.ndarray, mask: np.ndarray):
    """(sum(mask (y-t)^2), mean loss, gradient of the mean loss w.r.t. the parameter vector)."""
    W1, b1, W2, b2, W3, b3 = unpack(p, K)
    x = np.asarray(x, np.float64)
    y, (z1, h1, z2, h2) = forward(p, K, x)
    mask = np.asarray(mask, np.float64)
    d = (y - np.asarray(t, np.float64)) * mask
    total = float(np.sum(d * d))
    count = float(np.sum(mask))
    dy = 2.0 * d / count if count > 0 else np.zeros_like(d)
    gW3 = dy.T @ h2
    gb3 = dy.sum(axis=0)
    dh2 = (dy @ W3) * (z2 > 0)
    gW2 = dh2.T @ h1
    gb2 = dh2.sum(axis=0)
    dh1 = (dh2 @ W2) * (z1 > 0)
    gW1 = dh1.T @ x
    gb1 = dh1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return total, (total / count if count > 0 else 0.0), g


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update; returns (p, m, v)."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    p = p - lr * (m / (1.0 - b1 ** step)) / (np.sqrt(v / (1.0 - b2 ** step)) + eps)
    return p, m, v
