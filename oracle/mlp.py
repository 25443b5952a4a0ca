"""Test infrastructure only: a float64 numpy restatement of the served
scheduler's service-time network (efficient-sumcheck_amd/csrc/learn.hip), the
checker for the GPU forward and backward kernels.  It has no counterpart in
the reference (the reference has no learned component).

Network: x -> tanh(W1 x + b1) -> tanh(W2 h1 + b2) -> y = w3 . h2 + b3; loss
= mean((y - t)^2).  Flat parameter layout: W1 [H][D], b1 [H], W2 [H][H],
b2 [H], w3 [H], b3.  Pinned by central finite differences and by closed forms
(tests/test_oracle_mlp.py).
"""
import numpy as np


def count(D, H):
    return H * D + 3 * H + H * H + 1


def unpack(p, D, H):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + H * D].reshape(H, D); o += H * D
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def forward(p, X, D, H):
    W1, b1, W2, b2, w3, b3 = unpack(p, D, H)
    X = np.asarray(X, dtype=np.float64).reshape(-1, D)
    h1 = np.tanh(X @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return h2 @ w3 + b3


def loss_grad(p, X, T, D, H):
    """Mean squared error and its gradient in the flat layout."""
    W1, b1, W2, b2, w3, b3 = unpack(p, D, H)
    X = np.asarray(X, dtype=np.float64).reshape(-1, D)
    T = np.asarray(T, dtype=np.float64)
    n = len(T)
    h1 = np.tanh(X @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    y = h2 @ w3 + b3
    r = y - T
    loss = float(np.mean(r * r))
    dy = 2.0 * r / n                      # [n]
    g_b3 = dy.sum()
    g_w3 = h2.T @ dy                      # [H]
    d2 = np.outer(dy, w3) * (1 - h2 * h2)  # [n, H]
    g_W2 = d2.T @ h1
    g_b2 = d2.sum(0)
    d1 = (d2 @ W2) * (1 - h1 * h1)
    g_W1 = d1.T @ X
    g_b1 = d1.sum(0)
    g = np.concatenate([g_W1.ravel(), g_b1, g_W2.ravel(), g_b2, g_w3, [g_b3]])
    return loss, g
