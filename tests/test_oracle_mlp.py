"""The numpy service-time network (the checker of the GPU kernels) pinned by
central finite differences and closed forms."""
import numpy as np

import mlp as M


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    D, H, n = 5, 7, 11
    p = rng.normal(0, 0.5, M.count(D, H))
    X = rng.normal(0, 1, (n, D))
    T = rng.normal(0, 1, n)
    _, g = M.loss_grad(p, X, T, D, H)
    eps = 1e-6
    num = np.empty_like(p)
    for i in range(len(p)):
        q = p.copy()
        q[i] += eps
        lp, _ = M.loss_grad(q, X, T, D, H)
        q[i] -= 2 * eps
        lm, _ = M.loss_grad(q, X, T, D, H)
        num[i] = (lp - lm) / (2 * eps)
    assert np.allclose(g, num, rtol=1e-5, atol=1e-8)


def test_closed_forms():
    D, H = 3, 4
    p = np.zeros(M.count(D, H))
    p[-1] = 0.75  # b3: the network is the constant 0.75
    X = np.ones((6, D))
    T = np.arange(6.0)
    assert np.allclose(M.forward(p, X, D, H), 0.75)
    loss, g = M.loss_grad(p, X, T, D, H)
    assert np.isclose(loss, np.mean((0.75 - T) ** 2))
    assert np.isclose(g[-1], 2 * np.mean(0.75 - T))
    # with zero hidden activations every other gradient vanishes except w3's, which is h2 = 0
    assert np.allclose(g[:-1], 0)
    # one hidden unit, identity-like: y = w3 tanh(W2 tanh(W1 x))
    p = np.zeros(M.count(1, 1))
    p[0], p[2], p[4] = 0.5, 0.25, 2.0  # W1, W2, w3
    x = np.array([[0.3]])
    y = M.forward(p, x, 1, 1)[0]
    assert np.isclose(y, 2.0 * np.tanh(0.25 * np.tanh(0.5 * 0.3)))
