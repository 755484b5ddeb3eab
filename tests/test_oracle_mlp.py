"""The FP64 MLP oracle against central finite differences and closed forms
(CPU only)."""

import numpy as np
import pytest

from oracle import mlp


def _setup(K=7, n=5, seed=0):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, mlp.n_params(K))
    x = rng.normal(0, 2, (n, mlp.IN))
    t = rng.normal(0, 1, (n, K))
    mask = (rng.random((n, K)) < 0.8).astype(np.float64)
    mask[0, 0] = 1.0
    xm = rng.normal(0, 1, mlp.IN)
    xs = rng.uniform(0.5, 2.0, mlp.IN)
    return p, x, t, mask, xm, xs, 0.3, 1.7


def test_backward_matches_finite_differences():
    K = 7
    p, x, t, mask, xm, xs, ym, ys = _setup(K)
    loss, g = mlp.loss_and_grad(p, K, x, t, mask, xm, xs, ym, ys)
    assert loss == pytest.approx(mlp.loss_only(p, K, x, t, mask, xm, xs, ym, ys), rel=1e-14)
    rng = np.random.default_rng(1)
    idx = rng.choice(p.size, 300, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (mlp.loss_only(p + e, K, x, t, mask, xm, xs, ym, ys)
              - mlp.loss_only(p - e, K, x, t, mask, xm, xs, ym, ys)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_masked_entries_have_no_effect():
    K = 7
    p, x, t, mask, xm, xs, ym, ys = _setup(K)
    t2 = t.copy()
    t2[mask == 0] += 100.0
    l1, g1 = mlp.loss_and_grad(p, K, x, t, mask, xm, xs, ym, ys)
    l2, g2 = mlp.loss_and_grad(p, K, x, t2, mask, xm, xs, ym, ys)
    assert l1 == l2 and np.array_equal(g1, g2)


def test_linear_regime_closed_form():
    """With W1, W2 = identity-like positive paths and K = 1 the network is
    affine in x: y = c . xn + d; check the output against that closed form."""
    K = 1
    W1 = np.zeros((mlp.HID, mlp.IN))
    W1[np.arange(mlp.IN), np.arange(mlp.IN)] = 1.0
    b1 = np.zeros(mlp.HID)
    b1[:mlp.IN] = 100.0  # keeps every unit in the positive (linear) regime
    W2 = np.eye(mlp.HID)
    b2 = np.zeros(mlp.HID)
    rng = np.random.default_rng(3)
    W3 = rng.normal(size=(K, mlp.HID))
    b3 = np.array([0.5])
    p = mlp.pack(W1, b1, W2, b2, W3, b3)
    x = rng.normal(size=(4, mlp.IN))
    y = mlp.forward(p, K, x, np.zeros(mlp.IN), np.ones(mlp.IN), 0.0, 1.0)
    exp = x @ W3[0, :mlp.IN] + 100.0 * W3[0, :mlp.IN].sum() + 0.5
    assert np.allclose(y[:, 0], exp, rtol=1e-12, atol=1e-9)


def test_adam_first_step_is_sign_times_lr():
    """At t = 1 Adam moves every parameter by lr * g / (|g| + eps/...):
    the bias-corrected step is lr * sign(g) up to eps."""
    rng = np.random.default_rng(4)
    p = rng.normal(size=50)
    g = rng.normal(size=50)
    p1, _, _ = mlp.adam(p, g, np.zeros(50), np.zeros(50), 1, lr=0.01, eps=1e-12)
    assert np.allclose(p - p1, 0.01 * np.sign(g), rtol=1e-9)
