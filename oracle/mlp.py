"""ORACLE - test infrastructure only.

FP64 numpy restatement of the selector MLP (24 -> 64 -> 64 -> K, ReLU),
its masked mean-squared-error loss, its gradients and one Adam step, over the
same flat parameter layout as the library:

    W1[64][24] b1[64] W2[64][64] b2[64] W3[K][64] b3[K]   (row-major [out][in])

Inputs are standardised (x - x_mean) / x_std; targets t (ln ms) are
standardised (t - y_mean) / y_std; the loss is
sum(mask * (y - t_s)^2) / sum(mask) over the standardised prediction y.

The reference has no learned selector, so nothing here restates reference
code. The backward pass is pinned by central finite differences
(tests/test_oracle_mlp.py), and the library's HIP kernels are checked against
this module on the GPU. Only tests/ may import it.
"""

from __future__ import annotations

import numpy as np

IN, HID = 24, 64


def n_params(K: int) -> int:
    return HID * IN + HID + HID * HID + HID + K * HID + K


def unpack(p, K):
    p = np.asarray(p, np.float64)
    o = 0

    def take(n, shape):
        nonlocal o
        a = p[o:o + n].reshape(shape)
        o += n
        return a

    W1 = take(HID * IN, (HID, IN))
    b1 = take(HID, (HID,))
    W2 = take(HID * HID, (HID, HID))
    b2 = take(HID, (HID,))
    W3 = take(K * HID, (K, HID))
    b3 = take(K, (K,))
    assert o == p.size
    return W1, b1, W2, b2, W3, b3


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([a.ravel() for a in (W1, b1, W2, b2, W3, b3)])


def forward(p, K, x, x_mean, x_std, y_mean=0.0, y_std=1.0, destandardise=True, keep=False):
    W1, b1, W2, b2, W3, b3 = unpack(p, K)
    xn = (np.asarray(x, np.float64) - np.asarray(x_mean, np.float64)) / np.asarray(x_std, np.float64)
    z1 = xn @ W1.T + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0.0)
    y = h2 @ W3.T + b3
    out = y * y_std + y_mean if destandardise else y
    if keep:
        return out, (xn, h1, h2)
    return out


def loss_and_grad(p, K, x, t, mask, x_mean, x_std, y_mean, y_std):
    W1, b1, W2, b2, W3, b3 = unpack(p, K)
    y, (xn, h1, h2) = forward(p, K, x, x_mean, x_std, y_mean, y_std, destandardise=False, keep=True)
    mask = np.asarray(mask, np.float64)
    ts = (np.asarray(t, np.float64) - y_mean) / y_std
    cnt = mask.sum()
    e = np.where(mask != 0, y - ts, 0.0)
    loss = float((mask * e * e).sum() / cnt)
    d3 = 2.0 * mask * e / cnt
    d2 = (d3 @ W3) * (h2 > 0)
    d1 = (d2 @ W2) * (h1 > 0)
    g = pack(d1.T @ xn, d1.sum(0), d2.T @ h1, d2.sum(0), d3.T @ h2, d3.sum(0))
    return loss, g


def loss_only(p, K, x, t, mask, x_mean, x_std, y_mean, y_std):
    y = forward(p, K, x, x_mean, x_std, y_mean, y_std, destandardise=False)
    mask = np.asarray(mask, np.float64)
    ts = (np.asarray(t, np.float64) - y_mean) / y_std
    e = np.where(mask != 0, y - ts, 0.0)
    return float((mask * e * e).sum() / mask.sum())


def adam(p, g, m, v, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step at step count t (1-based); returns (p, m, v)."""
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    c1 = 1 - beta1 ** t
    c2 = 1 - beta2 ** t
    return p - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v
