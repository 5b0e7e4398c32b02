"""ORACLE (test infrastructure only) -- masked autoregressive proposal network.

FP64 numpy restatement of the learned Metropolis proposal used by the
device kernels: a MADE (Germain, Gregor, Murray, Larochelle, ICML 2015) over
the D = 2N doubled binary variables (bit d of the packed configuration key),
two hidden ReLU layers of width H, sigmoid Bernoulli conditionals

    q(x) = prod_d sigma(l_d)^x_d (1 - sigma(l_d))^(1 - x_d),
    l = W3 relu(W2 relu(W1 x + b1) + b2) + b3,

with the masks of the natural ordering: input degree d + 1, hidden degrees
m(k) = (k mod (D - 1)) + 1, W1[k, d] kept iff m1(k) >= d + 1, W2[k', k] kept
iff m2(k') >= m1(k), W3[d, k'] kept iff d + 1 > m2(k').  The reference has no
such network (it is an addition of libnqhip, plugged in through the sampler's
`log_prob_bias` hook, src/Samplers/Metropolis.jl:54-67,149).

Parity status: pinned by normalisation (sum_x q(x) = 1 over all 2^D
configurations for D <= 16), the autoregressive property (l_d does not
depend on x_>=d) and finite differences of the log-likelihood.
"""
from __future__ import annotations

import numpy as np

FIELDS = ("W1", "b1", "W2", "b2", "W3", "b3")


def degrees(D: int, H: int):
    m0 = np.arange(1, D + 1)
    mh = (np.arange(H) % max(D - 1, 1)) + 1
    return m0, mh, mh


def masks(D: int, H: int):
    m0, m1, m2 = degrees(D, H)
    M1 = (m1[:, None] >= m0[None, :]).astype(np.float64)       # H x D
    M2 = (m2[:, None] >= m1[None, :]).astype(np.float64)       # H x H
    M3 = (m0[:, None] > m2[None, :]).astype(np.float64)        # D x H
    return M1, M2, M3


def shapes(D: int, H: int):
    return {"W1": (H, D), "b1": (H,), "W2": (H, H), "b2": (H,), "W3": (D, H), "b3": (D,)}


def n_params(D: int, H: int) -> int:
    return sum(int(np.prod(s)) for s in shapes(D, H).values())


def unflatten(theta, D, H):
    out, o = {}, 0
    for k in FIELDS:
        s = shapes(D, H)[k]
        n = int(np.prod(s))
        out[k] = theta[o:o + n].reshape(s)          # row-major: W[row, col] at row * ncols + col
        o += n
    return out


def init(D, H, rng, scale=None):
    """Masked Gaussian initialisation (the masked entries are exactly zero)."""
    M1, M2, M3 = masks(D, H)
    th = {}
    th["W1"] = rng.standard_normal((H, D)) * (scale or 1.0 / np.sqrt(D)) * M1
    th["b1"] = np.zeros(H)
    th["W2"] = rng.standard_normal((H, H)) * (scale or 1.0 / np.sqrt(H)) * M2
    th["b2"] = np.zeros(H)
    th["W3"] = rng.standard_normal((D, H)) * (scale or 1.0 / np.sqrt(H)) * M3
    th["b3"] = np.zeros(D)
    return np.concatenate([th[k].ravel() for k in FIELDS])


def bits(keys, D):
    keys = np.asarray(keys, dtype=np.uint64)
    return ((keys[:, None] >> np.arange(D, dtype=np.uint64)[None, :]) & np.uint64(1)).astype(np.float64)


def forward(theta, X, D, H):
    p = unflatten(theta, D, H)
    M1, M2, M3 = masks(D, H)
    a1 = X @ (p["W1"] * M1).T + p["b1"]
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ (p["W2"] * M2).T + p["b2"]
    h2 = np.maximum(a2, 0.0)
    logits = h2 @ (p["W3"] * M3).T + p["b3"]
    return logits, (a1, h1, a2, h2)


def log_q(theta, keys, D, H):
    X = bits(keys, D)
    l, _ = forward(theta, X, D, H)
    return np.sum(X * l - np.logaddexp(0.0, l), axis=1)


def nll_and_grad(theta, keys, D, H):
    """Mean negative log-likelihood over the batch and its gradient (masked)."""
    X = bits(keys, D)
    B = X.shape[0]
    p = unflatten(theta, D, H)
    M1, M2, M3 = masks(D, H)
    l, (a1, h1, a2, h2) = forward(theta, X, D, H)
    nll = -np.mean(np.sum(X * l - np.logaddexp(0.0, l), axis=1))
    dl = (1.0 / (1.0 + np.exp(-l)) - X) / B
    g = {}
    g["W3"] = (dl.T @ h2) * M3
    g["b3"] = dl.sum(0)
    dh2 = dl @ (p["W3"] * M3)
    da2 = dh2 * (a2 > 0)
    g["W2"] = (da2.T @ h1) * M2
    g["b2"] = da2.sum(0)
    dh1 = da2 @ (p["W2"] * M2)
    da1 = dh1 * (a1 > 0)
    g["W1"] = (da1.T @ X) * M1
    g["b1"] = da1.sum(0)
    return nll, np.concatenate([g[k].ravel() for k in FIELDS])


def adam_step(theta, grad, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """Adam (Kingma & Ba, 2015), bias-corrected, t counted from 1."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v


def sample(theta, n, D, H, rng):
    """Ancestral sampling, d = 0 .. D-1; returns packed keys and log q."""
    X = np.zeros((n, D))
    for d in range(D):
        l, _ = forward(theta, X, D, H)
        pr = 1.0 / (1.0 + np.exp(-l[:, d]))
        X[:, d] = (rng.random(n) < pr).astype(np.float64)
    keys = (X.astype(np.uint64) << np.arange(D, dtype=np.uint64)[None, :]).sum(axis=1).astype(np.uint64)
    return keys, log_q(theta, keys, D, H)
