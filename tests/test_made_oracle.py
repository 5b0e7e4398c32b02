"""Pins of the MADE proposal oracle (CPU): normalisation, the autoregressive
property, finite-difference gradients, and Adam's first step closed form."""
import numpy as np

from oracle import made_ref as mr


def test_made_is_normalised_over_all_configurations():
    D, H = 8, 16
    th = mr.init(D, H, np.random.default_rng(0), scale=0.7)
    th[-D:] = np.random.default_rng(1).normal(size=D)       # nonzero output biases
    keys = np.arange(2 ** D, dtype=np.uint64)
    q = np.exp(mr.log_q(th, keys, D, H))
    assert abs(q.sum() - 1.0) < 1e-12


def test_made_is_autoregressive():
    D, H = 10, 24
    th = mr.init(D, H, np.random.default_rng(2), scale=0.8)
    rng = np.random.default_rng(3)
    X = rng.integers(0, 2, size=(64, D)).astype(float)
    l, _ = mr.forward(th, X, D, H)
    for d in range(D):
        Y = X.copy()
        Y[:, d:] = rng.integers(0, 2, size=(64, D - d))     # change x_>=d only
        l2, _ = mr.forward(th, Y, D, H)
        assert np.max(np.abs(l2[:, d] - l[:, d])) < 1e-13


def test_made_gradient_finite_differences():
    D, H = 6, 10
    th = mr.init(D, H, np.random.default_rng(4), scale=0.9)
    # random biases keep every pre-activation away from the ReLU kink at 0
    p = mr.unflatten(th, D, H)
    rb = np.random.default_rng(8)
    p["b1"][:] = rb.normal(0, 0.5, H)
    p["b2"][:] = rb.normal(0, 0.5, H)
    p["b3"][:] = rb.normal(0, 0.5, D)
    keys = np.random.default_rng(5).integers(0, 2 ** D, size=40, dtype=np.uint64)
    _, g = mr.nll_and_grad(th, keys, D, H)
    M = np.concatenate([m.ravel() if m.ndim else m for m in
                        [mr.masks(D, H)[0], np.ones(H), mr.masks(D, H)[1], np.ones(H), mr.masks(D, H)[2], np.ones(D)]])
    h = 1e-6
    for k in np.flatnonzero(M):
        e = np.zeros_like(th)
        e[k] = h
        fd = (mr.nll_and_grad(th + e, keys, D, H)[0] - mr.nll_and_grad(th - e, keys, D, H)[0]) / (2 * h)
        assert abs(fd - g[k]) < 1e-7
    assert np.all(g[M == 0] == 0)


def test_adam_first_step_is_signed_lr():
    th = np.zeros(5)
    g = np.array([1.0, -2.0, 3.0, -4.0, 0.5])
    th2, m, v = mr.adam_step(th, g, np.zeros(5), np.zeros(5), 1, lr=1e-3)
    assert np.allclose(th2, -1e-3 * np.sign(g), atol=1e-9)


def test_made_sampling_matches_density():
    D, H = 4, 8
    th = mr.init(D, H, np.random.default_rng(6), scale=1.2)
    keys, lq = mr.sample(th, 40000, D, H, np.random.default_rng(7))
    assert np.allclose(lq, mr.log_q(th, keys, D, H))
    q = np.exp(mr.log_q(th, np.arange(16, dtype=np.uint64), D, H))
    hist = np.bincount(keys.astype(np.int64), minlength=16) / len(keys)
    assert 0.5 * np.abs(hist - q).sum() < 0.02
