"""Learned proposal on the device: the MADE kernels against the FP64 oracle
(oracle/made_ref.py), the exact normalisation and autoregressive invariants,
Adam against its closed form, and stationarity of the mixed local + global
Metropolis chains against the exact distribution."""
import os

import numpy as np
import pytest

from _helpers import ising, params, rel_err
from nqhip import Context, _lib
from oracle import made_ref, nqref

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
H = 128


def golden():
    with np.load(os.path.join(HERE, "golden", "made_d8.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def made_ctx(N, seed=1, mix=0.0, **kw):
    ctx = Context(0, seed)
    nh = na = N
    flat, _, _ = ising(N)
    ctx.set_problem(flat)
    ctx.set_ndm(N, nh, na)
    ctx.set_params(params(ctx.np_, seed=3, sigma=0.1))
    ctx.made_setup(mix=mix, **kw)
    return ctx


def test_made_logq_normalised_and_matches_oracle(gpu_ok):
    """sum_x q(x) = 1 over all 2^8 configurations (exact for any logits when
    the masks make l_d depend on x_<d only), and log q within the BF16
    operand error of the FP64 oracle (0.05 absolute)."""
    g = golden()
    ctx = made_ctx(4, buffer_cap=0)
    ctx.made_set_params(g["theta"])
    lq = ctx.made_logq(g["keys"])
    assert abs(np.exp(lq.astype(np.float64)).sum() - 1.0) < 1e-4
    assert np.max(np.abs(lq - g["logq"])) < 5e-2
    ctx.close()


def test_made_gradient_matches_oracle(gpu_ok):
    """Mean-NLL gradient of the golden batch against the FP64 oracle: within
    3e-2 of max |grad| (BF16 operands, FP32 accumulation); masked entries
    exactly zero; NLL within 1e-2 relative."""
    g = golden()
    ctx = made_ctx(4, buffer_cap=0)
    ctx.made_set_params(g["theta"])
    nll, gr = ctx.made_grad(g["batch"])
    assert rel_err(gr, g["grad"]) < 3e-2
    assert abs(nll - float(g["nll"])) < 1e-2 * abs(float(g["nll"]))
    D = int(g["D"])
    M1, M2, M3 = made_ref.masks(D, H)
    mask = np.concatenate([M1.ravel(), np.ones(H), M2.ravel(), np.ones(H), M3.ravel(), np.ones(D)])
    assert np.all(gr[mask == 0] == 0.0)
    ctx.close()


def test_made_adam_step_closed_form(gpu_ok):
    """Two Adam steps on the device against made_ref.adam_step driven by the
    device's own gradients (FP32 tolerance)."""
    g = golden()
    ctx = made_ctx(4, buffer_cap=0)
    th = g["theta"].astype(np.float32)
    ctx.made_set_params(th)
    batch = g["batch"]
    m = np.zeros(len(th))
    v = np.zeros(len(th))
    ref = th.astype(np.float64)
    for t in (1, 2):
        _, gd = ctx.made_grad(batch)
        ctx.made_train(batch, steps=1, batch=len(batch), lr=1e-3)
        ref, m, v = made_ref.adam_step(ref, gd.astype(np.float64), m, v, t, lr=1e-3)
        got = ctx.made_get_params()
        assert np.max(np.abs(got - ref)) < 2e-6
    ctx.close()


def test_made_sampling_matches_logq(gpu_ok):
    """Ancestral samples: their log q equals the evaluation kernel's bit for
    bit (the exact Hastings ratio needs that), and their histogram is within
    total variation 0.03 of q itself (2^17 samples over 2^8 states)."""
    g = golden()
    ctx = made_ctx(4, buffer_cap=0)
    ctx.made_set_params(g["theta"])
    keys, lq = ctx.made_sample(1 << 17, it=5, tag=1)
    assert np.array_equal(ctx.made_logq(keys), lq)
    q = np.exp(ctx.made_logq(g["keys"]).astype(np.float64))
    hist = np.bincount(keys.astype(np.int64), minlength=256) / len(keys)
    assert 0.5 * np.abs(hist - q).sum() < 0.03
    k2, _ = ctx.made_sample(1 << 17, it=5, tag=1)
    assert np.array_equal(keys, k2)
    ctx.close()


def test_made_training_reduces_nll(gpu_ok):
    """Training on samples of a product Bernoulli distribution drives the
    NLL from its initial value to within 0.2 nat (minibatch noise) of that distribution's
    entropy."""
    D = 8
    rng = np.random.default_rng(4)
    pb = rng.uniform(0.1, 0.9, D)
    X = rng.random((8192, D)) < pb
    keys = (X.astype(np.uint64) << np.arange(D, dtype=np.uint64)).sum(1).astype(np.uint64)
    ent = -np.sum(pb * np.log(pb) + (1 - pb) * np.log(1 - pb))
    ctx = made_ctx(4, buffer_cap=0)
    nll0, _ = ctx.made_grad(keys[:1024])
    nll = ctx.made_train(keys, steps=400, batch=512, lr=3e-3)
    assert nll < nll0
    assert nll < ent + 0.2
    ctx.close()


def exact_pdf(ref, N, nh, na, p):
    lp = ref.logpsi(N, nh, na, p, nqref.all_keys(N))
    w = np.exp(2 * (lp.real - lp.real.max()))
    return w / w.sum()


@pytest.mark.parametrize("mix", [0.0, 1.0])
def test_mixed_sampler_is_stationary(gpu_ok, ref, mix):
    """Persistent chains on the 2N = 6 doubled space, with (mix = 1) and
    without (mix = 0) a global move before every stored slice: the pooled
    histogram of 8 calls x 512 chains x 32 slices is within total variation
    0.03 of the exact |rho|^2 distribution."""
    N, nh, na = 3, 3, 3
    flat, _, _ = ising(N)
    p = params(ref.ndm_np(N, nh, na), seed=21, sigma=0.3)
    ctx = Context(0, 9)
    ctx.set_problem(flat)
    ctx.set_ndm(N, nh, na)
    ctx.set_params(p)
    ctx.set_sampler(512, 32, 2 * N + 1, 20, _lib.SAMPLER_PERSISTENT)
    ctx.made_setup(mix=mix, retrain_every=0, train_steps=0, batch=64, buffer_cap=0)
    pex = exact_pdf(ref, N, nh, na, p)
    # a proposal fitted to samples of the exact distribution
    rng = np.random.default_rng(2)
    fit = rng.choice(len(pex), size=16384, p=pex).astype(np.uint64)
    ctx.made_train(fit, steps=300, batch=512, lr=3e-3)
    hist = np.zeros(len(pex))
    acc_g = 0
    for it in range(8):
        keys, _ = ctx.sample()
        hist += np.bincount(keys.astype(np.int64), minlength=len(pex))
        prop, acc, _ = ctx.made_stats()
        assert prop == (512 * 32 if mix == 1.0 else 0)
        acc_g += acc
    hist /= hist.sum()
    assert 0.5 * np.abs(hist - pex).sum() < 0.03
    if mix == 1.0:
        assert acc_g > 0.5 * 8 * 512 * 32      # the fitted proposal is accepted most of the time
    ctx.close()


def test_mixed_sampler_graph_replay_matches_eager(gpu_ok, ref):
    """SR iterations with the learned proposal (global moves every other
    slice, retraining every 2 calls from the sample buffer): graph replays
    give the same samples, statistics and parameters as eager execution."""
    N, nh, na = 4, 4, 4
    flat, _, _ = ising(N)
    p = params(ref.ndm_np(N, nh, na), seed=5, sigma=0.05)
    out = []
    for graph in (False, True):
        ctx = Context(0, 13)
        ctx.set_problem(flat)
        ctx.set_ndm(N, nh, na)
        ctx.set_params(p)
        ctx.set_sampler(64, 8, 9, 10, _lib.SAMPLER_PERSISTENT)
        ctx.set_iteration(0)
        ctx.made_setup(mix=0.5, retrain_every=2, train_steps=5, batch=128, lr=1e-3, buffer_cap=1024)
        ctx.set_graph(graph)
        rec = []
        for _ in range(5):
            st = ctx.sr_iteration(0.01, 0.01)
            rec.append((st, ctx.get_samples(), ctx.made_stats()))
        out.append((rec, ctx.get_params(), ctx.made_get_params()))
        ctx.close()
    (ra, pa, ma), (rb, pb, mb) = out
    for (sa, ka, xa), (sb, kb, xb) in zip(ra, rb):
        assert rel_err(sb, sa) < 1e-12
        assert np.array_equal(ka, kb)
        assert xa == xb
    assert rel_err(pb, pa) < 1e-12
    assert np.array_equal(ma, mb)


def tau_int(series):
    """Integrated autocorrelation time (Sokal window, c = 5) of per-chain
    series (chains x T), in stored samples; 1/2 for independent samples."""
    x = series - series.mean(axis=1, keepdims=True)
    var = np.mean(x * x)
    tau = 0.5
    for t in range(1, x.shape[1]):
        tau += np.mean(x[:, :-t] * x[:, t:]) / var
        if t >= 5 * tau:
            break
    return max(tau, 0.5)


def test_learned_proposal_shortens_autocorrelation(gpu_ok, ref):
    """What the trained network gains: with one local proposal per stored
    sample, chains on a peaked 2N = 8 distribution are strongly correlated;
    a global move from a proposal fitted to the distribution before every
    stored sample cuts the integrated autocorrelation time of log p by more
    than half, with the same stationary distribution (TV of 65536 pooled
    samples over the 256 configurations below 0.04)."""
    N, nh, na = 4, 4, 4
    flat, _, _ = ising(N)
    p = params(ref.ndm_np(N, nh, na), seed=23, sigma=0.6)
    pex = exact_pdf(ref, N, nh, na, p)
    rng = np.random.default_rng(3)
    fit = rng.choice(len(pex), size=32768, p=pex).astype(np.uint64)
    taus, tvs = [], []
    for mix in (0.0, 1.0):
        ctx = Context(0, 17)
        ctx.set_problem(flat)
        ctx.set_ndm(N, nh, na)
        ctx.set_params(p)
        ctx.set_sampler(1024, 64, 1, 50, 0)
        ctx.made_setup(mix=mix, retrain_every=0, train_steps=0, batch=512, buffer_cap=0)
        ctx.made_train(fit, steps=400, batch=512, lr=3e-3)
        keys, _ = ctx.sample()
        lp = 2.0 * ctx.logpsi(keys).real.reshape(64, 1024).T
        taus.append(tau_int(lp))
        h = np.bincount(keys.astype(np.int64), minlength=len(pex)) / len(keys)
        tvs.append(0.5 * np.abs(h - pex).sum())
        ctx.close()
    assert taus[1] < 0.5 * taus[0], taus
    assert tvs[1] < 0.04, tvs
