"""The benchmark configurations at their own shapes on the device, each
checked against the oracle or an exact full sum.

C1  N = 6 open dissipative Ising chain, NDM alpha = 1, 16 chains x 500
    samples, 100 burn-in calls, 13 passes, shift 0.01, lr 0.01, 200 SR
    iterations: the sampled loss against the exact <L^dag L> over the 4096
    doubled configurations, and the loss falls.  (The equal-loss comparison
    of the two forms is test_gpu_loop.py::test_mixed_form_reaches_the_fp64_loss.)
C2  N = 12 periodic, alpha = 2 (Np = 1248), 1024 chains x 16 samples, 50
    burn-in calls, shift 0.01, lr 0.01: the network, the estimator and the
    gradient statistics against the oracle at the full network shape, then
    300 SR iterations of the plain FP64 form and of the chosen form, the
    sampled loss against the exact <L^dag L> over all 2^24 doubled
    configurations.
C3  4x4 XYZ (Np = 2176), 4096 chains x 4: the FP32 network's errors against
    the FP64 oracle at parameters reached by training, not at the initial
    N(0, 0.01^2) draw.
C5  64 dissipative Ising chains, N ~ U{6..16}, g ~ U[0.5, 2] (seed 2024),
    alpha = 1, 256 chains x 8: the batch graph against every member run
    alone, and members against the oracle.
(C4, the learned proposal, is tests/test_gpu_made.py.)"""
import math

import numpy as np
import pytest

from _helpers import ising, params, random_keys, rel_err, xyz
from nqhip import Batch, Context, _lib, operators, problems
from oracle import nqref

pytestmark = pytest.mark.gpu


def sampled_vs_exact(ctx):
    """A fresh gradient pass at the current parameters: the sampled loss and
    its error (stats.jl:52-77) next to the exact full sum."""
    st = ctx.sample_grad()
    ex = ctx.exact()[0]
    return st[0], st[1], ex


def test_c1_loop_against_the_exact_sum(gpu_ok):
    N = 6
    _, _, _, L = problems.dissipative_ising_chain(N, V=2.0, g=1.0, gamma=1.0, periodic=False)
    ctx = Context(0, 4321)
    ctx.set_problem(operators.flatten(L))
    ctx.set_ndm(N, N, N)
    ctx.set_params(np.random.default_rng(1234).standard_normal(5 * N + 4 * N * N) * 0.01)
    ctx.set_precision(_lib.PREC_FP64, _lib.PREC_FP64)
    ctx.set_sampler(16, 500, 2 * N + 1, 100)
    ctx.set_graph(True)
    start = ctx.exact()[0]
    for _ in range(200):
        ctx.sr_iteration(0.01, 0.01)
    mean, err, ex = sampled_vs_exact(ctx)
    assert ex < start / 3, (start, ex)
    assert abs(mean - ex) <= 5 * err + 1e-3 * ex, (mean, err, ex)
    ctx.close()


def c2_ctx(prec, flags, seed=77):
    N = 12
    flat, _, _ = ising(N, periodic=True)
    ctx = Context(0, seed)
    ctx.set_problem(flat)
    ctx.set_ndm(N, 2 * N, 2 * N)
    ctx.set_params(params(ctx.np_, seed=1234, sigma=0.01))
    ctx.set_precision(*prec)
    ctx.set_sampler(1024, 16, 2 * N + 1, 50, flags)
    return ctx


def test_c2_network_and_gradient_against_the_oracle(gpu_ok, ref):
    """Np = 1248 (nh = na = 24), FP64: log rho and O of 512 configurations,
    L_loc and dL_loc of 128, and the statistics, F and S of 256 chains x 4
    sampled configurations, against the oracle within 1e-10."""
    N, nh, na = 12, 24, 24
    flat, tab, _ = ising(N, periodic=True)
    p = params(ref.ndm_np(N, nh, na), seed=5, sigma=0.05)
    ctx = Context(0, 9)
    ctx.set_problem(flat)
    ctx.set_ndm(N, nh, na)
    ctx.set_params(p)
    keys = random_keys(N, 512, seed=6)
    lp, O = ctx.logpsi_grad(keys)
    lp_r, O_r = ref.logpsi_grad(N, nh, na, p, keys)
    assert rel_err(lp, lp_r) < 1e-10 and rel_err(O, O_r) < 1e-10
    L, G = ctx.local_estimator(keys[:128], grad=True)
    L_r, G_r = ref.local_estimator(tab, nh, na, p, keys[:128])
    assert rel_err(L, L_r) < 1e-10 and rel_err(G, G_r) < 1e-10
    ctx.set_sampler(256, 4, 2 * N + 1, 50)
    ctx.set_iteration(3)
    skeys, _ = ctx.sample()
    st, S, F, _ = ctx.grad_from_samples(skeys, 256, 4)
    _, O2 = ref.logpsi_grad(N, nh, na, p, skeys)
    L2, dL2 = ref.local_estimator(tab, nh, na, p, skeys)
    S_r, F_r, _, st_r = ref.grad_step(256, 4, O2, L2, dL2)
    assert rel_err(st, st_r) < 1e-10 and rel_err(F, F_r) < 1e-10 and rel_err(S, S_r) < 1e-10
    ctx.close()


@pytest.mark.parametrize("form", ["fp64", "chosen"])
def test_c2_loop_against_the_exact_sum_over_2_24_configurations(gpu_ok, form):
    """300 SR iterations in the plain FP64 form (chains restarted and burned
    in every iteration, the reference's schedule) and in the chosen form
    (FP32 network, BF16 S build, persistent chains); then the sampled loss of
    16384 fresh samples against the exact sum over all 2^24 doubled
    configurations (nqh_exact, FP64 network for the chosen form's
    parameters too), within 5 of its standard errors."""
    prec, flags = (((_lib.PREC_FP64, _lib.PREC_FP64), 0) if form == "fp64"
                   else ((_lib.PREC_FP32, _lib.PREC_BF16), _lib.SAMPLER_PERSISTENT))
    ctx = c2_ctx(prec, flags)
    ctx.set_graph(True)
    ctx.set_precision(_lib.PREC_FP64, _lib.PREC_FP64)
    start = ctx.exact()[0]
    ctx.set_precision(*prec)
    for _ in range(300):
        ctx.sr_iteration(0.01, 0.01)
    ctx.set_precision(_lib.PREC_FP64, _lib.PREC_FP64)
    mean, err, ex = sampled_vs_exact(ctx)
    assert ex < 0.9 * start, (start, ex)
    assert abs(mean - ex) <= 5 * err, (mean, err, ex)
    ctx.close()


def test_c3_fp32_errors_at_trained_parameters(gpu_ok, ref):
    """The chosen form trained for 100 iterations on the 4x4 XYZ workload;
    at the parameters it reaches, the FP32 network against the FP64 oracle on
    512 of its own samples: log rho within 1e-4 absolute, O within 1e-5 and
    L_loc within 1e-4 relative to their largest entries."""
    N, nh, na = 16, 32, 32
    flat, tab, _ = xyz(4, 4)
    ctx = Context(0, 4321)
    ctx.set_problem(flat)
    ctx.set_ndm(N, nh, na)
    p0 = params(ctx.np_, seed=1234, sigma=0.01)
    ctx.set_params(p0)
    ctx.set_precision(_lib.PREC_FP32, _lib.PREC_BF16)
    ctx.set_sampler(4096, 4, 2 * N + 1, 50, _lib.SAMPLER_PERSISTENT)
    ctx.set_graph(True)
    first = ctx.sr_iteration(0.01, 0.005)[0]
    for _ in range(99):
        last = ctx.sr_iteration(0.01, 0.005)[0]
    assert last < first
    p = ctx.get_params()
    assert np.linalg.norm(p - p0) > 2 * np.linalg.norm(p0)   # far from the initial N(0, 0.01^2) draw
    keys = ctx.get_samples()[:512]
    lp, O = ctx.logpsi_grad(keys)
    lp_r, O_r = ref.logpsi_grad(N, nh, na, p, keys)
    assert np.max(np.abs(lp - lp_r)) < 1e-4
    assert rel_err(O, O_r) < 1e-5
    L = ctx.local_estimator(keys)
    assert rel_err(L, ref.local_estimator(tab, nh, na, p, keys, grad=False)) < 1e-4
    ctx.close()


C5 = dict(instances=64, nmin=6, nmax=16, gmin=0.5, gmax=2.0, seed=2024, chains=256, chain_length=8, burn=50)


def c5_members(prec):
    rng = np.random.default_rng(C5["seed"])
    ns = rng.integers(C5["nmin"], C5["nmax"] + 1, size=C5["instances"])
    gs = rng.uniform(C5["gmin"], C5["gmax"], size=C5["instances"])
    out = []
    for i, (N, g) in enumerate(zip(ns, gs)):
        N = int(N)
        _, _, _, L = problems.dissipative_ising_chain(N, V=2.0, g=float(g), gamma=1.0, periodic=True)
        ctx = Context(0, C5["seed"] + i)
        ctx.set_problem(operators.flatten(L))
        ctx.set_ndm(N, N, N)
        ctx.set_params(np.random.default_rng(C5["seed"] + 1000 + i).standard_normal(ctx.np_) * math.sqrt(1e-4))
        ctx.set_precision(*prec)
        ctx.set_sampler(C5["chains"], C5["chain_length"], 2 * N + 1, C5["burn"], _lib.SAMPLER_PERSISTENT)
        ctx.set_iteration(0)
        out.append((ctx, N, float(g)))
    return out


def test_c5_all_instances_batched_against_alone_and_the_oracle(gpu_ok, ref):
    """All 64 members in one batch graph (FP64, so the members can be held to
    the oracle) against the same 64 contexts run alone: bit-identical samples
    and statistics for three iterations; then the smallest and the largest
    member's gradient statistics, F and S on their own samples against the
    oracle within 1e-10."""
    prec = (_lib.PREC_FP64, _lib.PREC_FP64)
    inb, alone = c5_members(prec), c5_members(prec)
    batch = Batch([c for c, _, _ in inb])
    try:
        for it in range(3):
            st = batch.sr_iteration(0.01, 0.01)
            for i, (c, _, _) in enumerate(alone):
                sa = c.sr_iteration(0.01, 0.01)
                assert np.array_equal(st[i], sa), (it, i)
                assert np.array_equal(inb[i][0].get_samples(), c.get_samples()), (it, i)
        sizes = [N for _, N, _ in inb]
        for i in (int(np.argmin(sizes)), int(np.argmax(sizes))):
            c, N, g = inb[i]
            keys = c.get_samples()
            st, S, F, _ = c.grad_from_samples(keys, C5["chains"], C5["chain_length"])
            _, tab, _ = ising(N, periodic=True, g=g)
            p = c.get_params()
            _, O = ref.logpsi_grad(N, N, N, p, keys)
            Ll, dL = ref.local_estimator(tab, N, N, p, keys)
            S_r, F_r, _, st_r = ref.grad_step(C5["chains"], C5["chain_length"], O, Ll, dL)
            assert rel_err(st, st_r) < 1e-10 and rel_err(F, F_r) < 1e-10 and rel_err(S, S_r) < 1e-10, i
    finally:
        batch.close()
        for c, _, _ in inb + alone:
            c.close()
