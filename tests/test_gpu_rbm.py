"""RBMSplit on the device against the oracle (oracle/rbm_ref.py): log rho, O,
L_loc, dL_loc, the complex gradient and SR system, the solve, the sampler's
trajectories, whole SR iterations, graph replay, and the FP32 form's stated
errors."""
import numpy as np
import pytest

from _helpers import ising, rel_err
from nqhip import Context, _lib
from oracle import nqref, ops_ref, rbm_ref

pytestmark = pytest.mark.gpu


def cparams(N, nh, seed, scale=0.1):
    rng = np.random.default_rng(seed)
    n = rbm_ref.n_params(N, nh)
    return scale * (rng.standard_normal(n) + 1j * rng.standard_normal(n))


def rbm_ctx(N, nh, p, seed=1, periodic=True, prec=_lib.PREC_FP64, sprec=_lib.PREC_FP64):
    flat, _, _ = ising(N, periodic=periodic)
    ctx = Context(0, seed)
    ctx.set_problem(flat)
    ctx.set_rbmsplit(N, nh)
    ctx.set_precision(prec, sprec)
    ctx.set_params(p.view(np.float64))
    return ctx


@pytest.mark.parametrize("N,nh", [(3, 3), (4, 8), (6, 12)])
def test_rbm_logpsi_grad(gpu_ok, N, nh):
    p = cparams(N, nh, 1)
    ctx = rbm_ctx(N, nh, p)
    rng = np.random.default_rng(2)
    keys = rng.integers(0, 4 ** N, size=200, dtype=np.uint64)
    lp, O = ctx.logpsi_grad(keys)
    lp_r, O_r = rbm_ref.logpsi_grad(p, keys, N, nh)
    # imaginary parts compared modulo 2 pi (branch of the complex logarithm)
    assert np.max(np.abs(lp.real - lp_r.real)) < 1e-10
    assert np.max(np.abs(np.angle(np.exp(1j * (lp.imag - lp_r.imag))))) < 1e-10
    assert rel_err(O, O_r) < 1e-10
    ctx.close()


@pytest.mark.parametrize("N,nh,periodic", [(3, 3, True), (4, 8, True), (6, 6, False)])
def test_rbm_local_estimator(gpu_ok, N, nh, periodic):
    p = cparams(N, nh, 3)
    model = ops_ref.ising_chain_model(N, periodic=periodic)
    ctx = rbm_ctx(N, nh, p, periodic=periodic)
    rng = np.random.default_rng(4)
    keys = rng.integers(0, 4 ** N, size=96, dtype=np.uint64)
    L, dL = ctx.local_estimator(keys, grad=True)
    L_r, dL_r = rbm_ref.local_estimator(model, p, keys, N, nh)
    assert rel_err(L, L_r) < 1e-10
    assert rel_err(dL, dL_r) < 1e-10
    ctx.close()


def test_rbm_grad_and_solve(gpu_ok, ref):
    """Statistics, F, O-bar and the embedded S against the oracle's complex
    quantities (1e-10), and the shifted solve against scipy's complex
    Cholesky (1e-9)."""
    N, nh = 4, 8
    p = cparams(N, nh, 5)
    model = ops_ref.ising_chain_model(N)
    ctx = rbm_ctx(N, nh, p)
    B, Lc = 32, 4
    keys = np.random.default_rng(6).integers(0, 4 ** N, size=B * Lc, dtype=np.uint64)
    st, M, F, Ob = ctx.grad_from_samples(keys, B, Lc)
    _, O = rbm_ref.logpsi_grad(p, keys, N, nh)
    L, dL = rbm_ref.local_estimator(model, p, keys, N, nh)
    S_r, F_r, Ob_r, st_r = rbm_ref.grad_step(B, Lc, O, L, dL, stat_analysis=ref.stat_analysis)
    n = len(F_r)
    assert rel_err(st, st_r) < 1e-10
    assert rel_err(F, F_r) < 1e-10
    assert rel_err(Ob, Ob_r) < 1e-10
    Mr = np.block([[S_r.real, -S_r.imag], [S_r.imag, S_r.real]])
    assert rel_err(M, Mr) < 1e-10
    d = ctx.precondition(0.01)
    assert rel_err(d, rbm_ref.sr_solve(S_r, F_r, 0.01)) < 1e-9
    assert n == ctx.nd
    ctx.close()


@pytest.mark.parametrize("it", [0, 3])
def test_rbm_sampler_trajectory_bit_exact(gpu_ok, ref, it):
    N, nh = 3, 3
    p = cparams(N, nh, 7, scale=0.3)
    ctx = rbm_ctx(N, nh, p, seed=99)
    chains, Lc, passes, burn = 8, 4, 7, 3
    ctx.set_sampler(chains, Lc, passes, burn, 0)
    ctx.set_iteration(it)
    keys, acc = ctx.sample()
    keys_r, _, acc_r = rbm_ref.metropolis(p, N, nh, chains, Lc, passes, burn, 99, it, ref.philox)
    assert np.array_equal(keys, keys_r)
    assert acc == acc_r
    ctx.close()


def test_rbm_sampler_stationary(gpu_ok):
    """Pooled samples of persistent chains against the exact |rho|^2 (TV < 0.03)."""
    N, nh = 3, 3
    p = cparams(N, nh, 8, scale=0.3)
    ctx = rbm_ctx(N, nh, p, seed=5)
    ctx.set_sampler(512, 32, 2 * N + 1, 20, _lib.SAMPLER_PERSISTENT)
    hist = np.zeros(4 ** N)
    for _ in range(8):
        keys, _ = ctx.sample()
        hist += np.bincount(keys.astype(np.int64), minlength=4 ** N)
    hist /= hist.sum()
    lp = rbm_ref.logpsi(p, np.arange(4 ** N, dtype=np.uint64), N, nh)
    pex = np.exp(2 * (lp.real - lp.real.max()))
    pex /= pex.sum()
    assert 0.5 * np.abs(hist - pex).sum() < 0.03
    ctx.close()


def test_rbm_sr_iterations_match_oracle(gpu_ok, ref):
    N, nh = 4, 4
    p = cparams(N, nh, 9)
    model = ops_ref.ising_chain_model(N)
    ctx = rbm_ctx(N, nh, p, seed=77)
    chains, Lc, passes, burn, eps, lr = 16, 8, 9, 6, 0.01, 0.01
    ctx.set_sampler(chains, Lc, passes, burn, 0)
    ctx.set_iteration(0)
    pr = p.copy()
    for it in range(3):
        st = ctx.sr_iteration(eps, lr)
        keys, _, _ = rbm_ref.metropolis(pr, N, nh, chains, Lc, passes, burn, 77, it, ref.philox)
        _, O = rbm_ref.logpsi_grad(pr, keys, N, nh)
        L, dL = rbm_ref.local_estimator(model, pr, keys, N, nh)
        S, F, _, st_r = rbm_ref.grad_step(chains, Lc, O, L, dL, stat_analysis=ref.stat_analysis)
        pr = pr - lr * rbm_ref.sr_solve(S, F, eps)
        assert rel_err(st, st_r) < 1e-10
        assert rel_err(ctx.get_params().view(np.complex128), pr) < 1e-10
    ctx.close()


def test_rbm_graph_replay_matches_eager(gpu_ok):
    N, nh = 4, 8
    p = cparams(N, nh, 10, scale=0.05)
    runs = []
    for graph in (False, True):
        ctx = rbm_ctx(N, nh, p, seed=11)
        ctx.set_sampler(32, 8, 9, 10, _lib.SAMPLER_PERSISTENT)
        ctx.set_iteration(0)
        ctx.set_graph(graph)
        rec = [(ctx.sr_iteration(0.01, 0.01), ctx.get_samples()) for _ in range(4)]
        runs.append((rec, ctx.get_params()))
        ctx.close()
    (ra, pa), (rb, pb) = runs
    for (sa, ka), (sb, kb) in zip(ra, rb):
        assert rel_err(sb, sa) < 1e-12
        assert np.array_equal(ka, kb)
    assert rel_err(pb, pa) < 1e-12


def test_rbm_fp32_and_bf16_errors(gpu_ok, ref):
    """FP32 network (the reference's default ComplexF32 precision) within 1e-5
    (log rho real part, absolute) / 1e-4 (L_loc, relative), BF16 S build within
    2e-2 of max |S|."""
    N, nh = 6, 12
    p = cparams(N, nh, 12)
    model = ops_ref.ising_chain_model(N)
    ctx = rbm_ctx(N, nh, p, prec=_lib.PREC_FP32, sprec=_lib.PREC_BF16)
    B, Lc = 64, 8
    keys = np.random.default_rng(13).integers(0, 4 ** N, size=B * Lc, dtype=np.uint64)
    lp = ctx.logpsi(keys)
    lp_r = rbm_ref.logpsi(p, keys, N, nh)
    assert np.max(np.abs(lp.real - lp_r.real)) < 1e-5
    L = ctx.local_estimator(keys)
    L_r, dL_r = rbm_ref.local_estimator(model, p, keys, N, nh)
    assert rel_err(L, L_r) < 1e-4
    _, M, F, _ = ctx.grad_from_samples(keys, B, Lc)
    _, O = rbm_ref.logpsi_grad(p, keys, N, nh)
    S_r, F_r, _, _ = rbm_ref.grad_step(B, Lc, O, L_r, dL_r)
    Mr = np.block([[S_r.real, -S_r.imag], [S_r.imag, S_r.real]])
    assert rel_err(M, Mr) < 2e-2
    assert rel_err(F, F_r) < 1e-4
    ctx.close()
