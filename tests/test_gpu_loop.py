"""Device SR loop against the oracle's composition of the reference's steps
(sample! -> precondition! -> update!), graph replays against eager
execution, and the stated errors of the reduced-precision forms."""
import numpy as np
import pytest

from _helpers import ising, params, random_keys, rel_err, xyz
from nqhip import Context, _lib
from oracle import nqref

pytestmark = pytest.mark.gpu


def make(flat, N, nh, na, p, chains, Lc, passes, burn, seed, flags=0):
    ctx = Context(0, seed)
    ctx.set_problem(flat)
    ctx.set_ndm(N, nh, na)
    ctx.set_params(p)
    ctx.set_sampler(chains, Lc, passes, burn, flags)
    ctx.set_iteration(0)
    return ctx


def oracle_iteration(ref, tab, N, nh, na, p, chains, Lc, passes, burn, seed, it, eps, lr):
    keys, _ = ref.metropolis(N, nh, na, p, chains, Lc, passes, burn, seed, it=it)
    _, O = ref.logpsi_grad(N, nh, na, p, keys)
    L, dL = ref.local_estimator(tab, nh, na, p, keys)
    S, F, _, st = ref.grad_step(chains, Lc, O, L, dL)
    delta = nqref.sr_solve(S, F, eps)
    return p - lr * delta, st


@pytest.mark.parametrize("N,nh,na", [(4, 4, 4), (6, 6, 6)])
def test_sr_iterations_match_oracle(gpu_ok, ref, N, nh, na):
    flat, tab, _ = ising(N)
    p = params(ref.ndm_np(N, nh, na), seed=3, sigma=0.1)
    chains, Lc, passes, burn, seed, eps, lr = 16, 8, 2 * N + 1, 6, 77, 0.01, 0.01
    ctx = make(flat, N, nh, na, p, chains, Lc, passes, burn, seed)
    pr = p.copy()
    for it in range(3):
        st = ctx.sr_iteration(eps, lr)
        pr, st_r = oracle_iteration(ref, tab, N, nh, na, pr, chains, Lc, passes, burn, seed, it, eps, lr)
        assert rel_err(st, st_r) < 1e-10
        assert rel_err(ctx.get_params(), pr) < 1e-10
    ctx.close()


@pytest.mark.parametrize("flags", [0, _lib.SAMPLER_PERSISTENT])
def test_graph_replay_matches_eager(gpu_ok, ref, flags):
    N, nh, na = 6, 6, 6
    flat, _, _ = ising(N)
    p = params(ref.ndm_np(N, nh, na), seed=5, sigma=0.05)
    args = (flat, N, nh, na, p, 32, 8, 13, 10, 11, flags)
    a = make(*args)
    b = make(*args)
    b.set_graph(True)
    for _ in range(4):
        sa = a.sr_iteration(0.01, 0.01)
        sb = b.sr_iteration(0.01, 0.01)
        assert rel_err(sb, sa) < 1e-12
        assert np.array_equal(a.get_samples(), b.get_samples())
    assert rel_err(b.get_params(), a.get_params()) < 1e-12
    a.close()
    b.close()


@pytest.mark.parametrize("prec", [(_lib.PREC_FP64, _lib.PREC_FP64), (_lib.PREC_FP32, _lib.PREC_BF16)])
def test_forces_beside_s_build_match_in_order(gpu_ok, ref, prec):
    """4096 samples: the forces launch on the side stream beside the S build
    (graph replays and eager) gives what the same launches give in order (the
    profiled path): no race between the two streams."""
    N, nh, na = 6, 6, 6
    flat, _, _ = ising(N)
    p = params(ref.ndm_np(N, nh, na), seed=8, sigma=0.05)
    args = (flat, N, nh, na, p, 512, 8, 13, 10, 21)
    a, b, c = make(*args), make(*args), make(*args)
    for x in (a, b, c):
        x.set_precision(*prec)
    a.set_profiling(True)
    c.set_graph(True)
    for _ in range(3):
        sa = a.sr_iteration(0.01, 0.01)
        sb = b.sr_iteration(0.01, 0.01)
        sc = c.sr_iteration(0.01, 0.01)
        assert rel_err(sb, sa) < 1e-12
        assert rel_err(sc, sa) < 1e-12
    assert rel_err(b.get_params(), a.get_params()) < 1e-12
    assert rel_err(c.get_params(), a.get_params()) < 1e-12
    for x in (a, b, c):
        x.close()


def test_fp32_network_errors_stated(gpu_ok, ref):
    """FP32 network kernels against the FP64 oracle at the 4x4 XYZ shapes:
    log rho within 1e-5 absolute, O within 1e-5 relative, L_loc within 1e-4
    relative (the FP32 form's stated accuracy)."""
    N, nh, na = 16, 32, 32
    flat, tab, _ = xyz(4, 4)
    p = params(ref.ndm_np(N, nh, na), seed=7, sigma=0.01)
    keys = random_keys(N, 512, seed=12)
    ctx = Context(0, 1)
    ctx.set_problem(flat)
    ctx.set_ndm(N, nh, na)
    ctx.set_params(p)
    ctx.set_precision(_lib.PREC_FP32, _lib.PREC_FP32)
    lp, O = ctx.logpsi_grad(keys)
    lp_r, O_r = ref.logpsi_grad(N, nh, na, p, keys)
    assert np.max(np.abs(lp - lp_r)) < 1e-5
    assert rel_err(O, O_r) < 1e-5
    L = ctx.local_estimator(keys)
    L_r = ref.local_estimator(tab, nh, na, p, keys, grad=False)
    assert rel_err(L, L_r) < 1e-4
    ctx.close()


@pytest.mark.parametrize("sprec,tol", [(_lib.PREC_FP32, 1e-4), (_lib.PREC_BF16, 2e-2)])
def test_reduced_precision_s_build_error(gpu_ok, ref, sprec, tol):
    """S from FP32 / BF16 MFMA on centred O against the FP64 oracle
    (max abs error relative to max |S|)."""
    N, nh, na = 6, 6, 6
    flat, tab, _ = ising(N)
    p = params(ref.ndm_np(N, nh, na), seed=9, sigma=0.1)
    chains, Lc = 64, 16
    keys = random_keys(N, chains * Lc, seed=13)
    ctx = Context(0, 1)
    ctx.set_problem(flat)
    ctx.set_ndm(N, nh, na)
    ctx.set_params(p)
    ctx.set_precision(_lib.PREC_FP32, sprec)
    _, S, F, _ = ctx.grad_from_samples(keys, chains, Lc)
    _, O = ref.logpsi_grad(N, nh, na, p, keys)
    L, dL = ref.local_estimator(tab, nh, na, p, keys)
    S_r, F_r, _, _ = ref.grad_step(chains, Lc, O, L, dL)
    assert rel_err(S, S_r) < tol
    assert rel_err(F, F_r) < 1e-4
    ctx.close()


def test_mixed_form_reaches_the_fp64_loss(gpu_ok):
    """The chosen form (FP32 network, BF16 S build, persistent chains) and the
    plain FP64 form (restart + burn-in every iteration) on the N = 6 open
    dissipative Ising chain, NDM alpha = 1, 512 chains x 32, 300 SR
    iterations (shift 0.01, lr 0.01) from the same parameters: the exact
    <L^dag L> of the final parameters (full sum over the 4096 doubled
    configurations) agrees within 1 %, after falling by a factor > 20."""
    from nqhip import operators, problems
    N = 6
    _, _, _, L = problems.dissipative_ising_chain(N, V=2.0, g=1.0, gamma=1.0, periodic=False)
    flat = operators.flatten(L)
    p0 = np.random.default_rng(1234).standard_normal(5 * N + 4 * N * N) * 0.01
    final = []
    for prec, flags in (((_lib.PREC_FP64, _lib.PREC_FP64), 0), ((_lib.PREC_FP32, _lib.PREC_BF16),
                                                                _lib.SAMPLER_PERSISTENT)):
        ctx = Context(0, 4321)
        ctx.set_problem(flat)
        ctx.set_ndm(N, N, N)
        ctx.set_params(p0)
        ctx.set_precision(*prec)
        ctx.set_sampler(512, 32, 2 * N + 1, 100, flags)
        ctx.set_graph(True)
        start = ctx.exact()[0]
        for _ in range(300):
            ctx.sr_iteration(0.01, 0.01)
        final.append(ctx.exact()[0])
        assert final[-1] < start / 20, (start, final[-1])
        ctx.close()
    assert abs(final[1] - final[0]) < 0.01 * final[0], final
