"""Hamiltonian ground-state path on the device (ket RBM, local energy,
energy gradient, complex SR) against the oracle (oracle/rbm_ref.py), and a
short optimisation that approaches the exact ground-state energy."""
import numpy as np
import pytest

from _helpers import rel_err
from nqhip import Context, _lib, operators, problems
from oracle import ops_ref, rbm_ref

pytestmark = pytest.mark.gpu


def cparams(N, nh, seed, scale=0.1):
    rng = np.random.default_rng(seed)
    n = rbm_ref.ket_n_params(N, nh)
    return scale * (rng.standard_normal(n) + 1j * rng.standard_normal(n))


def ket_ctx(N, nh, p, seed=1, periodic=True, g=1.0, V=2.0, prec=_lib.PREC_FP64, sprec=_lib.PREC_FP64):
    _, H, _, _ = problems.dissipative_ising_chain(N, V=V, g=g, periodic=periodic)
    ctx = Context(0, seed)
    ctx.set_problem(operators.flatten(H))
    ctx.set_rbm(N, nh)
    ctx.set_precision(prec, sprec)
    ctx.set_params(p.view(np.float64))
    return ctx


def model(N, periodic=True, g=1.0, V=2.0):
    return ops_ref.ising_chain_model(N, V=V, g=g, periodic=periodic)


@pytest.mark.parametrize("N,nh", [(4, 4), (8, 16)])
def test_ket_logpsi_grad_and_energy(gpu_ok, N, nh):
    p = cparams(N, nh, 1)
    ctx = ket_ctx(N, nh, p)
    keys = np.random.default_rng(2).integers(0, 2 ** N, size=128, dtype=np.uint64)
    lp, O = ctx.logpsi_grad(keys)
    lp_r, O_r = rbm_ref.ket_logpsi_grad(p, keys, N, nh)
    assert np.max(np.abs(lp.real - lp_r.real)) < 1e-10
    assert np.max(np.abs(np.angle(np.exp(1j * (lp.imag - lp_r.imag))))) < 1e-10
    assert rel_err(O, O_r) < 1e-10
    E = ctx.local_estimator(keys)
    assert rel_err(E, rbm_ref.local_energy(model(N), p, keys, N, nh)) < 1e-10
    ctx.close()


def test_ket_gradient_and_solve(gpu_ok):
    N, nh = 6, 6
    p = cparams(N, nh, 3)
    ctx = ket_ctx(N, nh, p)
    B, Lc = 32, 4
    keys = np.random.default_rng(4).integers(0, 2 ** N, size=B * Lc, dtype=np.uint64)
    _, M, F, Ob = ctx.grad_from_samples(keys, B, Lc)
    st = ctx.get_stats(6)
    _, O = rbm_ref.ket_logpsi_grad(p, keys, N, nh)
    E = rbm_ref.local_energy(model(N), p, keys, N, nh)
    S_r, F_r, Ob_r, st_r = rbm_ref.energy_grad_step(B, Lc, O, E)
    # device order [Re mean, err, var, tau, R, Im mean]; oracle [Re, Im, err, var, tau, R]
    assert rel_err(st, np.array([st_r[0], st_r[2], st_r[3], st_r[4], st_r[5], st_r[1]])) < 1e-10
    assert rel_err(F, F_r) < 1e-10
    assert rel_err(Ob, Ob_r) < 1e-10
    assert rel_err(M, np.block([[S_r.real, -S_r.imag], [S_r.imag, S_r.real]])) < 1e-10
    d = ctx.precondition(0.01)
    assert rel_err(d, rbm_ref.sr_solve(S_r, F_r, 0.01)) < 1e-9
    ctx.close()


def test_ket_sampler_trajectory_bit_exact(gpu_ok, ref):
    N, nh = 5, 5
    p = cparams(N, nh, 5, scale=0.3)
    ctx = ket_ctx(N, nh, p, seed=42)
    chains, Lc, passes, burn = 8, 4, 5, 3
    ctx.set_sampler(chains, Lc, passes, burn, 0)
    ctx.set_iteration(2)
    keys, acc = ctx.sample()
    keys_r, acc_r = rbm_ref.ket_metropolis(p, N, nh, chains, Lc, passes, burn, 42, 2, ref.philox)
    assert np.array_equal(keys, keys_r)
    assert acc == acc_r
    ctx.close()


def test_ket_sr_iterations_match_oracle(gpu_ok, ref):
    N, nh = 4, 4
    p = cparams(N, nh, 6)
    ctx = ket_ctx(N, nh, p, seed=7)
    chains, Lc, passes, burn, eps, lr = 16, 8, 5, 4, 0.01, 0.05
    ctx.set_sampler(chains, Lc, passes, burn, 0)
    ctx.set_iteration(0)
    pr = p.copy()
    for it in range(3):
        ctx.sr_iteration(eps, lr)
        st = ctx.get_stats(6)
        keys, _ = rbm_ref.ket_metropolis(pr, N, nh, chains, Lc, passes, burn, 7, it, ref.philox)
        _, O = rbm_ref.ket_logpsi_grad(pr, keys, N, nh)
        E = rbm_ref.local_energy(model(N), pr, keys, N, nh)
        S, F, _, st_r = rbm_ref.energy_grad_step(chains, Lc, O, E)
        pr = pr - lr * rbm_ref.sr_solve(S, F, eps)
        assert abs(st[0] - st_r[0]) < 1e-10 * max(1.0, abs(st_r[0]))
        assert rel_err(ctx.get_params().view(np.complex128), pr) < 1e-10
    ctx.close()


def test_ket_optimisation_reaches_ground_state(gpu_ok):
    """Transverse-field Ising chain, N = 8: 150 SR iterations from a small
    random RBM bring <H> within 1 % of the exact ground-state energy."""
    N, nh = 8, 16
    H = rbm_ref.dense_hamiltonian(model(N, g=2.0, V=2.0))
    e0 = np.linalg.eigvalsh(H)[0]
    p = cparams(N, nh, 8, scale=0.01)
    ctx = ket_ctx(N, nh, p, seed=9, g=2.0, V=2.0)
    ctx.set_sampler(256, 8, 2 * N + 1, 20, _lib.SAMPLER_PERSISTENT)
    ctx.set_graph(True)
    es = []
    for _ in range(150):
        ctx.sr_iteration(0.01, 0.05)
        es.append(ctx.get_stats(1)[0])
    e = np.mean(es[-20:])
    assert abs(e - e0) < 0.01 * abs(e0), (e, e0)
    ctx.close()
