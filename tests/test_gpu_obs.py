"""Observables on the device (nqh_observable): the diagonal chain's
trajectory against the oracle bit for bit, the local values against the
oracle, and the mean against Tr(O rho)/Tr(rho)."""
import numpy as np
import pytest

from _helpers import ising, params, rel_err
from nqhip import Context, hilbert, operators
from oracle import obs_ref, ops_ref, rbm_ref

pytestmark = pytest.mark.gpu


def site_sum_op(N, which):
    h = hilbert.HomogeneousSpin(N)
    f = {"x": operators.sigmax, "y": operators.sigmay, "z": operators.sigmaz}[which]
    op = operators.LocalOperator(h)
    for i in range(1, N + 1):
        op = op + (1.0 / N) * f(h, i)
    return operators.flatten(op)


def site_sum_model(N, which):
    M = {"x": ops_ref.SX, "y": ops_ref.SY, "z": ops_ref.SZ}[which]
    return ops_ref.Model(N, [((i,), M / N) for i in range(1, N + 1)], [])


def ndm_ctx(N, p, seed=3):
    flat, _, _ = ising(N)
    ctx = Context(0, seed)
    ctx.set_problem(flat)
    ctx.set_ndm(N, N, N)
    ctx.set_params(p)
    return ctx


@pytest.mark.parametrize("which", ["x", "y", "z"])
def test_observable_ndm_against_oracle(gpu_ok, ref, which):
    N = 3
    p = params(ref.ndm_np(N, N, N), seed=5, sigma=0.3)
    ctx = ndm_ctx(N, p)
    logrho = lambda k: ref.logpsi(N, N, N, p, k)   # noqa: E731
    chains, Lc, passes, burn = 16, 4, 5, 3
    st = ctx.observable(site_sum_op(N, which), chains, Lc, passes, burn)
    keys, loc = ctx.observable_samples()
    keys_r, _ = obs_ref.diag_metropolis(logrho, N, chains, Lc, passes, burn, seed=3, it=0, philox=ref.philox)
    assert np.array_equal(keys, keys_r)
    sig = keys & np.uint64(2 ** N - 1)
    loc_r = obs_ref.local_values(logrho, obs_ref.ket_terms(site_sum_model(N, which)), sig, N)
    assert rel_err(loc, loc_r) < 1e-10
    assert rel_err(st, obs_ref.complex_stats(loc_r, chains, Lc)) < 1e-10
    ctx.close()


def test_observable_mean_matches_trace(gpu_ok, ref):
    """4096 x 32 samples: <Sx> within 5 standard errors of Tr(Sx rho)/Tr(rho)."""
    N = 4
    p = params(ref.ndm_np(N, N, N), seed=6, sigma=0.3)
    ctx = ndm_ctx(N, p)
    st = ctx.observable(site_sum_op(N, "x"), 4096, 32, 2 * N + 1, 20)
    ex = obs_ref.exact_expectation(lambda k: ref.logpsi(N, N, N, p, k), site_sum_model(N, "x"), N)
    assert abs(st[0] - ex.real) < 5 * st[1] + 1e-12
    assert abs(ex.imag) < 1e-12                    # Hermitian NDM and operator: real in expectation
    assert abs(st[5]) < 5 * st[1] + 1e-12          # the sample mean's imaginary part is noise
    ctx.close()


def test_observable_rbmsplit_local_values(gpu_ok, ref):
    N, nh = 3, 6
    rng = np.random.default_rng(7)
    n = rbm_ref.n_params(N, nh)
    p = 0.2 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    flat, _, _ = ising(N)
    ctx = Context(0, 4)
    ctx.set_problem(flat)
    ctx.set_rbmsplit(N, nh)
    ctx.set_params(p.view(np.float64))
    ctx.observable(site_sum_op(N, "x"), 16, 4, 5, 3)
    keys, loc = ctx.observable_samples()
    logrho = lambda k: rbm_ref.logpsi(p, k, N, nh)   # noqa: E731
    keys_r, _ = obs_ref.diag_metropolis(logrho, N, 16, 4, 5, 3, seed=4, it=0, philox=ref.philox)
    assert np.array_equal(keys, keys_r)
    sig = keys & np.uint64(2 ** N - 1)
    loc_r = obs_ref.local_values(logrho, obs_ref.ket_terms(site_sum_model(N, "x")), sig, N)
    assert rel_err(loc, loc_r) < 1e-10
    ctx.close()
