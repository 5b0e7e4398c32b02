"""Exact full sums on the device (nqh_exact, the batched ExactSampler with
the accumulator on top) against the oracle's full sums, the dense
Hamiltonian, and a large-N Monte Carlo estimate."""
import numpy as np
import pytest

from _helpers import ising, params, rel_err
from nqhip import Context, _lib, operators, problems
from oracle import nqref, ops_ref, rbm_ref

pytestmark = pytest.mark.gpu


def test_exact_ndm_matches_oracle(gpu_ok, ref):
    N, nh, na = 3, 3, 3
    flat, tab, _ = ising(N)
    p = params(ref.ndm_np(N, nh, na), seed=1, sigma=0.2)
    ctx = Context(0, 1)
    ctx.set_problem(flat)
    ctx.set_ndm(N, nh, na)
    ctx.set_params(p)
    C, Lm, logZ, logp = ctx.exact(want_logp=True)
    ex = nqref.exact_quantities(ref, tab, nh, na, p)
    assert abs(C - ex["C"]) < 1e-10 * ex["C"]
    assert np.max(np.abs(np.exp(logp) - ex["prob"])) < 1e-12
    assert abs(Lm - np.sum(ex["prob"] * ex["Lloc"])) < 1e-10 * max(1.0, abs(Lm))
    ctx.close()


def test_exact_rbmsplit_matches_oracle(gpu_ok):
    N, nh = 3, 6
    rng = np.random.default_rng(2)
    n = rbm_ref.n_params(N, nh)
    p = 0.2 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    flat, _, _ = ising(N)
    ctx = Context(0, 1)
    ctx.set_problem(flat)
    ctx.set_rbmsplit(N, nh)
    ctx.set_params(p.view(np.float64))
    C, _, _, logp = ctx.exact(want_logp=True)
    ex = rbm_ref.exact(ops_ref.ising_chain_model(N), p, N, nh)
    assert abs(C - ex["C"]) < 1e-10 * ex["C"]
    assert np.max(np.abs(np.exp(logp) - ex["prob"])) < 1e-12
    ctx.close()


def test_exact_energy_matches_dense_hamiltonian(gpu_ok):
    N, nh = 8, 8
    rng = np.random.default_rng(3)
    n = rbm_ref.ket_n_params(N, nh)
    p = 0.2 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    _, H, _, _ = problems.dissipative_ising_chain(N, V=2.0, g=1.0)
    ctx = Context(0, 1)
    ctx.set_problem(operators.flatten(H))
    ctx.set_rbm(N, nh)
    ctx.set_params(p.view(np.float64))
    _, E, _ = ctx.exact()
    psi = np.exp(rbm_ref.ket_logpsi(p, np.arange(2 ** N, dtype=np.uint64), N, nh))
    Hd = rbm_ref.dense_hamiltonian(ops_ref.ising_chain_model(N))
    e_ref = (np.conj(psi) @ Hd @ psi) / np.sum(np.abs(psi) ** 2)
    assert abs(E - e_ref) < 1e-10 * abs(e_ref)
    ctx.close()


def test_exact_large_n_agrees_with_sampling(gpu_ok):
    """N = 10 (2^20 doubled configurations): the exact <L^dag L> and the
    Monte Carlo mean of |L_loc|^2 over 4096 x 16 samples agree within 5
    standard errors."""
    N, nh, na = 10, 10, 10
    flat, _, _ = ising(N)
    ctx = Context(0, 5)
    ctx.set_problem(flat)
    ctx.set_ndm(N, nh, na)
    ctx.set_params(params(ctx.np_, seed=4, sigma=0.05))
    C, _, _ = ctx.exact()
    ctx.set_sampler(4096, 16, 2 * N + 1, 50, 0)
    st = ctx.sample_grad()
    assert abs(st[0] - C) < 5 * st[1] + 1e-12, (st[0], st[1], C)
    ctx.close()
