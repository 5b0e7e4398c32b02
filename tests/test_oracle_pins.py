"""Pins of the oracle itself (CPU): every function of oracle/ is checked here
against something that does not restate its formula -- a published known
answer, a closed form of different construction, the textbook Lindblad
superoperator, finite differences, brute force over small spaces, or a
library routine.  The reference ships no golden vectors for this path."""
import math

import numpy as np
import pytest
import scipy.linalg as sl

from _helpers import ising, params, random_keys, rel_err, xyz
from oracle import nqref, ops_ref


# ---------------------------------------------------------------- RNG
def test_philox_known_answers(ref):
    # Random123 kat_vectors for philox4x32-10
    assert ref.philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert ref.philox([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert ref.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


# ---------------------------------------------------------------- NDM
@pytest.mark.parametrize("N,nh,na", [(2, 2, 2), (4, 8, 4), (4, 4, 12)])
def test_ndm_logpsi_vs_numpy_closed_form(ref, N, nh, na):
    p = params(ref.ndm_np(N, nh, na), seed=N + nh + na, sigma=0.4)
    keys = nqref.all_keys(N)
    s = ((keys[None, :] >> np.arange(N, dtype=np.uint64)[:, None]) & np.uint64(1)).astype(float) * 2 - 1
    e = ((keys[None, :] >> (np.arange(N, dtype=np.uint64) + np.uint64(N))[:, None]) & np.uint64(1)).astype(float) * 2 - 1
    cf = nqref.ndm_closed_form(N, nh, na, p, s, e)
    assert rel_err(ref.logpsi(N, nh, na, p, keys), cf) < 1e-13


def test_ndm_parameter_count_matches_reference_formula(ref):
    for N, ah, aa in [(6, 1, 1), (12, 2, 2), (16, 2, 2)]:
        nh, na = ah * N, aa * N
        assert ref.ndm_np(N, nh, na) == 2 * N + 2 * nh + na + 2 * nh * N + 2 * na * N
    assert ref.ndm_np(16, 32, 32) == 2176 and ref.ndm_np(6, 6, 6) == 174 and ref.ndm_np(12, 24, 24) == 1248


def test_ndm_gradient_vs_finite_differences(ref):
    N, nh, na = 3, 4, 3
    p = params(ref.ndm_np(N, nh, na), seed=5, sigma=0.3)
    keys = nqref.all_keys(N)
    _, O = ref.logpsi_grad(N, nh, na, p, keys)
    h = 1e-6
    for k in range(len(p)):
        dp = np.zeros_like(p)
        dp[k] = h
        fd = (ref.logpsi(N, nh, na, p + dp, keys) - ref.logpsi(N, nh, na, p - dp, keys)) / (2 * h)
        assert np.max(np.abs(fd - O[:, k])) < 1e-8


def test_ndm_hermiticity_and_block_structure(ref):
    N, nh, na = 4, 4, 4
    p = params(ref.ndm_np(N, nh, na), seed=9, sigma=0.3)
    keys = nqref.all_keys(N)
    mask = np.uint64(2 ** N - 1)
    swapped = (keys >> np.uint64(N)) | ((keys & mask) << np.uint64(N))
    lp, O = ref.logpsi_grad(N, nh, na, p, keys)
    # rho(e, s) = conj(rho(s, e)): the NDM is Hermitian by construction
    assert np.max(np.abs(ref.logpsi(N, nh, na, p, swapped) - np.conj(lp))) < 1e-13
    off, _ = nqref.ndm_layout(N, nh, na)
    lam = np.concatenate([np.arange(off[f][0], off[f][0] + off[f][1]) for f in ("b_la", "h_la", "w_la")])
    mu = np.concatenate([np.arange(off[f][0], off[f][0] + off[f][1]) for f in ("b_mu", "h_mu", "w_mu")])
    assert np.all(O[:, lam].imag == 0) and np.all(O[:, mu].real == 0)
    diag = keys[(keys & mask) == (keys >> np.uint64(N))]
    assert np.max(np.abs(ref.logpsi(N, nh, na, p, diag).imag)) < 1e-14


# ---------------------------------------------------------------- Liouvillian
@pytest.mark.parametrize("n,periodic,fields_first", [(3, True, False), (4, True, True), (4, False, False)])
def test_connections_equal_kronecker_liouvillian(n, periodic, fields_first):
    m = ops_ref.ising_chain_model(n, V=2.0, g=0.7, periodic=periodic, fields_first=fields_first)
    assert np.max(np.abs(ops_ref.dense_from_connections(m) - ops_ref.dense_liouvillian_kron(m))) < 1e-14


def test_xyz_connections_equal_kronecker_liouvillian():
    m = ops_ref.xyz_model(2, 2)
    assert np.max(np.abs(ops_ref.dense_from_connections(m) - ops_ref.dense_liouvillian_kron(m))) < 1e-14


def test_kronecker_liouvillian_is_textbook_lindblad():
    """The Kronecker form applied to vec(rho) equals -i[H, rho] + sum L rho L^+
    - 1/2 {L^+ L, rho} computed with plain matrix products (reference
    convention: Hermitian H, symmetric HnH)."""
    m = ops_ref.ising_chain_model(3, V=2.0, g=0.7)
    d = 2 ** m.n
    H = sum(ops_ref.full_op(m.n, s, M) for s, M in m.h_terms)
    Ls = [ops_ref.full_op(m.n, s, L) for s, L in m.jumps]
    rng = np.random.default_rng(0)
    rho = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    lind = -1j * (H @ rho - rho @ H)
    for L in Ls:
        lind += L @ rho @ L.conj().T - 0.5 * (L.conj().T @ L @ rho + rho @ L.conj().T @ L)
    vec = ops_ref.dense_liouvillian_kron(m) @ rho.reshape(-1, order="F")
    assert np.max(np.abs(vec - lind.reshape(-1, order="F"))) < 1e-12


def test_steady_state_trace_preserving():
    """sum over the diagonal of L rho vanishes for any rho (trace preservation)."""
    m = ops_ref.ising_chain_model(3)
    d = 2 ** m.n
    Lv = ops_ref.dense_liouvillian_kron(m)
    diag_rows = [i + d * i for i in range(d)]
    assert np.max(np.abs(Lv[diag_rows].sum(axis=0))) < 1e-12


# ---------------------------------------------------------------- estimator
# ---------------------------------------------------------------- NDM with af_logcosh
def test_logcosh_activation_known_values():
    """The oracle's logcosh restates activation.jl:5-10 (a branch at |x| = 12,
    a complex form through tanh); numpy's log(cosh z) is a different
    construction of the same function.  Known values: log cosh 1 =
    0.43378083048302..., log cosh(i pi/4) = log cos(pi/4) = -log(2)/2, and
    the |x| > 12 branch |x| - log 2, which differs from log cosh x by
    log1p(e^-2|x|) (< 3.8e-11), is taken there and only there."""
    assert abs(math.log(math.cosh(1.0)) - 0.4337808304830271) < 1e-15
    r = nqref.Ref(act=1)
    N, nh, na = 1, 1, 1
    off, npar = nqref.ndm_layout(N, nh, na)
    # a single Pi unit with argument d_la + i u_mu (s - e)/2: s = +1, e = -1 gives z = d + i u
    for d, u in [(0.0, math.pi / 4), (1.0, 0.3), (13.5, -0.7), (-14.0, 2.0)]:
        p = np.zeros(npar)
        p[off["d_la"][0]] = d
        p[off["u_mu"][0]] = u
        key = np.array([1], dtype=np.uint64)          # s = +1, e = -1
        base = 2 * math.log(math.cosh(0.0))            # the lambda / mu units at zero angles
        z = complex(d, u)
        want = complex(np.log(np.cosh(z))) + base
        got = r.logpsi(N, nh, na, p, key)[0]
        # beyond |x| = 12 the reference's branch drops log1p(e^-2|x|) < 3.8e-11
        branch = math.log1p(math.exp(-2 * abs(d))) if abs(d) > 12 else 0.0
        assert abs(got - want) < 1e-13 * max(1.0, abs(want)) + 1.01 * branch
        assert branch == 0.0 or abs(got - want) > 0.5 * branch
    assert abs(r.logpsi(1, 1, 1, np.eye(1, npar, off["u_mu"][0]).ravel() * (math.pi / 4),
                        np.array([1], dtype=np.uint64))[0] - (-0.5 * math.log(2.0))) < 1e-15


@pytest.mark.parametrize("N,nh,na,sigma", [(2, 2, 2, 0.4), (4, 8, 4, 0.4), (3, 3, 6, 6.0)])
def test_ndm_logcosh_vs_numpy_closed_form(N, nh, na, sigma):
    """sigma = 6 puts many angles beyond the |x| = 12 branch point."""
    r = nqref.Ref(act=1)
    p = params(r.ndm_np(N, nh, na), seed=N + nh + na + 1, sigma=sigma)
    keys = nqref.all_keys(N)
    s = ((keys[None, :] >> np.arange(N, dtype=np.uint64)[:, None]) & np.uint64(1)).astype(float) * 2 - 1
    e = ((keys[None, :] >> (np.arange(N, dtype=np.uint64) + np.uint64(N))[:, None]) & np.uint64(1)).astype(float) * 2 - 1
    cf = nqref.ndm_closed_form(N, nh, na, p, s, e, act=1)
    # the imaginary part is defined mod 2 pi per Pi unit: compare exp(i Im)
    got = r.logpsi(N, nh, na, p, keys)
    # the reference's branch beyond |x| = 12 drops log1p(e^-2|x|) < 3.8e-11 per unit
    units = 2 * 2 * nh + na
    assert np.max(np.abs(got.real - cf.real)) < 1e-13 * np.max(np.abs(cf.real)) + units * 3.8e-11
    # (Gamma_mu, a sum of real f, is part of Im log rho: the same branch bound)
    assert np.max(np.abs(np.exp(1j * got.imag) - np.exp(1j * cf.imag))) < 1e-12 + units * 3.8e-11
    # and it differs from softplus
    assert rel_err(nqref.Ref(act=0).logpsi(N, nh, na, p, keys).real, cf.real) > 1e-3


def test_ndm_logcosh_gradient_vs_finite_differences():
    """O = d log rho / dp with f' = tanh (fwd_der(logL2) = dlogL2, activation.jl:26)."""
    r = nqref.Ref(act=1)
    N, nh, na = 3, 4, 3
    p = params(r.ndm_np(N, nh, na), seed=6, sigma=0.5)
    keys = nqref.all_keys(N)
    _, O = r.logpsi_grad(N, nh, na, p, keys)
    h = 1e-6
    for k in range(len(p)):
        dp = np.zeros_like(p)
        dp[k] = h
        fd = (r.logpsi(N, nh, na, p + dp, keys) - r.logpsi(N, nh, na, p - dp, keys)) / (2 * h)
        assert np.max(np.abs(fd - O[:, k])) < 1e-8


def test_ndm_logcosh_hermitian_and_sampled(ref):
    """rho(e, s) = conj(rho(s, e)) with logcosh too, and the oracle's Metropolis
    samples |rho|^2 of NDM(2, 1, 1, af_logcosh) (TV against the exact pdf)."""
    r = nqref.Ref(act=1)
    N, nh, na = 2, 2, 2
    p = params(r.ndm_np(N, nh, na), seed=41, sigma=0.8)
    keys = nqref.all_keys(N)
    mask = np.uint64(2 ** N - 1)
    swapped = (keys >> np.uint64(N)) | ((keys & mask) << np.uint64(N))
    lp = r.logpsi(N, nh, na, p, keys)
    assert np.max(np.abs(np.exp(r.logpsi(N, nh, na, p, swapped)) - np.conj(np.exp(lp)))) < 1e-13
    w = np.exp(2 * lp.real)
    exact = w / w.sum()
    out, _ = r.metropolis(N, nh, na, p, 64, 400, 5, 20, seed=98)
    hist = np.bincount(out.astype(np.int64), minlength=16) / len(out)
    assert 0.5 * np.abs(hist - exact).sum() < 0.02


@pytest.mark.parametrize("model", ["ising3", "xyz22"])
def test_local_estimator_is_L_rho_over_rho(ref, model):
    if model == "ising3":
        _, tab, m = ising(3)
        N, nh, na = 3, 3, 6
    else:
        _, tab, m = xyz(2, 2)
        N, nh, na = 4, 4, 4
    p = params(ref.ndm_np(N, nh, na), seed=13, sigma=0.3)
    keys = nqref.all_keys(N)
    lp = ref.logpsi(N, nh, na, p, keys)
    rho = np.exp(lp)
    Lloc = ref.local_estimator(tab, nh, na, p, keys, grad=False)
    assert rel_err(Lloc, (ops_ref.dense_liouvillian_kron(m) @ rho) / rho) < 1e-12


def test_estimator_gradient_is_half_exact_loss_gradient(ref):
    """With every configuration weighted by p, 2 Re(grad C) equals the finite
    difference gradient of C = |L rho|^2 / |rho|^2."""
    N, nh, na = 3, 3, 3
    _, tab, m = ising(3)
    p = params(ref.ndm_np(N, nh, na), seed=17, sigma=0.2)
    keys = nqref.all_keys(N)
    Lv = ops_ref.dense_liouvillian_kron(m)

    def C(pp):
        r = np.exp(ref.logpsi(N, nh, na, pp, keys))
        return np.linalg.norm(Lv @ r) ** 2 / np.linalg.norm(r) ** 2

    ex = nqref.exact_quantities(ref, tab, nh, na, p)
    assert abs(ex["C"] - C(p)) < 1e-12 * C(p)
    h = 1e-6
    g = np.array([(C(p + h * e) - C(p - h * e)) / (2 * h) for e in np.eye(len(p))])
    assert rel_err(2 * ex["gradC"].real, g) < 1e-7


# ---------------------------------------------------------------- statistics / SR
def test_stat_analysis_known_answer(ref):
    vals = np.array([[1.0, 3.0], [5.0, 7.0]])     # chains x slices
    st = ref.stat_analysis(vals)
    assert np.allclose(st, [4.0, 2.0, 2.0, 0.7, math.sqrt(1.7)], rtol=0, atol=1e-15)


def test_stat_analysis_iid_limit(ref):
    vals = np.random.default_rng(1).normal(size=(200, 400))
    st = ref.stat_analysis(vals)
    assert abs(st[3]) < 0.2 and abs(st[4] - 1.0) < 0.01


def test_grad_step_covariance_and_force(ref):
    """S, F and Obar of the C oracle against numpy linear algebra on the same
    per-sample arrays (covariance identity S = Re(X X^H)/Ns with X centred)."""
    N, nh, na = 4, 4, 4
    _, tab, _ = ising(4)
    p = params(ref.ndm_np(N, nh, na), seed=23)
    B, Lc = 8, 6
    keys = random_keys(N, B * Lc, seed=2)
    _, O = ref.logpsi_grad(N, nh, na, p, keys)
    L, dL = ref.local_estimator(tab, nh, na, p, keys)
    S, F, Ob, st = ref.grad_step(B, Lc, O, L, dL)
    Ns = B * Lc
    X = O - O.mean(axis=0)
    assert rel_err(S, (X.T @ X.conj()).real / Ns) < 1e-13
    C = np.mean(np.abs(L) ** 2)
    assert abs(st[0] - C) < 1e-13 * C
    assert rel_err(F, ((np.conj(L) @ dL) / Ns - C * O.mean(axis=0)).real) < 1e-12
    assert np.allclose(S, S.T) and np.min(np.linalg.eigvalsh(S)) > -1e-12


def test_sr_solve_is_shifted_cholesky():
    rng = np.random.default_rng(3)
    X = rng.normal(size=(50, 30))
    S = X.T @ X / 50
    F = rng.normal(size=30)
    d = nqref.sr_solve(S, F, 0.01)
    assert np.max(np.abs((S + 0.01 * np.eye(30)) @ d - F)) < 1e-11
    with pytest.raises(np.linalg.LinAlgError):
        nqref.sr_solve(-np.eye(3), np.ones(3), 0.0)


# ---------------------------------------------------------------- sampler
def test_metropolis_stationary_distribution(ref):
    """TV distance of the sampled doubled configurations to the exact |rho|^2
    distribution over all 16 states of N = 2 (test_samplers.jl:26-31 pattern)."""
    N, nh, na = 2, 2, 2
    p = params(ref.ndm_np(N, nh, na), seed=29, sigma=0.8)
    keys = nqref.all_keys(N)
    lp = ref.logpsi(N, nh, na, p, keys)
    w = np.exp(2 * lp.real)
    exact = w / w.sum()
    out, acc = ref.metropolis(N, nh, na, p, 64, 400, 5, 20, seed=99)
    hist = np.bincount(out.astype(np.int64), minlength=16) / len(out)
    assert 0.5 * np.abs(hist - exact).sum() < 0.02
    assert 0 < acc < 64 * 420 * 5


def test_metropolis_compat_stale_last_slice(ref):
    """Defect D1 reproduced on request: the last slice is never sampled and
    keeps the all -1 fill (key 0); the other slices are unchanged."""
    N, nh, na = 3, 3, 3
    p = params(ref.ndm_np(N, nh, na), seed=31)
    a, _ = ref.metropolis(N, nh, na, p, 8, 5, 3, 4, seed=5, compat=True)
    b, _ = ref.metropolis(N, nh, na, p, 8, 5, 3, 4, seed=5, compat=False)
    assert np.all(a[4 * 8:] == 0)
    assert np.array_equal(a[:4 * 8], b[:4 * 8])


def test_metropolis_detailed_balance_two_state_chain(ref):
    """With N = 1 (4 doubled states) the acceptance rate matches the exact
    Metropolis transition matrix of a single-flip uniform proposal."""
    N, nh, na = 1, 1, 1
    p = params(ref.ndm_np(N, nh, na), seed=37, sigma=0.9)
    keys = nqref.all_keys(N)
    w = np.exp(2 * ref.logpsi(N, nh, na, p, keys).real)
    pi = w / w.sum()
    acc_exact = 0.0
    for k in range(4):
        for site in range(2):
            k2 = k ^ (1 << site)
            acc_exact += pi[k] * 0.5 * min(1.0, w[k2] / w[k])
    chains, Lc, passes, burn = 256, 200, 1, 50
    _, acc = ref.metropolis(N, nh, na, p, chains, Lc, passes, burn, seed=7)
    rate = acc / (chains * (Lc + burn) * passes)
    assert abs(rate - acc_exact) < 0.01
