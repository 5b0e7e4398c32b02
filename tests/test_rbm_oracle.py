"""Pins of the RBMSplit oracle (oracle/rbm_ref.py) by what the mathematics
fixes: the network formula evaluated configuration by configuration,
holomorphic finite differences of log rho, L_loc = (L vec rho)/rho with the
dense Kronecker Liouvillian, the exact gradient of C = ||L rho||^2/||rho||^2
against finite differences, S Hermitian positive semi-definite, and the
complex Cholesky solve."""
import numpy as np
import pytest

from oracle import nqref, ops_ref, rbm_ref


def rand_params(N, nh, seed, scale=0.2):
    rng = np.random.default_rng(seed)
    n = rbm_ref.n_params(N, nh)
    return scale * (rng.standard_normal(n) + 1j * rng.standard_normal(n))


def direct_logpsi(p, key, N, nh):
    """log rho of one configuration straight from RBMSplit.jl:45-50."""
    ar, ac, b, Wr, Wc = rbm_ref.unpack(p, N, nh)
    sr = np.array([1.0 if (int(key) >> i) & 1 else -1.0 for i in range(N)])
    sc = np.array([1.0 if (int(key) >> (N + i)) & 1 else -1.0 for i in range(N)])
    out = 0j
    for i in range(N):
        out += ar[i] * sr[i] + ac[i] * sc[i]
    for h in range(nh):
        th = b[h] + sum(Wr[h, i] * sr[i] + Wc[h, i] * sc[i] for i in range(N))
        out += np.log(1 + np.exp(th))
    return out


def test_logpsi_matches_direct_formula():
    N, nh = 3, 5
    p = rand_params(N, nh, 1)
    keys = np.arange(4 ** N, dtype=np.uint64)
    lp = rbm_ref.logpsi(p, keys, N, nh)
    ref = np.array([direct_logpsi(p, k, N, nh) for k in keys])
    # imaginary parts may differ by 2 pi k (branch of log per unit)
    d = lp - ref
    assert np.max(np.abs(d.real)) < 1e-12
    assert np.max(np.abs(np.angle(np.exp(1j * d.imag)))) < 1e-12


def test_O_is_holomorphic_derivative():
    N, nh = 2, 3
    p = rand_params(N, nh, 2)
    keys = np.arange(4 ** N, dtype=np.uint64)
    _, O = rbm_ref.logpsi_grad(p, keys, N, nh)
    h = 1e-6
    for k in range(len(p)):
        e = np.zeros(len(p), dtype=complex)
        e[k] = h
        dre = (rbm_ref.logpsi(p + e, keys, N, nh) - rbm_ref.logpsi(p - e, keys, N, nh)) / (2 * h)
        dim = (rbm_ref.logpsi(p + 1j * e, keys, N, nh) - rbm_ref.logpsi(p - 1j * e, keys, N, nh)) / (2 * h)
        assert np.max(np.abs(dre - O[:, k])) < 1e-7
        assert np.max(np.abs(dim - 1j * O[:, k])) < 1e-7     # Cauchy-Riemann


@pytest.mark.parametrize("N,periodic", [(2, False), (3, True)])
def test_local_estimator_is_L_rho_over_rho(N, periodic):
    nh = N
    model = ops_ref.ising_chain_model(N, periodic=periodic)
    p = rand_params(N, nh, 3)
    keys = np.arange(4 ** N, dtype=np.uint64)
    rho = np.exp(rbm_ref.logpsi(p, keys, N, nh))
    Lrho = ops_ref.dense_liouvillian_kron(model) @ rho
    L, dL = rbm_ref.local_estimator(model, p, keys, N, nh)
    assert np.max(np.abs(L - Lrho / rho)) < 1e-10 * np.max(np.abs(L))
    # dL_loc(s) = sum_y L_sy rho_y O(y) / rho_s
    _, O = rbm_ref.logpsi_grad(p, keys, N, nh)
    Ld = ops_ref.dense_liouvillian_kron(model)
    ref = (Ld * rho[None, :]) @ O / rho[:, None]
    assert np.max(np.abs(dL - ref)) < 1e-10 * np.max(np.abs(ref))


def exact_C(model, p, N, nh):
    keys = np.arange(4 ** N, dtype=np.uint64)
    rho = np.exp(rbm_ref.logpsi(p, keys, N, nh))
    Lr = ops_ref.dense_liouvillian_kron(model) @ rho
    return np.sum(np.abs(Lr) ** 2) / np.sum(np.abs(rho) ** 2)


def test_exact_force_is_dC_dp():
    """With the exact |rho|^2 weights, F = sum p conj(L) dL - C sum p O equals
    dC/dp = (dC/dRe p - i dC/dIm p) / 2 (finite differences of C)."""
    N, nh = 2, 2
    model = ops_ref.ising_chain_model(N, periodic=False)
    p = rand_params(N, nh, 4)
    ex = rbm_ref.exact(model, p, N, nh)
    w = ex["prob"]
    F = (w * np.conj(ex["Lloc"])) @ ex["dLloc"] - ex["C"] * (w @ ex["O"])
    assert abs(ex["C"] - exact_C(model, p, N, nh)) < 1e-10 * ex["C"]
    h = 1e-6
    for k in range(len(p)):
        e = np.zeros(len(p), dtype=complex)
        e[k] = h
        dre = (exact_C(model, p + e, N, nh) - exact_C(model, p - e, N, nh)) / (2 * h)
        dim = (exact_C(model, p + 1j * e, N, nh) - exact_C(model, p - 1j * e, N, nh)) / (2 * h)
        assert abs(F[k] - 0.5 * (dre - 1j * dim)) < 1e-6 * max(1.0, abs(F[k]))


def test_grad_step_S_hermitian_psd_and_solve():
    N, nh = 3, 3
    model = ops_ref.ising_chain_model(N)
    p = rand_params(N, nh, 5)
    rng = np.random.default_rng(6)
    B, Lc = 8, 4
    keys = rng.integers(0, 4 ** N, size=B * Lc, dtype=np.uint64)
    _, O = rbm_ref.logpsi_grad(p, keys, N, nh)
    L, dL = rbm_ref.local_estimator(model, p, keys, N, nh)
    S, F, Ob, _ = rbm_ref.grad_step(B, Lc, O, L, dL)
    assert np.max(np.abs(S - S.conj().T)) < 1e-14
    assert np.min(np.linalg.eigvalsh(S)) > -1e-12
    Oc = O - O.mean(0)
    assert np.max(np.abs(S - np.einsum("sj,sk->jk", np.conj(Oc), Oc) / len(keys))) < 1e-13
    d = rbm_ref.sr_solve(S, F, 0.01)
    assert np.max(np.abs((S + 0.01 * np.eye(len(F))) @ d - F)) < 1e-10
    # the real embedding used by the device solves the same system
    A, Bm = S.real, S.imag
    M = np.block([[A, -Bm], [Bm, A]]) + 0.01 * np.eye(2 * len(F))
    xy = np.linalg.solve(M, np.concatenate([F.real, F.imag]))
    assert np.max(np.abs(xy[:len(F)] + 1j * xy[len(F):] - d)) < 1e-10


def test_metropolis_oracle_is_stationary():
    """The oracle's chains, pooled, reproduce the exact |rho|^2 distribution
    (2N = 4, 16 states)."""
    N, nh = 2, 2
    p = rand_params(N, nh, 7, scale=0.4)
    ref = nqref.Ref()
    keys, _, _ = rbm_ref.metropolis(p, N, nh, chains=64, chain_len=40, passes=5, burn=10, seed=11, it=0,
                                    philox=ref.philox)
    lp = rbm_ref.logpsi(p, np.arange(16, dtype=np.uint64), N, nh)
    pex = np.exp(2 * lp.real)
    pex /= pex.sum()
    hist = np.bincount(keys.astype(np.int64), minlength=16) / len(keys)
    assert 0.5 * np.abs(hist - pex).sum() < 0.05
