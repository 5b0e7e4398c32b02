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


# ---- ket RBM and the energy path -------------------------------------------

def kron_ising_hamiltonian(N, g, V, periodic):
    """H = sum g/2 sx_i + sum V/4 sz_i sz_j by explicit Kronecker products in
    the packed-key basis (site 1 = least significant bit, bit 0 <-> v = -1 <->
    the reference's local index 1, on which sigmaz = diag(1, -1))."""
    sx = np.array([[0, 1], [1, 0]], dtype=complex)
    sz = np.diag([1.0, -1.0]).astype(complex)

    def op(site_mats):
        out = np.eye(1, dtype=complex)
        for i in range(N, 0, -1):          # most significant bit first
            out = np.kron(out, site_mats.get(i, np.eye(2)))
        return out
    H = np.zeros((2 ** N, 2 ** N), dtype=complex)
    for i in range(1, N + 1):
        H += (g / 2) * op({i: sx})
    bonds = [(i, i + 1) for i in range(1, N)] + ([(1, N)] if periodic and N > 2 else [])
    for i, j in bonds:
        H += (V / 4) * op({i: sz, j: sz})
    return H


def test_ket_hamiltonian_matches_kronecker_form():
    for N, periodic in ((3, True), (4, False)):
        model = ops_ref.ising_chain_model(N, V=2.0, g=1.0, periodic=periodic)
        H = rbm_ref.dense_hamiltonian(model)
        assert np.allclose(H, kron_ising_hamiltonian(N, 1.0, 2.0, periodic), atol=1e-14)


def test_ket_rbm_O_and_local_energy():
    N, nh = 3, 4
    rng = np.random.default_rng(8)
    n = rbm_ref.ket_n_params(N, nh)
    p = 0.2 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    keys = np.arange(2 ** N, dtype=np.uint64)
    lp, O = rbm_ref.ket_logpsi_grad(p, keys, N, nh)
    h = 1e-6
    for k in range(n):
        e = np.zeros(n, dtype=complex)
        e[k] = h
        fd = (rbm_ref.ket_logpsi(p + e, keys, N, nh) - rbm_ref.ket_logpsi(p - e, keys, N, nh)) / (2 * h)
        assert np.max(np.abs(fd - O[:, k])) < 1e-7
    model = ops_ref.ising_chain_model(N)
    psi = np.exp(lp)
    E = rbm_ref.local_energy(model, p, keys, N, nh)
    Hk = kron_ising_hamiltonian(N, 1.0, 2.0, True)
    assert np.max(np.abs(E - (Hk @ psi) / psi)) < 1e-12


def test_energy_gradient_is_dE_dp():
    """With exact |psi|^2 weights, sum p E_loc conj(Oc) = dE/d conj(p) =
    (dE/dRe p + i dE/dIm p) / 2 for E = <psi|H|psi>/<psi|psi>."""
    N, nh = 3, 2
    rng = np.random.default_rng(9)
    n = rbm_ref.ket_n_params(N, nh)
    p = 0.2 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    keys = np.arange(2 ** N, dtype=np.uint64)
    model = ops_ref.ising_chain_model(N)
    H = kron_ising_hamiltonian(N, 1.0, 2.0, True)

    def energy(q):
        psi = np.exp(rbm_ref.ket_logpsi(q, keys, N, nh))
        return (np.conj(psi) @ H @ psi).real / np.sum(np.abs(psi) ** 2)
    lp, O = rbm_ref.ket_logpsi_grad(p, keys, N, nh)
    w = np.abs(np.exp(lp)) ** 2
    w /= w.sum()
    E = rbm_ref.local_energy(model, p, keys, N, nh)
    Oc = O - w @ O
    F = (w * E) @ np.conj(Oc)
    h = 1e-6
    for k in range(n):
        e = np.zeros(n, dtype=complex)
        e[k] = h
        dre = (energy(p + e) - energy(p - e)) / (2 * h)
        dim = (energy(p + 1j * e) - energy(p - 1j * e)) / (2 * h)
        assert abs(F[k] - 0.5 * (dre + 1j * dim)) < 1e-6


def test_energy_stats_known_answer():
    """Complex statistics: the mean, and the variances as mean |x - m|^2."""
    B, Lc = 3, 4
    rng = np.random.default_rng(10)
    E = rng.standard_normal(B * Lc) + 1j * rng.standard_normal(B * Lc)
    O = np.zeros((B * Lc, 1), dtype=complex)
    _, _, _, st = rbm_ref.energy_grad_step(B, Lc, O, E)
    v = E.reshape(Lc, B).T
    assert abs(st[0] + 1j * st[1] - E.mean()) < 1e-14
    chain_var = np.array([np.var(v[j], ddof=1) for j in range(B)])   # numpy var of complex = mean |x-m|^2
    assert abs(st[3] - chain_var.mean()) < 1e-14
    assert abs(st[2] - np.sqrt(np.var(v.mean(1), ddof=1) / B)) < 1e-14
