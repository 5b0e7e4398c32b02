"""ORACLE (test infrastructure only) -- RBMSplit density-matrix network.

FP64/complex128 numpy restatement of the reference's RBMSplit on the
steady-state path:

  log rho(sr, sc) = ar . sr + ac . sc + sum_h logL(b_h + (Wr sr)_h + (Wc sc)_h)
  (src/Networks/MixedDensityMatrix/RBMSplit.jl:45-50,84-102; logL = log1p(exp)
  on complex arguments, src/Networks/activation.jl:18)

with complex parameters in functor order ar[N], ac[N], b[nh], Wr[nh x N],
Wc[nh x N] (column-major, RBMSplit.jl:3-10), nh = alpha N; the log-derivative
O (holomorphic: the network is analytic, RBMSplit.jl:45) is
(sr, sc, f'(theta), f'(theta) sr^T, f'(theta) sc^T) (RBMSplit.jl:104-123);
the local estimator is the one of AccumulatorObsGrad.jl:39-127 with this
network; the gradient and the complex SR system follow
BatchedGradSampler.jl:99-118 and SRDirect.jl:26-46 for a complex parameter
type: F = conj(L)^T dL / Ns - C Obar (complex), S_jk = sum_s conj(Oc_js) Oc_ks
/ Ns (Hermitian), Delta = (S + eps I)^-1 F.

Parity status: pinned by the closed form above (checked against a direct
per-configuration evaluation), finite differences of log rho in the real and
imaginary directions of every parameter (holomorphy), the identity
L_loc = (L vec rho) / rho with the dense Kronecker Liouvillian, and the
complex Cholesky solve against scipy (library routine).
"""
from __future__ import annotations

import numpy as np

from oracle import ops_ref


def n_params(N, nh):
    return 2 * N + nh + 2 * nh * N


def unpack(p, N, nh):
    p = np.asarray(p, dtype=np.complex128)
    ar, ac, b = p[:N], p[N:2 * N], p[2 * N:2 * N + nh]
    o = 2 * N + nh
    Wr = p[o:o + nh * N].reshape((nh, N), order="F")
    Wc = p[o + nh * N:o + 2 * nh * N].reshape((nh, N), order="F")
    return ar, ac, b, Wr, Wc


def states(keys, N):
    keys = np.asarray(keys, dtype=np.uint64)
    bits = (keys[:, None] >> np.arange(2 * N, dtype=np.uint64)[None, :]) & np.uint64(1)
    v = 2.0 * bits.astype(np.float64) - 1.0
    return v[:, :N], v[:, N:]


def logL(z):
    return np.log1p(np.exp(z))


def dlogL(z):
    return 1.0 / (1.0 + np.exp(-z))


def logpsi(p, keys, N, nh):
    ar, ac, b, Wr, Wc = unpack(p, N, nh)
    sr, sc = states(keys, N)
    th = b[None, :] + sr @ Wr.T + sc @ Wc.T
    return sr @ ar + sc @ ac + logL(th).sum(1)


def logpsi_grad(p, keys, N, nh):
    ar, ac, b, Wr, Wc = unpack(p, N, nh)
    sr, sc = states(keys, N)
    th = b[None, :] + sr @ Wr.T + sc @ Wc.T
    lp = sr @ ar + sc @ ac + logL(th).sum(1)
    fp = dlogL(th)                                       # B x nh
    B = len(lp)
    O = np.empty((B, n_params(N, nh)), dtype=np.complex128)
    O[:, :N] = sr
    O[:, N:2 * N] = sc
    O[:, 2 * N:2 * N + nh] = fp
    o = 2 * N + nh
    # column-major (h fastest) within Wr, Wc
    O[:, o:o + nh * N] = (fp[:, :, None] * sr[:, None, :]).transpose(0, 2, 1).reshape(B, nh * N)
    O[:, o + nh * N:] = (fp[:, :, None] * sc[:, None, :]).transpose(0, 2, 1).reshape(B, nh * N)
    return lp, O


def local_estimator(model, p, keys, N, nh, grad=True):
    """L_loc(s) = sum_c mel_c exp(log rho_c - log rho_s) and
    dL_loc(s) = sum_c mel_c exp(..) O(c) (no -O(s) term)."""
    keys = np.asarray(keys, dtype=np.uint64)
    terms = ops_ref.terms(model)
    lp = logpsi(p, keys, N, nh)
    L = np.zeros(len(keys), dtype=np.complex128)
    dL = np.zeros((len(keys), n_params(N, nh)), dtype=np.complex128) if grad else None
    for i, k in enumerate(keys):
        conn = ops_ref.connections(model, int(k), _terms=terms)
        if not conn:
            continue
        mel = np.array([c[0] for c in conn], dtype=np.complex128)
        kc = np.array([c[1] for c in conn], dtype=np.uint64)
        if grad:
            lpc, Oc = logpsi_grad(p, kc, N, nh)
        else:
            lpc = logpsi(p, kc, N, nh)
        w = mel * np.exp(lpc - lp[i])
        L[i] = w.sum()
        if grad:
            dL[i] = w @ Oc
    return (L, dL) if grad else L


def grad_step(B, Lc, O, L, dL, stat_analysis=None):
    """Complex-parameter gradient and SR matrix (samples ordered s = chain +
    slice * B): F = sum conj(L) dL / Ns - C Obar, S = Oc^H Oc / Ns."""
    Ns = O.shape[0]
    assert Ns == B * Lc
    C = np.mean(np.abs(L) ** 2)
    Ob = O.mean(0)
    Oc = O - Ob
    F = (np.conj(L) @ dL) / Ns - C * Ob
    S = np.conj(Oc).T @ Oc / Ns
    st = stat_analysis(np.abs(L.reshape(Lc, B).T) ** 2) if stat_analysis else None
    return S, F, Ob, st


def sr_solve(S, F, eps):
    """Delta = (S + eps I)^-1 F by complex Cholesky (SRDirect.jl:60-81)."""
    import scipy.linalg as sl
    A = np.array(S, dtype=np.complex128, copy=True)
    A[np.diag_indices_from(A)] += eps
    c = sl.cho_factor(A, lower=True, check_finite=True)
    return sl.cho_solve(c, F)


def metropolis(p, N, nh, chains, chain_len, passes, burn, seed, it, philox, persistent_state=None):
    """LocalRule Metropolis on |rho|^2 with the device's random streams
    (Philox4x32-10 key = seed, counter (0, it, chain, 0) for the uniform start,
    (t, it, chain, 1) for proposal t: site = (x0 * 2N) >> 32, u from x1, x2 as
    a 53-bit uniform; reject iff u - exp(delta log p) >= 0, Metropolis.jl:148-154).
    `philox(ctr, key)` returns the four output words."""
    k0, k1 = seed & 0xffffffff, seed >> 32
    m2 = (1 << (2 * N)) - 1
    keys = np.zeros(chains, dtype=np.uint64)
    for c in range(chains):
        if persistent_state is None:
            o = philox([0, it, c, 0], [k0, k1])
            keys[c] = np.uint64((o[0] | (o[1] << 32)) & m2)
        else:
            keys[c] = persistent_state[c]
    out = np.zeros((chain_len, chains), dtype=np.uint64)
    calls = (burn if persistent_state is None else 0) + chain_len
    b0 = calls - chain_len
    t = 0
    acc = 0
    for call in range(calls):
        for _ in range(passes):
            lp = 2.0 * logpsi(p, keys, N, nh).real
            for c in range(chains):
                o = philox([t, it, c, 1], [k0, k1])
                site = (o[0] * 2 * N) >> 32
                u = ((o[1] >> 5) * 67108864.0 + (o[2] >> 6)) * (1.0 / 9007199254740992.0)
                kn = np.uint64(int(keys[c]) ^ (1 << site))
                lpn = 2.0 * logpsi(p, np.array([kn]), N, nh).real[0]
                if not (u - np.exp(lpn - lp[c]) >= 0.0):
                    keys[c] = kn
                    acc += 1
            t += 1
        if call >= b0:
            out[call - b0] = keys
    return out.reshape(-1), keys, acc


def exact(model, p, N, nh):
    """Full-sum quantities over all 4^N doubled configurations."""
    keys = np.arange(4 ** N, dtype=np.uint64)
    lp, O = logpsi_grad(p, keys, N, nh)
    L, dL = local_estimator(model, p, keys, N, nh)
    w = np.exp(2.0 * (lp.real - lp.real.max()))
    prob = w / w.sum()
    return dict(keys=keys, logpsi=lp, O=O, Lloc=L, dLloc=dL, prob=prob,
                C=float(np.sum(prob * np.abs(L) ** 2)))


# ---------------------------------------------------------------------------
# Ket RBM and the energy path (src/Networks/ClosedSystems/RBM.jl,
# BatchedValSampler.jl:70-139, AccumulatorObsScalar.jl).
#
#   log psi(s) = a . s + sum_h logL(b_h + (W s)_h),  parameters a[N] b[nh] W[nh x N]
#   (RBM.jl:28-61, activation af_softplus = logL), O = (s, f'(theta), f'(theta) s^T);
#   E_loc(s) = sum_s' <s|H|s'> psi(s') / psi(s);  grad C = sum_s E_loc(s) conj(Oc(s)) / Ns
#   (BatchedValSampler.jl:100-116: Cr * Or' with Oc centred, no conjugation after);
#   S_jk = sum conj(Oc_j) Oc_k / Ns (SRDirect.jl:35-45, complex parameters).

def ket_n_params(N, nh):
    return N + nh + nh * N


def ket_unpack(p, N, nh):
    p = np.asarray(p, dtype=np.complex128)
    return p[:N], p[N:N + nh], p[N + nh:].reshape((nh, N), order="F")


def ket_states(keys, N):
    keys = np.asarray(keys, dtype=np.uint64)
    bits = (keys[:, None] >> np.arange(N, dtype=np.uint64)[None, :]) & np.uint64(1)
    return 2.0 * bits.astype(np.float64) - 1.0


def ket_logpsi_grad(p, keys, N, nh):
    a, b, W = ket_unpack(p, N, nh)
    s = ket_states(keys, N)
    th = b[None, :] + s @ W.T
    lp = s @ a + logL(th).sum(1)
    fp = dlogL(th)
    B = len(lp)
    O = np.empty((B, ket_n_params(N, nh)), dtype=np.complex128)
    O[:, :N] = s
    O[:, N:N + nh] = fp
    O[:, N + nh:] = (fp[:, :, None] * s[:, None, :]).transpose(0, 2, 1).reshape(B, nh * N)
    return lp, O


def ket_logpsi(p, keys, N, nh):
    return ket_logpsi_grad(p, keys, N, nh)[0]


def ket_terms(model):
    """The Hamiltonian of `model` (an ops_ref.Model; its jumps are ignored) as
    the rows of its same-site-merged KLocalOperatorSum terms: [(sites, rows)],
    each row diagonal first, then ascending columns with |mel| >= 1e-5
    (KLocalOperator.jl:75-101, KLocalOperatorSum.jl:65-85)."""
    merged = ops_ref.Model(model.n, model.h_terms, []).hnh_terms()
    return [(s, ops_ref.rows_of(M)) for s, M in merged]


def ket_connections(terms, key):
    """(mel, new key) of <key|H|key'> in the reference's order."""
    out = []
    for sites, rows in terms:
        for mel, fm in rows[ops_ref._local_row(int(key), sites)]:
            out.append((mel, int(key) ^ ops_ref._global_mask(fm, sites)))
    return out


def local_energy(model, p, keys, N, nh):
    """E_loc(s) = sum_s' <s|H|s'> psi(s') / psi(s), zero mels skipped
    (AccumulatorObsScalar.jl)."""
    keys = np.asarray(keys, dtype=np.uint64)
    terms = ket_terms(model)
    lp = ket_logpsi(p, keys, N, nh)
    E = np.zeros(len(keys), dtype=np.complex128)
    for i, k in enumerate(keys):
        conn = [c for c in ket_connections(terms, k) if c[0] != 0]
        mel = np.array([c[0] for c in conn], dtype=np.complex128)
        kc = np.array([c[1] for c in conn], dtype=np.uint64)
        E[i] = np.sum(mel * np.exp(ket_logpsi(p, kc, N, nh) - lp[i]))
    return E


def dense_hamiltonian(model):
    """<s|H|s'> over the 2^N packed keys (site i = bit i - 1), by Kronecker
    products of the local matrices."""
    n = model.n
    D = 2 ** n
    H = np.zeros((D, D), dtype=complex)
    for sites, M in model.h_terms:
        for k in range(D):
            r = ops_ref._local_row(k, sites)
            for c in range(M.shape[1]):
                if M[r, c] != 0:
                    H[k, k ^ ops_ref._global_mask(r ^ c, sites)] += M[r, c]
    return H


def energy_grad_step(B, Lc, O, E):
    """grad C = sum E conj(Oc) / Ns and S = Oc^H Oc / Ns; statistics of the
    complex E over chains (stats.jl:52-77 with complex values: var = mean |x - m|^2)."""
    Ns = O.shape[0]
    Ob = O.mean(0)
    Oc = O - Ob
    F = (E @ np.conj(Oc)) / Ns
    S = np.conj(Oc).T @ Oc / Ns
    v = E.reshape(Lc, B).T                     # [chain, slice]
    mch = v.mean(1)
    mu = mch.mean()
    var_ch = (np.abs(v - mch[:, None]) ** 2).sum(1) / (Lc - 1)
    var_mu_ch = (np.abs(mch - mch.mean()) ** 2).sum() / (B - 1)
    var_mu = (np.abs(v - mu) ** 2).sum() / (B * Lc - 1)
    t = var_mu_ch / var_mu
    st = np.array([mu.real, mu.imag, np.sqrt(var_mu_ch / B), var_ch.mean(), max(0.0, 0.5 * (t * Lc - 1)),
                   np.sqrt((Lc - 1) / Lc + t)])
    return S, F, Ob, st


def ket_metropolis(p, N, nh, chains, chain_len, passes, burn, seed, it, philox):
    """LocalRule Metropolis on |psi|^2 over N sites with the device's streams
    (site = (x0 N) >> 32; other conventions as `metropolis`)."""
    k0, k1 = seed & 0xffffffff, seed >> 32
    m1 = (1 << N) - 1
    keys = np.zeros(chains, dtype=np.uint64)
    for c in range(chains):
        o = philox([0, it, c, 0], [k0, k1])
        keys[c] = np.uint64((o[0] | (o[1] << 32)) & m1)
    out = np.zeros((chain_len, chains), dtype=np.uint64)
    t = 0
    acc = 0
    for call in range(burn + chain_len):
        for _ in range(passes):
            lp = 2.0 * ket_logpsi(p, keys, N, nh).real
            for c in range(chains):
                o = philox([t, it, c, 1], [k0, k1])
                site = (o[0] * N) >> 32
                u = ((o[1] >> 5) * 67108864.0 + (o[2] >> 6)) * (1.0 / 9007199254740992.0)
                kn = np.uint64(int(keys[c]) ^ (1 << site))
                lpn = 2.0 * ket_logpsi(p, np.array([kn]), N, nh).real[0]
                if not (u - np.exp(lpn - lp[c]) >= 0.0):
                    keys[c] = kn
                    acc += 1
            t += 1
        if call >= burn:
            out[call - burn] = keys
    return out.reshape(-1), acc
