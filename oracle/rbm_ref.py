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
