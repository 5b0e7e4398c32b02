"""ORACLE (test infrastructure only) -- observables of a density-matrix network.

Restates BatchedObsDMSampler.compute_observables
(src/IterativeInterface/Samplers/Obs/BatchedObsDMSampler.jl:59-105): a
Metropolis chain on the physical configurations sigma with p(sigma) proportional to
rho(sigma, sigma), then the local value of a ket operator O

    O_loc(sigma) = sum_sigma' <sigma|O|sigma'> rho(sigma', sigma) / rho(sigma, sigma),

whose p-weighted mean is Tr(O rho) / Tr(rho), with the statistics of
stats.jl:52-77 on the (complex) local values.

Two reference behaviours are corrected here (and in the device code):
  * the diagonal chain's log-probability is abs(log rho(sigma, sigma))
    (base_batched_networks.jl:249-253); it must be Re log rho(sigma, sigma);
  * the accumulator evaluates the network at (sigma', sigma') after applying
    the operator's changes to a physical state (AccumulatorLogPsi.jl:85-91 with
    logpsi!(.., sigma) = logpsi!(.., sigma, sigma) for matrix networks,
    base_batched_networks.jl:242-244), which gives Tr(O rho) only for diagonal O;
    the restatement uses rho(sigma', sigma).
The chain uses the device's random streams: Philox4x32-10 key = seed,
counter (0, it, chain, 16) for the uniform start (N bits), (t, it, chain, 17)
for proposal t: site = (x0 N) >> 32 flipped in both sigma and eta, u from x1, x2
(53 bits), rejected iff u - exp(delta log p) >= 0 (Metropolis.jl:148-154).

Parity status: O_loc pinned by Tr(O rho)/Tr(rho) from the dense density
matrix (full sums over p(sigma)); the chain pinned by its stationary
distribution (TV against rho(sigma, sigma) / Tr rho).
"""
from __future__ import annotations

import numpy as np

from oracle import ops_ref


def diag_keys(sig, N):
    s = np.asarray(sig, dtype=np.uint64)
    return s | (s << np.uint64(N))


def ket_terms(model):
    merged = ops_ref.Model(model.n, model.h_terms, []).hnh_terms()
    return [(s, ops_ref.rows_of(M)) for s, M in merged]


def local_values(logrho, terms, sig, N):
    """O_loc(sigma) for physical keys `sig`; logrho(doubled keys) -> log rho."""
    sig = np.asarray(sig, dtype=np.uint64)
    lp = logrho(diag_keys(sig, N))
    out = np.zeros(len(sig), dtype=complex)
    for i, s in enumerate(sig):
        conn = []
        for sites, rows in terms:
            for mel, fm in rows[ops_ref._local_row(int(s), sites)]:
                if mel != 0:
                    conn.append((mel, int(s) ^ ops_ref._global_mask(fm, sites)))
        mel = np.array([c[0] for c in conn], dtype=complex)
        kc = np.array([c[1] | (int(s) << N) for c in conn], dtype=np.uint64)   # (sigma', sigma)
        out[i] = np.sum(mel * np.exp(logrho(kc) - lp[i]))
    return out


def exact_expectation(logrho, model, N):
    """Tr(O rho) / Tr(rho) from the dense rho (site i = bit i - 1)."""
    D = 2 ** N
    keys = np.arange(D, dtype=np.uint64)
    rho = np.exp(logrho((keys[:, None] | (keys[None, :] << np.uint64(N))).reshape(-1))).reshape(D, D)
    O = np.zeros((D, D), dtype=complex)
    for sites, M in model.h_terms:
        for k in range(D):
            r = ops_ref._local_row(k, sites)
            for c in range(M.shape[1]):
                if M[r, c] != 0:
                    O[k, k ^ ops_ref._global_mask(r ^ c, sites)] += M[r, c]
    return np.trace(O @ rho) / np.trace(rho)


def diag_metropolis(logrho, N, chains, chain_len, passes, burn, seed, it, philox, tag=1):
    k0, k1 = seed & 0xffffffff, seed >> 32
    c3 = 16 * tag
    m1 = (1 << N) - 1
    sig = np.zeros(chains, dtype=np.uint64)
    for c in range(chains):
        o = philox([0, it, c, c3], [k0, k1])
        sig[c] = np.uint64((o[0] | (o[1] << 32)) & m1)
    out = np.zeros((chain_len, chains), dtype=np.uint64)
    t = 0
    acc = 0
    for call in range(burn + chain_len):
        for _ in range(passes):
            lp = logrho(diag_keys(sig, N)).real
            for c in range(chains):
                o = philox([t, it, c, c3 + 1], [k0, k1])
                site = (o[0] * N) >> 32
                u = ((o[1] >> 5) * 67108864.0 + (o[2] >> 6)) * (1.0 / 9007199254740992.0)
                sn = np.uint64(int(sig[c]) ^ (1 << site))
                lpn = logrho(diag_keys([sn], N)).real[0]
                if not (u - np.exp(lpn - lp[c]) >= 0.0):
                    sig[c] = sn
                    acc += 1
            t += 1
        if call >= burn:
            out[call - burn] = sig
    return diag_keys(out.reshape(-1), N), acc


def complex_stats(vals, B, Lc):
    """stat_analysis of complex values s = chain + slice B:
    [Re mean, error, mean chain variance, tau, R, Im mean]."""
    v = np.asarray(vals).reshape(Lc, B).T
    mch = v.mean(1)
    mu = mch.mean()
    var_ch = (np.abs(v - mch[:, None]) ** 2).sum(1) / (Lc - 1)
    var_mu_ch = (np.abs(mch - mch.mean()) ** 2).sum() / (B - 1)
    var_mu = (np.abs(v - mu) ** 2).sum() / (B * Lc - 1)
    t = var_mu_ch / var_mu
    return np.array([mu.real, np.sqrt(var_mu_ch / B), var_ch.mean(), max(0.0, 0.5 * (t * Lc - 1)),
                     np.sqrt((Lc - 1) / Lc + t), mu.imag])
