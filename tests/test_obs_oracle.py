"""Pins of the observables oracle (oracle/obs_ref.py): the p-weighted local
values equal Tr(O rho)/Tr(rho) from the dense density matrix, for diagonal
and off-diagonal operators, and the diagonal chain is stationary."""
import numpy as np

from oracle import nqref, obs_ref, ops_ref

SX = ops_ref.SX
SY = ops_ref.SY
SZ = ops_ref.SZ


def site_sum(N, M):
    return ops_ref.Model(N, [((i,), M / N) for i in range(1, N + 1)], [])


def test_local_values_average_to_trace(ref):
    N, nh, na = 3, 3, 3
    p = np.random.default_rng(1).normal(0, 0.3, ref.ndm_np(N, nh, na))
    logrho = lambda k: ref.logpsi(N, nh, na, p, k)   # noqa: E731
    sig = np.arange(2 ** N, dtype=np.uint64)
    w = np.exp(logrho(obs_ref.diag_keys(sig, N)).real)
    w /= w.sum()
    for M in (SX, SY, SZ):
        model = site_sum(N, M)
        loc = obs_ref.local_values(logrho, obs_ref.ket_terms(model), sig, N)
        ex = obs_ref.exact_expectation(logrho, model, N)
        assert abs(np.sum(w * loc) - ex) < 1e-12
    # a two-site operator
    model = ops_ref.Model(N, [((1, 2), ops_ref.pair_matrix(SX, SZ))], [])
    loc = obs_ref.local_values(logrho, obs_ref.ket_terms(model), sig, N)
    assert abs(np.sum(w * loc) - obs_ref.exact_expectation(logrho, model, N)) < 1e-12


def test_diagonal_chain_is_stationary(ref):
    N, nh, na = 3, 3, 3
    p = np.random.default_rng(2).normal(0, 0.4, ref.ndm_np(N, nh, na))
    logrho = lambda k: ref.logpsi(N, nh, na, p, k)   # noqa: E731
    keys, _ = obs_ref.diag_metropolis(logrho, N, 64, 40, 5, 10, seed=3, it=0, philox=ref.philox)
    sig = keys & np.uint64(2 ** N - 1)
    assert np.array_equal(keys >> np.uint64(N), sig)
    w = np.exp(logrho(obs_ref.diag_keys(np.arange(2 ** N, dtype=np.uint64), N)).real)
    w /= w.sum()
    hist = np.bincount(sig.astype(np.int64), minlength=2 ** N) / len(sig)
    assert 0.5 * np.abs(hist - w).sum() < 0.05
