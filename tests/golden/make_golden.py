"""Writes the golden fixtures of the hot path (tests/golden/*.npz).

The reference commits no golden vectors (its tests compute expectations on
the fly), so the fixtures are written from the FP64 oracle and every value is
cross-checked here, before it is saved, against an independent restatement:

* log rho against the numpy closed form of the NDM (NDM.jl:76-96);
* O against central finite differences of that closed form;
* the connection lists against the numpy rule-based enumeration and against
  the dense Kronecker Liouvillian (KLocalLiouvillian.jl:9-52);
* Lloc against (L vec rho) / rho with the dense Liouvillian
  (AccumulatorObsGrad.jl:39-127);
* dLloc against sum_c mel_c rho_c / rho_s O(c) from the cross-checked pieces;
* S, F against their definitions in numpy (BatchedGradSampler.jl:99-118,
  SRDirect.jl:26-46) and the shifted solve against scipy's Cholesky.

Run: python tests/golden/make_golden.py   (rewrites the .npz files)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "neuralquantum.jl_amd"))

from oracle import made_ref, nqref, ops_ref  # noqa: E402

# (name, N, periodic, nh, na, number of keys or None for all 4^N, seed)
CASES = [
    ("ising_n2_open", 2, False, 2, 2, None, 101),
    ("ising_n3_periodic", 3, True, 3, 3, None, 102),
    ("ising_n4_periodic", 4, True, 4, 8, 64, 103),
]


def _states(keys, N):
    b = ((keys[:, None] >> np.arange(2 * N, dtype=np.uint64)[None, :]) & np.uint64(1)).astype(np.float64)
    v = 2.0 * b - 1.0
    return v[:, :N].T, v[:, N:].T


def _check_case(ref, N, nh, na, p, keys, model, table, lp, O, L, dL):
    np_ = ref.ndm_np(N, nh, na)
    s, e = _states(keys, N)
    cf = nqref.ndm_closed_form(N, nh, na, p, s, e)
    assert np.max(np.abs(lp - cf)) < 1e-12
    h = 1e-6
    for k in range(np_):
        dp = np.zeros(np_)
        dp[k] = h
        fd = (nqref.ndm_closed_form(N, nh, na, p + dp, s, e) - nqref.ndm_closed_form(N, nh, na, p - dp, s, e)) / (2 * h)
        assert np.max(np.abs(fd - O[:, k])) < 1e-6
    terms = ops_ref.terms(model)
    Ld = ops_ref.dense_liouvillian_kron(model)
    assert np.allclose(Ld, ops_ref.dense_from_connections(model), atol=1e-13)
    ka = nqref.all_keys(N)
    rho = np.exp(ref.logpsi(N, nh, na, p, ka))
    Lrho = Ld @ rho
    for i, k in enumerate(keys):
        kc, mc = ref.connections(table, int(k))
        conn = ops_ref.connections(model, int(k), _terms=terms)
        mr = np.array([c[0] for c in conn], dtype=complex)
        kr = np.array([c[1] for c in conn], dtype=np.uint64)
        assert np.array_equal(kc, kr)
        assert np.allclose(mc, mr, atol=1e-15)
        assert abs(L[i] - Lrho[int(k)] / rho[int(k)]) < 1e-10 * max(1.0, abs(L[i]))
        lpc, Oc = ref.logpsi_grad(N, nh, na, p, kc)
        w = mc * np.exp(lpc - lp[i])
        assert np.max(np.abs(w @ Oc - dL[i])) < 1e-10 * max(1.0, np.max(np.abs(dL[i])))


def _grad_numpy(O, L, dL):
    Ns = O.shape[0]
    Ob = O.mean(0)
    Oc = O - Ob
    C = np.mean(np.abs(L) ** 2)
    S = np.real(Oc.T @ np.conj(Oc)) / Ns
    F = np.real(np.conj(L) @ dL / Ns - C * Ob)
    return S, F, Ob


def write_ndm_cases(ref):
    for name, N, periodic, nh, na, nkeys, seed in CASES:
        model = ops_ref.ising_chain_model(N, V=2.0, g=1.0, gamma=1.0, periodic=periodic)
        table = nqref.Table(ops_ref.table_from_rules(model))
        rng = np.random.default_rng(seed)
        p = rng.normal(0.0, 0.1, ref.ndm_np(N, nh, na))
        if nkeys is None:
            keys = nqref.all_keys(N)
        else:
            keys = np.sort(rng.choice(4 ** N, size=nkeys, replace=False)).astype(np.uint64)
        lp, O = ref.logpsi_grad(N, nh, na, p, keys)
        L, dL = ref.local_estimator(table, nh, na, p, keys, grad=True)
        _check_case(ref, N, nh, na, p, keys, model, table, lp, O, L, dL)
        B = len(keys)
        S, F, Ob, st = ref.grad_step(B // 4, 4, O, L, dL)   # 4 slices of B/4 chains
        S_n, F_n, Ob_n = _grad_numpy(O, L, dL)
        assert np.max(np.abs(S - S_n)) < 1e-12 * np.max(np.abs(S_n))
        assert np.max(np.abs(F - F_n)) < 1e-12 * np.max(np.abs(F_n))
        assert np.max(np.abs(Ob - Ob_n)) < 1e-13
        delta = nqref.sr_solve(S, F, 0.01)
        assert np.allclose((S + 0.01 * np.eye(len(F))) @ delta, F, atol=1e-10)
        # connections of the first four keys (flattened, with counts)
        ck, cm, cn = [], [], []
        for k in keys[:4]:
            kc, mc = ref.connections(table, int(k))
            ck.append(kc)
            cm.append(mc)
            cn.append(len(kc))
        np.savez_compressed(
            os.path.join(HERE, f"{name}.npz"),
            N=N, periodic=periodic, nh=nh, na=na, V=2.0, g=1.0, gamma=1.0, params=p, keys=keys,
            logpsi=lp, O=O, Lloc=L, dLloc=dL, S=S, F=F, Obar=Ob, stats=st, sr_eps=0.01, sr_delta=delta,
            conn_count=np.asarray(cn), conn_keys=np.concatenate(ck), conn_mel=np.concatenate(cm))
        print(f"wrote {name}: Np={len(p)} keys={B}")


def write_made_case():
    D, H = 8, 128
    rng = np.random.default_rng(7)
    th = made_ref.init(D, H, rng)
    keys = np.arange(2 ** D, dtype=np.uint64)
    lq = made_ref.log_q(th, keys, D, H)
    assert abs(np.exp(lq).sum() - 1.0) < 1e-12
    batch = rng.integers(0, 2 ** D, size=96, dtype=np.uint64)
    nll, g = made_ref.nll_and_grad(th, batch, D, H)
    np.savez_compressed(os.path.join(HERE, "made_d8.npz"), D=D, H=H, theta=th, keys=keys, logq=lq,
                        batch=batch, nll=nll, grad=g)
    print("wrote made_d8")


if __name__ == "__main__":
    ref = nqref.Ref()
    write_ndm_cases(ref)
    write_made_case()
