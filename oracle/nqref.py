"""ORACLE (test infrastructure only) -- Python front of the C restatement.

Only `tests/`, `__graft_entry__.smoke()` and `bench.py`'s CPU-baseline leg use
this module, always as the checker; the product never imports it.

* `Ref`: ctypes binding of oracle/libnqref.so (oracle/nqref.c).
* `ndm_closed_form`: a second, independent numpy restatement of the NDM
  (src/Networks/MixedDensityMatrix/NDM.jl:76-96) used to cross-check the C one.
* `exact_*`: full sums over all 4^N doubled configurations (the exact
  sampler / densitymatrix oracles, src/Samplers/Exact.jl:110-181,
  src/utils/densitymatrix.jl:9-32).
* `sr_solve`: additive-shift Cholesky solve with LAPACK through scipy
  (SRDirect.jl:51-90, `sr_shift` + `sr_cholesky`).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libnqref.so")

_dp = C.POINTER(C.c_double)
_u64p = C.POINTER(C.c_uint64)
_i32p = C.POINTER(C.c_int32)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class _Table(C.Structure):
    _fields_ = [("n_sites", C.c_int32), ("n_terms", C.c_int32),
                ("term_kind", _i32p), ("term_op_l", _i32p), ("term_op_r", _i32p),
                ("op_site_off", _i32p), ("op_sites", _i32p), ("op_row_off", _i32p),
                ("row_ent_off", _i32p), ("ent_mel", _dp), ("ent_flip", _u64p)]


def _ptr(a, t):
    return a.ctypes.data_as(t)


class Table:
    """Keeps the numpy arrays of a flat table alive behind the C struct."""

    def __init__(self, d):
        self.arrays = {
            "term_kind": np.ascontiguousarray(d["term_kind"], dtype=np.int32),
            "term_op_l": np.ascontiguousarray(d["term_op_l"], dtype=np.int32),
            "term_op_r": np.ascontiguousarray(d["term_op_r"], dtype=np.int32),
            "op_site_off": np.ascontiguousarray(d["op_site_off"], dtype=np.int32),
            "op_sites": np.ascontiguousarray(d["op_sites"], dtype=np.int32),
            "op_row_off": np.ascontiguousarray(d["op_row_off"], dtype=np.int32),
            "row_ent_off": np.ascontiguousarray(d["row_ent_off"], dtype=np.int32),
            "ent_mel": np.ascontiguousarray(np.asarray(d["ent_mel"], dtype=np.complex128).view(np.float64)),
            "ent_flip": np.ascontiguousarray(d["ent_flip"], dtype=np.uint64),
        }
        a = self.arrays
        self.n_sites = int(d["n_sites"])
        self.s = _Table(self.n_sites, len(a["term_kind"]),
                        _ptr(a["term_kind"], _i32p), _ptr(a["term_op_l"], _i32p),
                        _ptr(a["term_op_r"], _i32p), _ptr(a["op_site_off"], _i32p),
                        _ptr(a["op_sites"], _i32p), _ptr(a["op_row_off"], _i32p),
                        _ptr(a["row_ent_off"], _i32p), _ptr(a["ent_mel"], _dp),
                        _ptr(a["ent_flip"], _u64p))


class Ref:
    """act: the NDM's activation f (NDM.jl:50), 0 = softplus (af_softplus,
    the default), 1 = logcosh (af_logcosh, activation.jl:5-10); set before
    every network evaluation of this object."""

    def __init__(self, act=0):
        if not os.path.exists(_LIB):
            build()
        self.act = int(act)
        L = C.CDLL(_LIB)
        L.nqref_set_activation.argtypes = [C.c_int]
        L.nqref_set_threads.argtypes = [C.c_int]
        L.nqref_ndm_np.restype = C.c_int64
        L.nqref_ndm_np.argtypes = [C.c_int, C.c_int, C.c_int]
        L.nqref_ndm_logpsi.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_int64, _u64p, _dp]
        L.nqref_ndm_logpsi_grad.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_int64, _u64p, _dp, _dp]
        L.nqref_connections.restype = C.c_int
        L.nqref_connections.argtypes = [C.POINTER(_Table), C.c_uint64, C.c_int, C.c_int, _u64p, _dp]
        L.nqref_local_estimator.argtypes = [C.POINTER(_Table), C.c_int, C.c_int, _dp, C.c_int64,
                                            _u64p, _dp, _dp]
        L.nqref_stat_analysis.argtypes = [C.c_int64, C.c_int64, _dp, _dp]
        L.nqref_grad_step.argtypes = [C.c_int64, C.c_int64, C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.nqref_metropolis.restype = C.c_int64
        L.nqref_metropolis.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_int64, C.c_int64, C.c_int,
                                       C.c_int, C.c_uint64, C.c_uint32, C.c_int, _u64p]
        L.nqref_philox.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        self.L = L
        L.nqref_set_threads(1)

    def set_threads(self, n):
        """Threads of the loops over samples and chains (OpenMP); 1 by default."""
        self.L.nqref_set_threads(int(n))

    # -- network -----------------------------------------------------------
    def ndm_np(self, N, nh, na):
        return int(self.L.nqref_ndm_np(N, nh, na))

    def logpsi(self, N, nh, na, p, keys):
        p = np.ascontiguousarray(p, dtype=np.float64)
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        out = np.empty(2 * len(keys))
        self.L.nqref_set_activation(self.act)
        self.L.nqref_ndm_logpsi(N, nh, na, _ptr(p, _dp), len(keys), _ptr(keys, _u64p), _ptr(out, _dp))
        return out.view(np.complex128)

    def logpsi_grad(self, N, nh, na, p, keys):
        p = np.ascontiguousarray(p, dtype=np.float64)
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        np_ = self.ndm_np(N, nh, na)
        out = np.empty(2 * len(keys))
        O = np.empty(2 * len(keys) * np_)
        self.L.nqref_set_activation(self.act)
        self.L.nqref_ndm_logpsi_grad(N, nh, na, _ptr(p, _dp), len(keys), _ptr(keys, _u64p),
                                     _ptr(out, _dp), _ptr(O, _dp))
        return out.view(np.complex128), O.view(np.complex128).reshape(len(keys), np_)

    # -- operator ------------------------------------------------------------
    def connections(self, table: Table, key, skip_zero=True, cap=4096):
        ko = np.empty(cap, dtype=np.uint64)
        mo = np.empty(2 * cap)
        n = self.L.nqref_connections(C.byref(table.s), int(key), int(skip_zero), cap,
                                     _ptr(ko, _u64p), _ptr(mo, _dp))
        if n > cap:
            raise RuntimeError("capacity")
        return ko[:n].copy(), mo[:2 * n].view(np.complex128).copy()

    def local_estimator(self, table: Table, nh, na, p, keys, grad=True):
        p = np.ascontiguousarray(p, dtype=np.float64)
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        N = table.n_sites
        np_ = self.ndm_np(N, nh, na)
        Lloc = np.empty(2 * len(keys))
        dL = np.empty(2 * len(keys) * np_) if grad else None
        self.L.nqref_set_activation(self.act)
        self.L.nqref_local_estimator(C.byref(table.s), nh, na, _ptr(p, _dp), len(keys),
                                     _ptr(keys, _u64p), _ptr(Lloc, _dp),
                                     _ptr(dL, _dp) if grad else None)
        Lloc = Lloc.view(np.complex128)
        if grad:
            return Lloc, dL.view(np.complex128).reshape(len(keys), np_)
        return Lloc

    # -- statistics / gradient -------------------------------------------------
    def stat_analysis(self, vals_BxL):
        """vals indexed [chain j, slice i] (local_vals layout)."""
        v = np.asfortranarray(vals_BxL, dtype=np.float64)
        B, Lc = v.shape
        out = np.empty(5)
        flat = np.ascontiguousarray(v.reshape(-1, order="F"))
        self.L.nqref_stat_analysis(B, Lc, _ptr(flat, _dp), _ptr(out, _dp))
        return out

    def grad_step(self, B, Lc, O, Lloc, dLloc):
        Ns, np_ = O.shape
        assert Ns == B * Lc
        O = np.ascontiguousarray(O, dtype=np.complex128)
        Lloc = np.ascontiguousarray(Lloc, dtype=np.complex128)
        dLloc = np.ascontiguousarray(dLloc, dtype=np.complex128)
        S = np.empty((np_, np_), order="F")
        F = np.empty(np_)
        Ob = np.empty(2 * np_)
        st = np.empty(5)
        self.L.nqref_grad_step(np_, B, Lc, _ptr(O.view(np.float64), _dp), _ptr(Lloc.view(np.float64), _dp),
                               _ptr(dLloc.view(np.float64), _dp), _ptr(S, _dp), _ptr(F, _dp),
                               _ptr(Ob, _dp), _ptr(st, _dp))
        return S, F, Ob.view(np.complex128), st

    # -- sampler -----------------------------------------------------------------
    def metropolis(self, N, nh, na, p, chains, chain_len, passes, burn, seed, it=0, compat=False):
        p = np.ascontiguousarray(p, dtype=np.float64)
        out = np.empty(chains * chain_len, dtype=np.uint64)
        self.L.nqref_set_activation(self.act)
        acc = self.L.nqref_metropolis(N, nh, na, _ptr(p, _dp), chains, chain_len, passes, burn,
                                      C.c_uint64(seed), C.c_uint32(it), int(compat), _ptr(out, _u64p))
        return out, int(acc)

    def philox(self, ctr, key):
        c = (C.c_uint32 * 4)(*ctr)
        k = (C.c_uint32 * 2)(*key)
        o = (C.c_uint32 * 4)()
        self.L.nqref_philox(c, k, o)
        return list(o)


# ---------------------------------------------------------------------------
def ndm_layout(N, nh, na):
    """Offsets of the NDM blocks in the flat parameter vector (NDM.jl:17-18)."""
    sizes = [("b_mu", N), ("h_mu", nh), ("w_mu", nh * N), ("u_mu", na * N), ("b_la", N),
             ("h_la", nh), ("d_la", na), ("w_la", nh * N), ("u_la", na * N)]
    off, p = {}, 0
    for name, n in sizes:
        off[name] = (p, n)
        p += n
    return off, p


def ndm_closed_form(N, nh, na, p, s, e, act=0):
    """log rho(s, e) by the NDM formula with numpy, vectorised over columns of
    s, e (N x B); act 1 takes f = log cosh (numpy's complex log of cosh: the
    principal value, which the reference's logL2(a) + log(cos b + i tanh a
    sin b) equals since cosh a > 0)."""
    off, _ = ndm_layout(N, nh, na)

    def blk(name, shape=None):
        a, n = off[name]
        v = p[a:a + n]
        return v.reshape(shape, order="F") if shape else v

    def sp(x):
        return np.log(np.cosh(x)) if act else np.log1p(np.exp(x))
    wl, wm = blk("w_la", (nh, N)), blk("w_mu", (nh, N))
    ul, um = blk("u_la", (na, N)), blk("u_mu", (na, N))
    hl, hm, dl = blk("h_la")[:, None], blk("h_mu")[:, None], blk("d_la")[:, None]
    bl, bm = blk("b_la"), blk("b_mu")
    gl = 0.5 * (sp(wl @ s + hl).sum(0) + sp(wl @ e + hl).sum(0) + bl @ (s + e))
    gm = 0.5 * (sp(wm @ s + hm).sum(0) - sp(wm @ e + hm).sum(0) + bm @ (s - e))
    arg = 0.5 * (ul @ (s + e)) + dl + 0.5j * (um @ (s - e))
    return gl + 1j * gm + sp(arg).sum(0)


def all_keys(N):
    return np.arange(4 ** N, dtype=np.uint64)


def exact_quantities(ref: Ref, table: Table, nh, na, p):
    """Full-sum oracle over all doubled configurations: probabilities p(x),
    Lloc, dLloc, O, C_exact = sum p |Lloc|^2 and the p-weighted versions of
    O-bar and grad C (what the estimator converges to)."""
    N = table.n_sites
    keys = all_keys(N)
    lp, O = ref.logpsi_grad(N, nh, na, p, keys)
    Lloc, dL = ref.local_estimator(table, nh, na, p, keys, grad=True)
    w = np.exp(2.0 * (lp.real - lp.real.max()))
    prob = w / w.sum()
    C_ex = float(np.sum(prob * np.abs(Lloc) ** 2))
    Obar = prob @ O
    gradC = (prob * np.conj(Lloc)) @ dL - C_ex * Obar
    return dict(keys=keys, logpsi=lp, prob=prob, O=O, Lloc=Lloc, dLloc=dL, C=C_ex,
                Obar=Obar, gradC=gradC)


def sr_solve(S, F, eps):
    """Delta = (S + eps I)^-1 F by Cholesky (SRDirect.jl:60-81)."""
    import scipy.linalg as sl
    A = np.array(S, dtype=np.float64, copy=True)
    A[np.diag_indices_from(A)] += eps
    c = sl.cho_factor(A, lower=True, check_finite=True)
    return sl.cho_solve(c, F)
