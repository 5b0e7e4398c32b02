"""ctypes binding of libnqhip.so (include/nqhip.h).

The library is loaded from this package directory.  There is no fallback: if
the library is missing or cannot be loaded, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnqhip.so")

NQH_OK = 0
NQH_E_ARG = -1
NQH_E_HIP = -2
NQH_E_NOTPD = -3
NQH_E_NAN = -4
NQH_E_STATE = -5
NQH_E_NOMEM = -6
NQH_E_CAPACITY = -7

PREC_FP64, PREC_FP32, PREC_BF16 = 0, 1, 2
ACT_SOFTPLUS, ACT_LOGCOSH = 0, 1
SHIFT_ADD, SHIFT_DIAG = 0, 1
SOLVER_CHOLESKY, SOLVER_CG, SOLVER_MINRES, SOLVER_MINRES_QLP, SOLVER_LSQR = 0, 1, 2, 3, 4
SAMPLER_COMPAT_STALE_LAST_SLICE = 1
SAMPLER_PERSISTENT = 2
BATCH_PACKED, BATCH_GRAPH = 0, 1

_dp = C.POINTER(C.c_double)
_u64p = C.POINTER(C.c_uint64)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_fp = C.POINTER(C.c_float)


class NQHError(RuntimeError):
    """A non-zero nqh_status; `.status` holds the code."""

    def __init__(self, status, msg):
        super().__init__(f"nqhip error {status}: {msg}")
        self.status = status


class OpTable(C.Structure):
    _fields_ = [("n_sites", C.c_int32), ("n_terms", C.c_int32), ("n_ops", C.c_int32),
                ("n_rows", C.c_int32), ("n_entries", C.c_int32),
                ("term_kind", _i32p), ("term_op_l", _i32p), ("term_op_r", _i32p),
                ("op_site_off", _i32p), ("op_sites", _i32p), ("op_row_off", _i32p),
                ("row_ent_off", _i32p), ("ent_mel", _dp), ("ent_flip", _u64p)]


SIGNATURES = {
    "nqh_abi_version": (C.c_int32, []),
    "nqh_create": (C.c_int32, [C.POINTER(C.c_void_p), C.c_int32, C.c_uint64]),
    "nqh_destroy": (None, [C.c_void_p]),
    "nqh_last_error": (C.c_char_p, [C.c_void_p]),
    "nqh_set_problem": (C.c_int32, [C.c_void_p, C.POINTER(OpTable)]),
    "nqh_set_ndm": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "nqh_num_params": (C.c_int64, [C.c_void_p]),
    "nqh_num_derivs": (C.c_int64, [C.c_void_p]),
    "nqh_set_rbmsplit": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32]),
    "nqh_set_rbm": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32]),
    "nqh_get_stats": (C.c_int32, [C.c_void_p, _dp, C.c_int32]),
    "nqh_exact": (C.c_int32, [C.c_void_p, _dp, _dp]),
    "nqh_exact_noalloc": (C.c_int32, [C.c_void_p, _dp, _dp]),
    "nqh_state_bytes": (C.c_int64, [C.c_void_p]),
    "nqh_save_state": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int64]),
    "nqh_load_state": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int64]),
    "nqh_observable": (C.c_int32, [C.c_void_p, C.POINTER(OpTable), C.c_int64, C.c_int32, C.c_int32, C.c_int32,
                                   _dp]),
    "nqh_get_observable_samples": (C.c_int32, [C.c_void_p, _u64p, _dp, C.c_int64]),
    "nqh_set_params": (C.c_int32, [C.c_void_p, _dp, C.c_int64]),
    "nqh_get_params": (C.c_int32, [C.c_void_p, _dp, C.c_int64]),
    "nqh_set_precision": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32]),
    "nqh_logpsi": (C.c_int32, [C.c_void_p, _u64p, C.c_int64, _dp]),
    "nqh_logpsi_grad": (C.c_int32, [C.c_void_p, _u64p, C.c_int64, _dp, _dp]),
    "nqh_connections": (C.c_int32, [C.c_void_p, _u64p, C.c_int64, C.c_int32, _i32p, _u64p, _dp]),
    "nqh_local_estimator": (C.c_int32, [C.c_void_p, _u64p, C.c_int64, _dp, _dp]),
    "nqh_grad_from_samples": (C.c_int32, [C.c_void_p, _u64p, C.c_int64, C.c_int64, _dp, _dp, _dp, _dp]),
    "nqh_set_sampler": (C.c_int32, [C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "nqh_set_iteration": (C.c_int32, [C.c_void_p, C.c_uint32]),
    "nqh_sample": (C.c_int32, [C.c_void_p, _u64p, _i64p]),
    "nqh_get_chain_acceptance": (C.c_int32, [C.c_void_p, C.POINTER(C.c_uint32), C.c_int64]),
    "nqh_sample_grad": (C.c_int32, [C.c_void_p, _dp]),
    "nqh_precondition": (C.c_int32, [C.c_void_p, C.c_double, C.c_int32, C.c_int32, _dp]),
    "nqh_update": (C.c_int32, [C.c_void_p, C.c_double]),
    "nqh_sr_iteration": (C.c_int32, [C.c_void_p, C.c_double, C.c_int32, C.c_int32, C.c_double, _dp]),
    "nqh_get_samples": (C.c_int32, [C.c_void_p, _u64p, C.c_int64]),
    "nqh_get_local_values": (C.c_int32, [C.c_void_p, _dp, C.c_int64]),
    "nqh_set_profiling": (C.c_int32, [C.c_void_p, C.c_int32]),
    "nqh_get_phase_times": (C.c_int32, [C.c_void_p, _dp, _i64p]),
    "nqh_set_graph": (C.c_int32, [C.c_void_p, C.c_int32]),
    "nqh_set_launch_shape": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "nqh_get_launch_shape": (C.c_int32, [C.c_void_p, _i32p]),
    "nqh_sr_iteration_async": (C.c_int32, [C.c_void_p, C.c_double, C.c_int32, C.c_int32, C.c_double]),
    "nqh_wait": (C.c_int32, [C.c_void_p, _dp]),
    "nqh_set_sr_solver": (C.c_int32, [C.c_void_p, C.c_int32, C.c_double, C.c_int32]),
    "nqh_get_sr_solver_info": (C.c_int32, [C.c_void_p, _i32p, _dp]),
    "nqh_batch_create": (C.c_int32, [C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_int32]),
    "nqh_batch_destroy": (None, [C.c_void_p]),
    "nqh_batch_last_error": (C.c_char_p, [C.c_void_p]),
    "nqh_batch_sr_iteration": (C.c_int32, [C.c_void_p, C.c_double, C.c_int32, C.c_int32, C.c_double, _dp]),
    "nqh_batch_set_mode": (C.c_int32, [C.c_void_p, C.c_int32]),
    "nqh_batch_launches": (C.c_int32, [C.c_void_p]),
    "nqh_made_num_params": (C.c_int64, [C.c_int32]),
    "nqh_made_setup": (C.c_int32, [C.c_void_p, C.c_double, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                   C.c_int64]),
    "nqh_made_set_params": (C.c_int32, [C.c_void_p, _fp, C.c_int64]),
    "nqh_made_get_params": (C.c_int32, [C.c_void_p, _fp, C.c_int64]),
    "nqh_made_grad": (C.c_int32, [C.c_void_p, _u64p, C.c_int64, _fp, _dp]),
    "nqh_made_train": (C.c_int32, [C.c_void_p, _u64p, C.c_int64, C.c_int32, C.c_int32, C.c_double, _dp]),
    "nqh_made_logq": (C.c_int32, [C.c_void_p, _u64p, C.c_int64, _fp]),
    "nqh_made_sample": (C.c_int32, [C.c_void_p, C.c_int64, C.c_uint32, C.c_uint32, _u64p, _fp]),
    "nqh_made_stats": (C.c_int32, [C.c_void_p, _i64p, _i64p, _i64p]),
}

NUM_PHASES = 7
PHASES = ("sampler", "estimator", "post", "sbuild", "solve", "update", "factor")

_lib = None


def load():
    """Load libnqhip.so (raises OSError if it is missing: no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} not built; run neuralquantum.jl_amd/build.py")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


class Table:
    """ctypes view of a FlatOpTable (keeps the arrays alive)."""

    def __init__(self, flat):
        self.arrays = dict(
            term_kind=np.ascontiguousarray(flat.term_kind, dtype=np.int32),
            term_op_l=np.ascontiguousarray(flat.term_op_l, dtype=np.int32),
            term_op_r=np.ascontiguousarray(flat.term_op_r, dtype=np.int32),
            op_site_off=np.ascontiguousarray(flat.op_site_off, dtype=np.int32),
            op_sites=np.ascontiguousarray(flat.op_sites, dtype=np.int32),
            op_row_off=np.ascontiguousarray(flat.op_row_off, dtype=np.int32),
            row_ent_off=np.ascontiguousarray(flat.row_ent_off, dtype=np.int32),
            ent_mel=np.ascontiguousarray(np.asarray(flat.ent_mel, dtype=np.complex128).view(np.float64)),
            ent_flip=np.ascontiguousarray(flat.ent_flip, dtype=np.uint64),
        )
        a = self.arrays
        n_ops = len(a["op_site_off"]) - 1
        n_rows = len(a["row_ent_off"]) - 1
        self.s = OpTable(int(flat.n_sites), len(a["term_kind"]), n_ops, n_rows, len(a["ent_flip"]),
                         _p(a["term_kind"], _i32p), _p(a["term_op_l"], _i32p), _p(a["term_op_r"], _i32p),
                         _p(a["op_site_off"], _i32p), _p(a["op_sites"], _i32p), _p(a["op_row_off"], _i32p),
                         _p(a["row_ent_off"], _i32p), _p(a["ent_mel"], _dp), _p(a["ent_flip"], _u64p))


class Context:
    """Owns one nqh_ctx (one device, one stream, one SR loop)."""

    def __init__(self, device: int = 0, seed: int = 0):
        self.lib = load()
        h = C.c_void_p()
        st = self.lib.nqh_create(C.byref(h), int(device), C.c_uint64(seed))
        if st != NQH_OK:
            raise NQHError(st, "nqh_create failed (no HIP device?)")
        self.h = h
        self.np_ = None
        self.N = None

    def close(self):
        if getattr(self, "h", None):
            self.lib.nqh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, st):
        if st != NQH_OK:
            raise NQHError(st, self.lib.nqh_last_error(self.h).decode())

    # -- setup --------------------------------------------------------------
    def set_problem(self, flat):
        self._table = Table(flat)
        self._chk(self.lib.nqh_set_problem(self.h, C.byref(self._table.s)))

    def set_ndm(self, N, nh, na, act=0):
        self._chk(self.lib.nqh_set_ndm(self.h, N, nh, na, act))
        self.np_ = int(self.lib.nqh_num_params(self.h))
        self.nd = int(self.lib.nqh_num_derivs(self.h))
        self.N = N
        self.complex_params = False
        self.ket = False

    def set_rbmsplit(self, N, nh):
        """RBMSplit network; parameters are complex, passed interleaved
        (re, im) in the reference order ar, ac, b, Wr, Wc."""
        self._chk(self.lib.nqh_set_rbmsplit(self.h, N, nh))
        self.np_ = int(self.lib.nqh_num_params(self.h))
        self.nd = int(self.lib.nqh_num_derivs(self.h))
        self.N = N
        self.complex_params = True
        self.ket = False

    def set_rbm(self, N, nh):
        """Ket RBM for a Hamiltonian problem; complex parameters interleaved
        (re, im) in the reference order a, b, W."""
        self._chk(self.lib.nqh_set_rbm(self.h, N, nh))
        self.np_ = int(self.lib.nqh_num_params(self.h))
        self.nd = int(self.lib.nqh_num_derivs(self.h))
        self.N = N
        self.complex_params = True
        self.ket = True

    def exact(self, want_logp=False):
        """Exact sums over every configuration: (sum p |L|^2, sum p L, log Z[, log p])."""
        out = np.empty(4)
        bits = self.N if getattr(self, "ket", False) else 2 * self.N
        lp = np.empty(1 << bits) if want_logp else None
        self._chk(self.lib.nqh_exact(self.h, _p(out, _dp), _p(lp, _dp)))
        res = (float(out[0]), complex(out[1], out[2]), float(out[3]))
        return res + (lp,) if want_logp else res

    def observable(self, flat_op, chains, chain_length, passes, burn):
        """compute_observables for one ket operator (a FlatOpTable of kind-3
        terms): [Re mean, error, variance, tau, R, Im mean]."""
        t = Table(flat_op)
        out = np.empty(6)
        self._chk(self.lib.nqh_observable(self.h, C.byref(t.s), chains, chain_length, passes, burn, _p(out, _dp)))
        self._obs_n = int(chains) * int(chain_length)
        return out

    def observable_samples(self):
        n = self._obs_n
        keys = np.empty(n, dtype=np.uint64)
        L = np.empty(2 * n)
        self._chk(self.lib.nqh_get_observable_samples(self.h, _p(keys, _u64p), _p(L, _dp), n))
        return keys, L.view(np.complex128)

    def get_stats(self, n=6):
        out = np.empty(n)
        self._chk(self.lib.nqh_get_stats(self.h, _p(out, _dp), n))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64)
        self._chk(self.lib.nqh_set_params(self.h, _p(p, _dp), len(p)))

    def get_params(self):
        p = np.empty(self.np_)
        self._chk(self.lib.nqh_get_params(self.h, _p(p, _dp), self.np_))
        return p

    def set_precision(self, net_prec=PREC_FP64, sbuild_prec=PREC_FP64):
        self._chk(self.lib.nqh_set_precision(self.h, net_prec, sbuild_prec))

    # -- parity entry points ------------------------------------------------
    def logpsi(self, keys):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        out = np.empty(2 * len(keys))
        self._chk(self.lib.nqh_logpsi(self.h, _p(keys, _u64p), len(keys), _p(out, _dp)))
        return out.view(np.complex128)

    def logpsi_grad(self, keys):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        out = np.empty(2 * len(keys))
        O = np.empty(2 * len(keys) * self.nd)
        self._chk(self.lib.nqh_logpsi_grad(self.h, _p(keys, _u64p), len(keys), _p(out, _dp), _p(O, _dp)))
        return out.view(np.complex128), O.view(np.complex128).reshape(len(keys), self.nd)

    def connections(self, keys, cap):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        B = len(keys)
        cnt = np.empty(B, dtype=np.int32)
        ko = np.empty(B * cap, dtype=np.uint64)
        mo = np.empty(2 * B * cap)
        self._chk(self.lib.nqh_connections(self.h, _p(keys, _u64p), B, cap, _p(cnt, _i32p),
                                           _p(ko, _u64p), _p(mo, _dp)))
        return cnt, ko.reshape(B, cap), mo.view(np.complex128).reshape(B, cap)

    def local_estimator(self, keys, grad=False):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        B = len(keys)
        L = np.empty(2 * B)
        G = np.empty(2 * B * self.nd) if grad else None
        self._chk(self.lib.nqh_local_estimator(self.h, _p(keys, _u64p), B, _p(L, _dp), _p(G, _dp)))
        L = L.view(np.complex128)
        if grad:
            return L, G.view(np.complex128).reshape(B, self.nd)
        return L

    def grad_from_samples(self, keys, chains, chain_length, want_S=True):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        st = np.empty(5)
        S = np.empty((self.np_, self.np_), order="F") if want_S else None
        F = np.empty(self.np_)
        Ob = np.empty(2 * self.nd)
        self._chk(self.lib.nqh_grad_from_samples(self.h, _p(keys, _u64p), chains, chain_length,
                                                 _p(st, _dp), _p(S, _dp), _p(F, _dp), _p(Ob, _dp)))
        if self.complex_params:   # F complex; S returned as its real embedding [[A, -B], [B, A]]
            return st, S, F.view(np.complex128), Ob.view(np.complex128)
        return st, S, F, Ob.view(np.complex128)

    # -- loop -------------------------------------------------------------------
    def set_sampler(self, chains, chain_length, passes, burn, flags=0):
        self.chains, self.chain_length = int(chains), int(chain_length)
        self._chk(self.lib.nqh_set_sampler(self.h, chains, chain_length, passes, burn, flags))

    def set_iteration(self, it):
        self._chk(self.lib.nqh_set_iteration(self.h, C.c_uint32(it)))

    def sample(self):
        keys = np.empty(self.chains * self.chain_length, dtype=np.uint64)
        acc = C.c_int64()
        self._chk(self.lib.nqh_sample(self.h, _p(keys, _u64p), C.byref(acc)))
        return keys, int(acc.value)

    def save_state(self):
        """The chains, the iteration index and the learned proposal's state
        (nqh_save_state) as bytes."""
        n = int(self.lib.nqh_state_bytes(self.h))
        buf = np.empty(n, dtype=np.uint8)
        self._chk(self.lib.nqh_save_state(self.h, buf.ctypes.data_as(C.c_void_p), n))
        return buf.tobytes()

    def load_state(self, blob):
        buf = np.frombuffer(blob, dtype=np.uint8).copy()
        self._chk(self.lib.nqh_load_state(self.h, buf.ctypes.data_as(C.c_void_p), len(buf)))

    def chain_acceptance(self):
        """Accepted moves per chain in the last sampling call (0: a stuck chain)."""
        out = np.empty(self.chains, dtype=np.uint32)
        self._chk(self.lib.nqh_get_chain_acceptance(self.h, out.ctypes.data_as(C.POINTER(C.c_uint32)), self.chains))
        return out

    def sample_grad(self):
        st = np.empty(5)
        self._chk(self.lib.nqh_sample_grad(self.h, _p(st, _dp)))
        return st

    def precondition(self, eps, shift_kind=SHIFT_ADD, refine=0, want_delta=True):
        d = np.empty(self.np_) if want_delta else None
        self._chk(self.lib.nqh_precondition(self.h, eps, shift_kind, refine, _p(d, _dp)))
        if d is not None and self.complex_params:
            return d.view(np.complex128)
        return d

    def update(self, lr):
        self._chk(self.lib.nqh_update(self.h, lr))

    def sr_iteration(self, eps, lr, shift_kind=SHIFT_ADD, refine=0):
        st = np.empty(5)
        self._chk(self.lib.nqh_sr_iteration(self.h, eps, shift_kind, refine, lr, _p(st, _dp)))
        return st

    def sr_iteration_async(self, eps, lr, shift_kind=SHIFT_ADD, refine=0):
        self._chk(self.lib.nqh_sr_iteration_async(self.h, eps, shift_kind, refine, lr))

    def wait(self):
        st = np.empty(5)
        self._chk(self.lib.nqh_wait(self.h, _p(st, _dp)))
        return st

    def get_samples(self):
        keys = np.empty(self.chains * self.chain_length, dtype=np.uint64)
        self._chk(self.lib.nqh_get_samples(self.h, _p(keys, _u64p), len(keys)))
        return keys

    def set_sr_solver(self, solver=SOLVER_CHOLESKY, tol=1e-4, maxiter=0):
        self._chk(self.lib.nqh_set_sr_solver(self.h, solver, tol, maxiter))

    def sr_solver_info(self):
        it, res = C.c_int32(0), C.c_double(0.0)
        self._chk(self.lib.nqh_get_sr_solver_info(self.h, C.byref(it), C.byref(res)))
        return int(it.value), float(res.value)

    def set_graph(self, on=True):
        self._chk(self.lib.nqh_set_graph(self.h, 1 if on else 0))

    def set_launch_shape(self, est_samples_per_block=0, est_waves=0, gram_splits=0, sampler_waves=0):
        """Launch shapes (0 = chosen from the problem); see nqh_set_launch_shape."""
        self._chk(self.lib.nqh_set_launch_shape(self.h, int(est_samples_per_block), int(est_waves), int(gram_splits),
                                                int(sampler_waves)))

    def get_launch_shape(self):
        out = np.zeros(4, dtype=np.int32)
        self._chk(self.lib.nqh_get_launch_shape(self.h, _p(out, _i32p)))
        return [int(v) for v in out]

    def set_profiling(self, on=True):
        self._chk(self.lib.nqh_set_profiling(self.h, 1 if on else 0))

    def phase_times(self):
        ms = np.zeros(NUM_PHASES)
        n = np.zeros(NUM_PHASES, dtype=np.int64)
        self._chk(self.lib.nqh_get_phase_times(self.h, _p(ms, _dp), _p(n, _i64p)))
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(PHASES)}

    def get_local_values(self):
        n = self.chains * self.chain_length
        L = np.empty(2 * n)
        self._chk(self.lib.nqh_get_local_values(self.h, _p(L, _dp), n))
        return L.view(np.complex128)

    # -- learned proposal (MADE over the 2N doubled bits) --------------------
    def made_setup(self, mix=0.5, retrain_every=20, train_steps=50, batch=1024, lr=1e-3, buffer_cap=65536):
        self._chk(self.lib.nqh_made_setup(self.h, mix, retrain_every, train_steps, batch, lr, buffer_cap))
        D = self.N if getattr(self, "ket", False) else 2 * self.N
        self.made_np = int(self.lib.nqh_made_num_params(D))

    def made_set_params(self, th):
        th = np.ascontiguousarray(th, dtype=np.float32)
        self._chk(self.lib.nqh_made_set_params(self.h, _p(th, _fp), len(th)))

    def made_get_params(self):
        th = np.empty(self.made_np, dtype=np.float32)
        self._chk(self.lib.nqh_made_get_params(self.h, _p(th, _fp), len(th)))
        return th

    def made_grad(self, keys):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        g = np.empty(self.made_np, dtype=np.float32)
        nll = C.c_double(0.0)
        self._chk(self.lib.nqh_made_grad(self.h, _p(keys, _u64p), len(keys), _p(g, _fp), C.byref(nll)))
        return float(nll.value), g

    def made_train(self, keys, steps, batch, lr):
        """Adam steps on `keys` (None: on the context's sample buffer)."""
        k = None if keys is None else np.ascontiguousarray(keys, dtype=np.uint64)
        nll = C.c_double(0.0)
        self._chk(self.lib.nqh_made_train(self.h, _p(k, _u64p), 0 if k is None else len(k), steps, batch, lr,
                                          C.byref(nll)))
        return float(nll.value)

    def made_logq(self, keys):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        lq = np.empty(len(keys), dtype=np.float32)
        self._chk(self.lib.nqh_made_logq(self.h, _p(keys, _u64p), len(keys), _p(lq, _fp)))
        return lq

    def made_sample(self, n, it=0, tag=0):
        keys = np.empty(n, dtype=np.uint64)
        lq = np.empty(n, dtype=np.float32)
        self._chk(self.lib.nqh_made_sample(self.h, n, it, tag, _p(keys, _u64p), _p(lq, _fp)))
        return keys, lq

    def made_stats(self):
        a, b, c = C.c_int64(0), C.c_int64(0), C.c_int64(0)
        self._chk(self.lib.nqh_made_stats(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return int(a.value), int(b.value), int(c.value)


class Batch:
    """Owns one nqh_batch: the SR iterations of several contexts (independent
    problems on one device) captured as one graph, launched once per
    iteration.  The member contexts must stay configured and alive while the
    batch exists; close the batch first."""

    def __init__(self, members):
        self.lib = load()
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m.h.value for m in self.members])
        h = C.c_void_p()
        st = self.lib.nqh_batch_create(C.byref(h), arr, len(self.members))
        if st != NQH_OK:
            raise NQHError(st, "nqh_batch_create failed (members on different devices, duplicated or null)")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.nqh_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_mode(self, mode):
        """BATCH_PACKED (default: one launch per kernel for all qualifying
        members) or BATCH_GRAPH (every member's own launches in the graph)."""
        r = self.lib.nqh_batch_set_mode(self.h, mode)
        if r != NQH_OK:
            raise NQHError(r, self.lib.nqh_batch_last_error(self.h).decode())

    def launches(self):
        """Kernel launches per iteration of the packed graph (0: per-member
        graph, -1: nothing captured yet)."""
        return int(self.lib.nqh_batch_launches(self.h))

    def sr_iteration(self, eps, lr, shift_kind=SHIFT_ADD, refine=0):
        """One SR iteration of every member; returns an (n, 5) array of statistics."""
        st = np.empty((len(self.members), 5))
        r = self.lib.nqh_batch_sr_iteration(self.h, eps, shift_kind, refine, lr, _p(st, _dp))
        if r != NQH_OK:
            raise NQHError(r, self.lib.nqh_batch_last_error(self.h).decode())
        return st
