"""Host-side mirror of the reference's steady-state VMC API on libnqhip.

Julia (reference)                                   here
-------------------------------------------------   ---------------------------------------------
NDM(Float64, N, αh, αa, af_softplus)  NDM.jl:50     NDM(N, alpha_h, alpha_a, act="softplus")
  (af_logcosh, activation.jl:5-10)                    NDM(N, alpha_h, alpha_a, act="logcosh")
RBM(ComplexF64, N, α)                  RBM.jl:37-44  RBM(N, alpha)   (kets, with a Hamiltonian)
RBMSplit(ComplexF64, N, α)             RBMSplit.jl:33-40  RBMSplit(N, alpha)   (density matrices)
init_random_pars!(net; sigma)          rng.jl:20    net.init_random_pars(rng, sigma)
MetropolisSampler(LocalRule(), L, passes; burn,     MetropolisSampler(LocalRule(), L, passes,
                  seed)                Metropolis.jl:26                  burn=, seed=)
SR(; ϵ, algorithm=sr_cholesky,                       SR(eps=, precondition_type="shift")
   precondition_type=sr_shift)         SR.jl:50-57
Optimisers.Descent(η)                  rules.jl:11   Descent(eta)
BatchedSampler(net, sampl, liouv, algo;              BatchedSampler(net, sampl, liouv, algo,
               batch_sz)               build_Batched.jl:1-14        batch_sz=)
sample!(is) -> (L_stat, prec)          BatchedGradSampler.jl:77     sample_(is)
precondition!(prec, algo, i) -> grad   SRDirect.jl:51-90            precondition_(prec, algo, i)
Optimisers.update!(opt, net, grad)     apply.jl:9                   update_(opt, net, grad)
add_observable!(is, name, op)          BatchedGradSampler.jl:129    add_observable_(is, name, op)
compute_observables(is)                BatchedObsDMSampler.jl:59    compute_observables(is)

The loss statistics come back as `Measurement` (stats.jl:1-13).  Device state
(samples, O, S, the factor, the update) stays on the GPU; `net.flat` is the
canonical host copy: it is uploaded at every sample_ (so host-side changes
to the network take effect, as in the reference) and refreshed after every
update_.  precondition_ returns a `Gradient` (the step, reference order,
owned by the cache and overwritten by the next call as SRDirect.jl:89 does)
that knows its sampler, so update_(opt, net, grad) takes the reference's
arguments.
"""
from __future__ import annotations

import math
import warnings
from dataclasses import dataclass

import numpy as np

from . import _lib
from .operators import KLocalLiouvillian, KLocalOperator, KLocalOperatorSum, flatten

__all__ = ["NDM", "RBM", "RBMSplit", "LocalRule", "MetropolisSampler", "SR", "Descent", "Measurement", "Gradient",
           "BatchedSampler", "HIPBatchedSampler", "sample_", "precondition_", "update_",
           "init_random_pars_", "add_observable_", "compute_observables"]


class NDM:
    """Neural density matrix parameters in the reference's flat order
    (NDM.jl:17-18): b_mu, h_mu, w_mu, u_mu, b_la, h_la, d_la, w_la, u_la,
    matrices column-major (nh x N, na x N)."""

    FIELDS = ("b_mu", "h_mu", "w_mu", "u_mu", "b_la", "h_la", "d_la", "w_la", "u_la")

    ACTS = {"softplus": _lib.ACT_SOFTPLUS, "logcosh": _lib.ACT_LOGCOSH}

    def __init__(self, N: int, alpha_h: float, alpha_a: float, act: str = "softplus"):
        if act not in self.ACTS:
            raise ValueError("act must be 'softplus' (af_softplus) or 'logcosh' (af_logcosh)")
        self.act = act
        self.N = int(N)
        self.nh = int(alpha_h * N)
        self.na = int(alpha_a * N)
        if self.nh != alpha_h * N or self.na != alpha_a * N:
            raise ValueError("alpha * N must be an integer (convert(Int, αh*in), NDM.jl:55)")
        N, nh, na = self.N, self.nh, self.na
        sizes = [N, nh, nh * N, na * N, N, nh, na, nh * N, na * N]
        shapes = [(N,), (nh,), (nh, N), (na, N), (N,), (nh,), (na,), (nh, N), (na, N)]
        self.flat = np.zeros(sum(sizes))
        off = 0
        for name, n, shp in zip(self.FIELDS, sizes, shapes):
            setattr(self, name, self.flat[off:off + n].reshape(shp, order="F"))
            off += n

    @property
    def n_params(self) -> int:
        return len(self.flat)

    def init_random_pars(self, rng=None, sigma: float = 0.01):
        """`init_random_pars!(net; sigma)`: randn * sqrt(sigma) (rng.jl:20-31)."""
        rng = rng if rng is not None else np.random.default_rng()
        self.flat[:] = rng.standard_normal(self.n_params) * math.sqrt(sigma)
        return self


class RBM:
    """`RBM(ComplexF64, N, α)` on kets (RBM.jl:37-44, af_softplus): complex
    a[N], b[nh], W[nh x N] in the flat order a, b, W (W column-major),
    log psi(s) = a.s + sum softplus(b + W s)."""

    def __init__(self, N: int, alpha: float):
        self.N = int(N)
        self.nh = int(alpha * N)
        if self.nh != alpha * N:
            raise ValueError("alpha * N must be an integer (convert(Int, α*in), RBM.jl:44)")
        N, nh = self.N, self.nh
        self.flat = np.zeros(N + nh + nh * N, dtype=np.complex128)
        self.a = self.flat[:N]
        self.b = self.flat[N:N + nh]
        self.W = self.flat[N + nh:].reshape((nh, N), order="F")

    @property
    def n_params(self) -> int:
        return len(self.flat)

    def init_random_pars(self, rng=None, sigma: float = 0.01):
        """`init_random_pars!(net; sigma)`: complex randn (unit variance) *
        sqrt(sigma) (rng.jl:20-31)."""
        rng = rng if rng is not None else np.random.default_rng()
        z = rng.standard_normal(self.n_params) + 1j * rng.standard_normal(self.n_params)
        self.flat[:] = z * math.sqrt(sigma / 2)
        return self


class RBMSplit(RBM):
    """`RBMSplit(ComplexF64, N, α)` (RBMSplit.jl:33-40): a density matrix
    rho(s, e) from complex ar[N], ac[N], b[nh], Wr[nh x N], Wc[nh x N] in that
    flat order (matrices column-major)."""

    def __init__(self, N: int, alpha: float):
        self.N = int(N)
        self.nh = int(alpha * N)
        if self.nh != alpha * N:
            raise ValueError("alpha * N must be an integer (convert(Int, α*in), RBMSplit.jl:40)")
        N, nh = self.N, self.nh
        self.flat = np.zeros(2 * N + nh + 2 * nh * N, dtype=np.complex128)
        self.ar, self.ac = self.flat[:N], self.flat[N:2 * N]
        self.b = self.flat[2 * N:2 * N + nh]
        o = 2 * N + nh
        self.Wr = self.flat[o:o + nh * N].reshape((nh, N), order="F")
        self.Wc = self.flat[o + nh * N:].reshape((nh, N), order="F")


def init_random_pars_(net: NDM, rng=None, sigma: float = 0.01):
    return net.init_random_pars(rng, sigma)


class LocalRule:
    """Single-site flip proposal (LocalRule.jl:19-28)."""


class MetropolisSampler:
    """`MetropolisSampler(rule, chain_length, passes; burn=0, seed)`
    (Metropolis.jl:26-40).  An even `passes` is bumped to the next odd number,
    as the reference does."""

    def __init__(self, rule, chain_length: int, passes: int, burn: int = 0, seed: int | None = None,
                 compat_stale_last_slice: bool = False, persistent: bool = False):
        if not isinstance(rule, LocalRule):
            raise NotImplementedError("only LocalRule is implemented")
        if passes <= 0 or chain_length <= 0:
            raise ValueError("passes and chain_length must be positive")
        if passes % 2 == 0:
            warnings.warn(f"LocalRule sampling with an even number of switches per step is known to be "
                          f"non-ergodic; setting the number of switches to {passes + 1}.")
            passes += 1
        self.rule = rule
        self.chain_length = int(chain_length)
        self.passes = int(passes)
        self.burn = int(burn)
        self.seed = int(seed) if seed is not None else int(np.random.default_rng().integers(0, 2 ** 63))
        self.compat_stale_last_slice = compat_stale_last_slice
        self.persistent = persistent

    @property
    def flags(self):
        f = 0
        if self.compat_stale_last_slice:
            f |= _lib.SAMPLER_COMPAT_STALE_LAST_SLICE
        if self.persistent:
            f |= _lib.SAMPLER_PERSISTENT
        return f


class SR:
    """`SR(; ϵ=0.001, algorithm=sr_cholesky, precondition_type=sr_shift,
    precision=1e-4)` (SR.jl:37-57).  `algorithm` "cholesky", "diag" or "div"
    (SRDirect.jl; all three by the Cholesky factorization of the shifted S),
    "cg", "minres", "qlp" or "lsq" (SRIterative.jl with the matrix-free S of
    SR_notfull.jl: S is never formed; `precision` is the solver's tolerance,
    unused by "lsq" as in the reference; `maxiter` 0 means the reference's
    10 Np).  `precondition_type`
    (SR.jl:4, SRDirect.jl:56-70): "none" solves S itself; "shift" adds eps*I;
    "multiplicative" adds lambda*diag(S) with lambda = max(lambda0 * b^n,
    lambda_min) at iteration n (the reference's Cholesky path drops this
    shift, SRDirect.jl:67-78; here every solver applies it, unless `as_run`
    is set: then "multiplicative" with "cholesky" or "div" factors the
    unshifted S as the reference does as written); "diag" adds a fixed
    eps*diag(S).  "diag" (pinv of the preconditioned S) with "none" is
    refused: on a singular S the pseudo-inverse step differs from any
    factorization.  `refine` extra FP64 refinement steps of the Cholesky
    solve."""

    def __init__(self, eps: float = 0.001, algorithm: str = "cholesky", precondition_type: str = "shift",
                 refine: int = 0, precision: float = 1e-4, maxiter: int = 0, lambda0: float = 100.0,
                 b: float = 0.95, lambda_min: float = 1e-4, as_run: bool = False):
        if algorithm not in ("cholesky", "diag", "div", "cg", "minres", "qlp", "lsq"):
            # sr_pivcholesky is declared by the reference but throws "Alg not known" (SRDirect.jl:85-86)
            raise NotImplementedError("algorithm must be 'cholesky' (sr_cholesky), 'diag' (sr_diag), 'div' "
                                      "(sr_div), 'cg' (sr_cg), 'minres' (sr_minres), 'qlp' (sr_qlp) or 'lsq' (sr_lsq)")
        if precondition_type not in ("none", "shift", "multiplicative", "diag"):
            raise ValueError("precondition_type must be 'none', 'shift', 'multiplicative' or 'diag'")
        if algorithm == "diag" and precondition_type == "none":
            raise NotImplementedError("sr_diag with sr_none is the pseudo-inverse of the unshifted, possibly "
                                      "singular S (SRDirect.jl:73-75): not provided; use a shift")
        self.as_run = bool(as_run)
        self.eps = float(eps)
        self.algorithm = algorithm
        self.precondition_type = precondition_type
        self.refine = int(refine)
        self.precision = float(precision)
        self.maxiter = int(maxiter)
        self.lambda0, self.b, self.lambda_min = float(lambda0), float(b), float(lambda_min)

    def shift(self, it: int = 0):
        """(eps, shift kind) of iteration `it` (SRDirect.jl:56-70)."""
        if self.precondition_type == "none":
            return 0.0, _lib.SHIFT_ADD
        if self.precondition_type == "shift":
            return self.eps, _lib.SHIFT_ADD
        if self.precondition_type == "multiplicative":
            if self.as_run and self.algorithm in ("cholesky", "div"):
                # defect D4 as run: cholesky!(Hermitian(S)) / ldiv!(Hermitian(S)) take S, not Sprecond
                return 0.0, _lib.SHIFT_ADD
            return max(self.lambda0 * self.b ** it, self.lambda_min), _lib.SHIFT_DIAG
        return self.eps, _lib.SHIFT_DIAG

    @property
    def solver(self):
        # sr_diag (pseudo-inverse) and sr_div (Bunch-Kaufman) solve the same shifted,
        # positive definite S: the Cholesky factorization gives their step
        return {"cholesky": _lib.SOLVER_CHOLESKY, "diag": _lib.SOLVER_CHOLESKY, "div": _lib.SOLVER_CHOLESKY,
                "cg": _lib.SOLVER_CG, "minres": _lib.SOLVER_MINRES,
                "qlp": _lib.SOLVER_MINRES_QLP, "lsq": _lib.SOLVER_LSQR}[self.algorithm]

    @property
    def shift_kind(self):
        return self.shift(0)[1]


class Gradient(np.ndarray):
    """The preconditioned gradient precondition_ returns (reference order):
    an array that also knows the sampler whose device state holds it, so
    update_(opt, net, grad) applies it on the device (apply.jl:9)."""

    def __new__(cls, values, owner):
        obj = np.asarray(values).view(cls)
        obj.owner = owner
        return obj

    def __array_finalize__(self, obj):
        self.owner = getattr(obj, "owner", None)


class Descent:
    """`Optimisers.Descent(η)`: p -= η Δ (rules.jl:11-17)."""

    def __init__(self, eta: float):
        self.eta = float(eta)


@dataclass
class Measurement:
    """`Measurement` of stats.jl:1-13."""
    mean: float | complex
    error: float
    variance: float
    tau: float
    R: float

    @classmethod
    def from_array(cls, st):
        return cls(*(float(x) for x in st[:5]))

    @classmethod
    def from_observable(cls, st):
        """[Re mean, error, variance, tau, R, Im mean] of a complex observable."""
        return cls(complex(st[0], st[5]), *(float(x) for x in st[1:5]))

    def __str__(self):
        return (f"({self.mean:6.4f}) ± {self.error:6.4f} [var={self.variance:6.4f}, "
                f"tau={self.tau:6.4f}, R={self.R:6.4f}]")


_PREC = {"fp64": (_lib.PREC_FP64, _lib.PREC_FP64),
         "fp32": (_lib.PREC_FP32, _lib.PREC_FP32),
         "mixed": (_lib.PREC_FP32, _lib.PREC_BF16)}


class HIPBatchedSampler:
    """`BatchedGradSampler` (BatchedGradSampler.jl:29-75: an NDM and a
    KLocalLiouvillian, the L†L problem) or `BatchedValSampler`
    (BatchedValSampler.jl:70-139: a ket RBM and a Hamiltonian operator, the
    energy problem), as build_Batched.jl:1-14 dispatches, on libnqhip."""

    def __init__(self, net, sampler: MetropolisSampler, prob, algo: SR,
                 batch_sz: int = 16, device: int = 0, precision: str = "fp64"):
        ket = isinstance(prob, (KLocalOperator, KLocalOperatorSum))
        if isinstance(prob, KLocalLiouvillian):
            if not isinstance(net, (NDM, RBMSplit)):
                raise NotImplementedError("the L†L problem takes an NDM or an RBMSplit")
            n_sites = prob.phys.n_sites
        elif ket:
            if not isinstance(net, RBM) or isinstance(net, RBMSplit):
                raise NotImplementedError("a Hamiltonian takes a ket RBM")
            n_sites = prob.hilb.n_sites if hasattr(prob, "hilb") else net.N
        else:
            raise NotImplementedError("prob must be a KLocalLiouvillian or a Hamiltonian operator")
        if n_sites != net.N:
            raise ValueError("network and problem have different numbers of sites")
        if precision not in _PREC:
            raise ValueError(f"precision must be one of {sorted(_PREC)}")
        self.net, self.sampler, self.prob, self.algo = net, sampler, prob, algo
        self.batch_sz = int(batch_sz)
        self.precision = precision
        self.ctx = _lib.Context(device, sampler.seed)
        self.ctx.set_problem(flatten(prob))
        if ket:
            self.ctx.set_rbm(net.N, net.nh)
        elif isinstance(net, RBMSplit):
            self.ctx.set_rbmsplit(net.N, net.nh)
        else:
            self.ctx.set_ndm(net.N, net.nh, net.na, NDM.ACTS[net.act])
        self.ctx.set_precision(*_PREC[precision])
        self.ctx.set_params(net.flat.view(np.float64))
        self.ctx.set_sampler(self.batch_sz, sampler.chain_length, sampler.passes, sampler.burn, sampler.flags)
        self.ctx.set_sr_solver(algo.solver, algo.precision, algo.maxiter)
        self._pending = False
        self._iter = 0
        self.observables = {}   # name -> flattened ket operator

    @property
    def n_samples(self):
        return self.batch_sz * self.sampler.chain_length

    def push_params(self):
        self.ctx.set_params(self.net.flat.view(np.float64))

    def sample(self):
        self.push_params()         # the host network is canonical
        st = self.ctx.sample_grad()
        self._pending = True
        return Measurement.from_array(st), self

    def precondition(self, algo: SR, it: int = 0):
        if not self._pending:
            raise RuntimeError("call sample_(is) before precondition_")
        eps, kind = algo.shift(it)
        return Gradient(self.ctx.precondition(eps, kind, algo.refine), self)

    def add_observable(self, name: str, op):
        """`add_observable!(is, name, Ô)` (BatchedGradSampler.jl:129-131): a
        ket operator on the physical space, evaluated by compute_observables."""
        self.observables[name] = flatten(op)

    def compute_observables(self):
        """`compute_observables(is)` (BatchedObsDMSampler.jl:59-76): a fresh
        diagonal chain with the gradient sampler's settings, then every added
        observable's statistics; None when none was added."""
        if not self.observables:
            return None
        s = self.sampler
        return {name: Measurement.from_observable(
                    self.ctx.observable(op, self.batch_sz, s.chain_length, s.passes, s.burn))
                for name, op in self.observables.items()}

    def update(self, opt: Descent, grad=None):
        self.ctx.update(opt.eta)
        self.net.flat[:] = self.ctx.get_params().view(self.net.flat.dtype)
        self._pending = False

    def step(self, opt: Descent):
        """One fused SR iteration (sample!, precondition!, update!) without
        host round trips; returns the loss Measurement."""
        eps, kind = self.algo.shift(self._iter)
        st = self.ctx.sr_iteration(eps, opt.eta, kind, self.algo.refine)
        self._iter += 1
        return Measurement.from_array(st)

    def samples(self):
        return self.ctx.get_samples()

    def local_values(self):
        return self.ctx.get_local_values()

    def close(self):
        self.ctx.close()


def BatchedSampler(net, sampler, prob, algo, batch_sz: int = 16, **kw):  # noqa: N802
    return HIPBatchedSampler(net, sampler, prob, algo, batch_sz=batch_sz, **kw)


def sample_(is_: HIPBatchedSampler):
    """`sample!(is)` -> (L_stat, precond_cache)."""
    return is_.sample()


def precondition_(cache: HIPBatchedSampler, algo: SR, it: int = 0):
    """`precondition!(cache, algo, iter)` -> the preconditioned gradient
    (reference order)."""
    return cache.precondition(algo, it)


def update_(opt: Descent, net, grad=None):
    """`Optimisers.update!(opt, net, grad)` (apply.jl:9): p -= eta * grad,
    applied on the device by the sampler that produced `grad`, and the host
    network refreshed.  The sampler itself may stand in for `net`."""
    if isinstance(net, HIPBatchedSampler):
        net.update(opt, grad)
        return net.net
    owner = getattr(grad, "owner", None)
    if owner is None:
        raise TypeError("grad must come from precondition_ (or pass the HIPBatchedSampler)")
    if owner.net is not net:
        raise ValueError("grad was computed for another network")
    owner.update(opt, grad)
    return net


def add_observable_(is_: HIPBatchedSampler, name: str, op):
    """`add_observable!(is, name, Ô)`."""
    is_.add_observable(name, op)


def compute_observables(is_: HIPBatchedSampler):
    """`compute_observables(is)` -> {name: Measurement} (None without observables)."""
    return is_.compute_observables()
