"""Edge cases at the C ABI: empty batches, the widest configurations (N = 16,
nh = na = 64), argument and call-order errors reported as status codes with a
message, never as a fault."""
import numpy as np
import pytest

from _helpers import ising, params, random_keys, rel_err
from nqhip import Context, NQHError, _lib

pytestmark = pytest.mark.gpu


def ndm_ctx(N, alpha=1, seed=1):
    flat, tab, _ = ising(N)
    ctx = Context(0, seed)
    ctx.set_problem(flat)
    ctx.set_ndm(N, alpha * N, alpha * N)
    return ctx, tab


def test_empty_batches(gpu_ok):
    ctx, _ = ndm_ctx(4)
    ctx.set_params(params(ctx.np_, seed=2))
    none = np.zeros(0, dtype=np.uint64)
    assert ctx.logpsi(none).shape == (0,)
    lp, O = ctx.logpsi_grad(none)
    assert lp.shape == (0,)
    assert ctx.local_estimator(none).shape == (0,)
    cnt, ko, mo = ctx.connections(none, 8)
    assert cnt.shape == (0,)
    ctx.close()


def test_widest_keys_match_the_oracle(gpu_ok, ref):
    """N = 16, the widest NDM: sigma in bits 0..15 and eta in bits 16..31,
    all-ones and top-bit keys included; nh = na = 64, the widest layers."""
    N = 16
    ctx, tab = ndm_ctx(N, alpha=4)
    p = params(ctx.np_, seed=3, sigma=0.02)
    ctx.set_params(p)
    keys = random_keys(N, 16, seed=4)
    keys[0] = np.uint64(2 ** 32 - 1)
    keys[1] = np.uint64(2 ** 31)
    keys[2] = np.uint64(0)
    lp = ctx.logpsi(keys)
    lp_r = ref.logpsi(N, 4 * N, 4 * N, p, keys)
    assert rel_err(lp.real, lp_r.real) < 1e-10
    assert np.max(np.abs(np.angle(np.exp(1j * (lp.imag - lp_r.imag))))) < 1e-10
    L = ctx.local_estimator(keys)
    L_r = ref.local_estimator(tab, 4 * N, 4 * N, p, keys, grad=False)
    assert rel_err(L, L_r) < 1e-9
    ctx.close()


def test_argument_errors(gpu_ok):
    ctx = Context(0, 1)
    with pytest.raises(NQHError, match="1..16"):
        ctx.set_ndm(17, 4, 4)
    with pytest.raises(NQHError, match="1..64"):
        ctx.set_ndm(4, 0, 4)
    with pytest.raises(NQHError, match="1..64"):
        ctx.set_ndm(4, 4, 65)
    with pytest.raises(NQHError, match="1..16"):
        ctx.set_rbm(17, 4)
    ctx.close()
    ctx, _ = ndm_ctx(4)
    with pytest.raises(NQHError, match="parameter count"):
        ctx.set_params(np.zeros(ctx.np_ + 1))
    with pytest.raises(NQHError, match="unknown solver"):
        ctx.set_sr_solver(99, 1e-4, 0)
    with pytest.raises(NQHError, match="tol"):
        ctx.set_sr_solver(_lib.SOLVER_CG, 0.0, 0)
    with pytest.raises(NQHError, match="maxiter"):
        ctx.set_sr_solver(_lib.SOLVER_MINRES, 1e-4, -1)
    ctx.close()


def test_call_order_errors(gpu_ok):
    ctx, _ = ndm_ctx(4)
    ctx.set_params(params(ctx.np_, seed=5))
    with pytest.raises(NQHError, match="no gradient"):
        ctx.precondition(0.01)
    keys = random_keys(4, 16 * 4, seed=6)
    ctx.grad_from_samples(keys, 16, 4)
    with pytest.raises(NQHError, match="shift kind"):
        ctx.precondition(0.01, 7)
    d = ctx.precondition(0.01)
    assert np.all(np.isfinite(d))
    ctx.close()


def test_non_finite_parameters_are_reported(gpu_ok):
    ctx, _ = ndm_ctx(4)
    p = params(ctx.np_, seed=7)
    p[3] = np.nan
    with pytest.raises(NQHError, match="non-finite parameter"):
        ctx.set_params(p)
    ctx.close()
