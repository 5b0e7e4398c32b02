"""Matrix-free SR (conjugate gradients or MINRES on S v = Re(Oc^H Oc v)/Ns +
shift, S never formed) against the Cholesky solve of the same S and F, which
is itself pinned to LAPACK elsewhere, and MINRES against the MINRES oracle run
on the explicit S.  The solvers stop at ||r|| <= 1e-12 ||F||, so the
solutions agree to about cond(S + shift) x 1e-12 (a few 1e-9 with the
eps diag(S) shift at N = 6)."""
import numpy as np
import pytest

from _helpers import ising, params, random_keys, rel_err
from nqhip import Context, NQHError, _lib
from oracle.minres_ref import minres as minres_oracle

pytestmark = pytest.mark.gpu


def ndm(N, alpha, seed, prec=_lib.PREC_FP64, solver=_lib.SOLVER_CHOLESKY, tol=1e-4, maxiter=0):
    flat, _, _ = ising(N)
    ctx = Context(0, 1)
    ctx.set_problem(flat)
    ctx.set_ndm(N, alpha * N, alpha * N)
    ctx.set_params(params(ctx.np_, seed=seed, sigma=0.1))
    ctx.set_precision(prec, _lib.PREC_FP64)     # the Cholesky reference solves an FP64-built S
    ctx.set_sr_solver(solver, tol, maxiter)
    return ctx


@pytest.mark.parametrize("solver", [_lib.SOLVER_CG, _lib.SOLVER_MINRES])
@pytest.mark.parametrize("kind", [_lib.SHIFT_ADD, _lib.SHIFT_DIAG])
@pytest.mark.parametrize("N,alpha,prec,tol", [(6, 1, _lib.PREC_FP64, 1e-8), (12, 2, _lib.PREC_FP32, 1e-6)])
def test_cg_solution_matches_cholesky(gpu_ok, N, alpha, prec, tol, kind, solver):
    keys = random_keys(N, 128 * 8, seed=31)
    a = ndm(N, alpha, 4, prec)
    b = ndm(N, alpha, 4, prec, solver, 1e-12)
    a.grad_from_samples(keys, 128, 8, want_S=False)
    b.grad_from_samples(keys, 128, 8, want_S=False)
    da = a.precondition(0.01, kind)
    db = b.precondition(0.01, kind)
    it, res = b.sr_solver_info()
    assert 0 < it <= 10 * b.np_ and res <= 1e-12
    assert rel_err(db, da) < tol
    a.close()
    b.close()


@pytest.mark.parametrize("solver", [_lib.SOLVER_CG, _lib.SOLVER_MINRES])
def test_cg_sr_loop_tracks_cholesky_loop(gpu_ok, solver):
    N = 6
    a = ndm(N, 1, 5)
    b = ndm(N, 1, 5, solver=solver, tol=1e-12)
    for c in (a, b):
        c.set_sampler(64, 8, 13, 10)
        c.set_iteration(0)
    for _ in range(3):
        sa = a.sr_iteration(0.01, 0.01)
        sb = b.sr_iteration(0.01, 0.01)
        assert rel_err(sb[:2], sa[:2]) < 1e-8
    assert rel_err(b.get_params(), a.get_params()) < 1e-9
    a.close()
    b.close()


@pytest.mark.parametrize("solver", [_lib.SOLVER_CG, _lib.SOLVER_MINRES])
def test_cg_without_convergence_gives_zero_step(gpu_ok, solver):
    N = 6
    keys = random_keys(N, 64 * 8, seed=32)
    b = ndm(N, 1, 6, solver=solver, tol=1e-14, maxiter=2)
    b.grad_from_samples(keys, 64, 8, want_S=False)
    d = b.precondition(0.01)
    it, res = b.sr_solver_info()
    assert it == 2 and res > 1e-14
    assert not np.any(d)
    b.close()


def test_cg_does_not_form_s(gpu_ok):
    N = 6
    keys = random_keys(N, 64 * 8, seed=33)
    b = ndm(N, 1, 7, solver=_lib.SOLVER_CG, tol=1e-8)
    with pytest.raises(NQHError, match="not formed"):
        b.grad_from_samples(keys, 64, 8, want_S=True)
    b.close()


@pytest.mark.parametrize("kind", [_lib.SHIFT_ADD, _lib.SHIFT_DIAG])
def test_minres_follows_the_oracle(gpu_ok, kind):
    """The device MINRES against the oracle's MINRES on the explicit S + shift
    and F of the same samples: the same number of products (one more or less
    where the stopping test sits on a rounding edge) and, at equal counts,
    the same iterate."""
    N, eps = 6, 0.01
    keys = random_keys(N, 128 * 8, seed=34)
    a = ndm(N, 1, 8)
    b = ndm(N, 1, 8, solver=_lib.SOLVER_MINRES, tol=1e-6)
    _, S, F, _ = a.grad_from_samples(keys, 128, 8)
    b.grad_from_samples(keys, 128, 8, want_S=False)
    x = b.precondition(eps, kind)
    it, res = b.sr_solver_info()
    A = S + (eps * np.eye(len(F)) if kind == _lib.SHIFT_ADD else eps * np.diag(np.diag(S)))
    xr, itr, rr, ok = minres_oracle(A, F, 1e-6, 10 * len(F))
    assert ok and 0 < it and res <= 1e-6
    assert abs(it - itr) <= 1
    if it == itr:
        assert rel_err(x, xr) < 1e-8
    assert np.linalg.norm(F - A @ x) <= 1.01e-6 * np.linalg.norm(F)
    a.close()
    b.close()
