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
from oracle.minresqlp_ref import minresqlp as minresqlp_oracle
from oracle.lsqr_ref import lsqr as lsqr_oracle

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


@pytest.mark.parametrize("solver", [_lib.SOLVER_CG, _lib.SOLVER_MINRES, _lib.SOLVER_MINRES_QLP])
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


def perturbed(A, seed):
    """A + a symmetric perturbation at the level of the rounding of the
    device's products (1e-15 of the largest entry)."""
    E = np.random.default_rng(seed).standard_normal(A.shape) * 1e-15 * np.abs(A).max()
    return A + 0.5 * (E + E.T)


@pytest.mark.parametrize("kind", [_lib.SHIFT_ADD, _lib.SHIFT_DIAG])
def test_minres_follows_the_oracle(gpu_ok, kind):
    """The device MINRES against the oracle's MINRES on the explicit S + shift
    and F of the same samples.  Lanczos runs part by rounding once
    orthogonality is lost (the eps diag(S) shift needs ~Np products here), so
    the product count is held to the spread the oracle itself shows under a
    rounding-level perturbation of S (at least +-1), and the iterate is
    compared with the oracle's after the same number of products, within 10x
    that spread."""
    N, eps, tol = 6, 0.01, 1e-6
    keys = random_keys(N, 128 * 8, seed=34)
    A, F = explicit_system(N, 8, kind, eps, keys)
    b = ndm(N, 1, 8, solver=_lib.SOLVER_MINRES, tol=tol)
    b.grad_from_samples(keys, 128, 8, want_S=False)
    x = b.precondition(eps, kind)
    it, res = b.sr_solver_info()
    n = len(F)
    xr, itr, rr, ok = minres_oracle(A, F, tol, 10 * n)
    assert ok and 0 < it and res <= tol
    counts = [itr] + [minres_oracle(perturbed(A, s), F, tol, 10 * n)[1] for s in range(4)]
    assert abs(it - itr) <= max(1, 2 * (max(counts) - min(counts)))
    xk = minres_oracle(A, F, 0.0, it)[0]
    spread = max(rel_err(minres_oracle(perturbed(A, s), F, 0.0, it)[0], xk) for s in range(3))
    assert rel_err(x, xk) <= 10 * spread + 1e-12
    assert np.linalg.norm(F - A @ x) <= 1.01 * tol * np.linalg.norm(F)
    b.close()


def explicit_system(N, seed, kind, eps, keys):
    a = ndm(N, 1, seed)
    _, S, F, _ = a.grad_from_samples(keys, 128, 8)
    a.close()
    return S + (eps * np.eye(len(F)) if kind == _lib.SHIFT_ADD else eps * np.diag(np.diag(S))), F


@pytest.mark.parametrize("kind", [_lib.SHIFT_ADD, _lib.SHIFT_DIAG])
def test_minres_qlp_follows_the_reference_file(gpu_ok, kind):
    """sr_qlp against the oracle's restatement of minresqlp.jl (its QLP-branch
    updates, rotations, norm estimates and stopping flags) on the explicit
    S + shift, with the product count and the iterate held to the oracle's
    own spread under a rounding-level perturbation of S, as for MINRES.  The test's relres is relative to
    ||A|| ||x|| + ||F||, so the solution is held to that bound, not to the
    Cholesky solve."""
    N, eps, tol = 6, 0.01, 1e-6
    keys = random_keys(N, 128 * 8, seed=35)
    A, F = explicit_system(N, 9, kind, eps, keys)
    b = ndm(N, 1, 9, solver=_lib.SOLVER_MINRES_QLP, tol=tol)
    b.grad_from_samples(keys, 128, 8, want_S=False)
    x = b.precondition(eps, kind)
    it, relres = b.sr_solver_info()
    n = len(F)
    xr, sc = minresqlp_oracle(A, F, tol, 10 * n)
    assert sc.flag == 1 and 0 < it and relres <= tol
    counts = [sc.iters] + [minresqlp_oracle(perturbed(A, s), F, tol, 10 * n)[1].iters for s in range(4)]
    assert abs(it - sc.iters) <= max(1, 2 * (max(counts) - min(counts)))
    xk, _ = minresqlp_oracle(A, F, 0.0, it)
    spread = max(rel_err(minresqlp_oracle(perturbed(A, s), F, 0.0, it)[0], xk) for s in range(3))
    assert rel_err(x, xk) <= 10 * spread + 1e-12
    # relres bounds the true residual through the estimates of ||A|| and ||x||
    assert np.linalg.norm(F - A @ x) <= 1.01 * tol * (sc.anorm * np.linalg.norm(x) + np.linalg.norm(F))
    b.close()


def test_minres_qlp_maxiter_keeps_the_iterate(gpu_ok):
    """As minresqlp.jl is written, reaching maxiter (flag 8) counts as
    converged: the step is the last iterate, not zero."""
    N, eps = 6, 0.01
    keys = random_keys(N, 128 * 8, seed=36)
    A, F = explicit_system(N, 10, _lib.SHIFT_ADD, eps, keys)
    b = ndm(N, 1, 10, solver=_lib.SOLVER_MINRES_QLP, tol=1e-14, maxiter=3)
    b.grad_from_samples(keys, 128, 8, want_S=False)
    x = b.precondition(eps)
    it, _ = b.sr_solver_info()
    xr, sc = minresqlp_oracle(A, F, 1e-14, 3)
    assert it == sc.iters == 3 and sc.flag == 8
    assert np.any(x) and rel_err(x, xr) < 1e-10
    b.close()


@pytest.mark.parametrize("kind", [_lib.SHIFT_ADD, _lib.SHIFT_DIAG])
def test_lsqr_follows_the_oracle(gpu_ok, kind):
    """sr_lsq against the LSQR oracle on the explicit S + shift with the
    reference's default tolerances: the same stopping reason's outcome
    (converged), a step count within the oracle's own spread under a
    rounding-level perturbation of S (LSQR works on the normal equations,
    so runs part by rounding early), and the iterate after the same number of
    steps within 10x that spread."""
    N, eps = 6, 0.01
    keys = random_keys(N, 128 * 8, seed=37)
    A, F = explicit_system(N, 11, kind, eps, keys)
    b = ndm(N, 1, 11, solver=_lib.SOLVER_LSQR, tol=1e-4)
    b.grad_from_samples(keys, 128, 8, want_S=False)
    x = b.precondition(eps, kind)
    it, res = b.sr_solver_info()
    n = len(F)
    xr, itr, istop, ok = lsqr_oracle(A, F, 10 * n)
    assert ok and np.any(x)
    counts = [itr] + [lsqr_oracle(perturbed(A, s), F, 10 * n)[1] for s in range(4)]
    assert abs(it - itr) <= max(2, 2 * (max(counts) - min(counts)))
    xk = lsqr_oracle(A, F, it, atol=0.0, btol=0.0, conlim=0.0)[0]
    spread = max(rel_err(lsqr_oracle(perturbed(A, s), F, it, atol=0.0, btol=0.0, conlim=0.0)[0], xk)
                 for s in range(3))
    assert rel_err(x, xk) <= 10 * spread + 1e-10
    b.close()
