"""Matrix-free SR (conjugate gradients, MINRES, MINRES-QLP or LSQR on
S v = Re(Oc^H Oc v)/Ns + shift, S never formed) against the Cholesky solve
of the same S and F, which is itself pinned to LAPACK elsewhere, and each
Krylov solver against its oracle run on the explicit S.  The solvers stop at
||r|| <= tol ||F||, so the solutions agree to about cond(S + shift) x tol.
Where Lanczos loses orthogonality (the eps diag(S) shift, cond ~4e5 at
N = 6) rounding-level changes of S or F move the iterates, so those cases
are held to bounds rounding cannot move (residuals, distance to the LAPACK
solve) and to count windows measured with the oracle on the same inputs."""
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


def perturbed(A, F, seed):
    """A and F with perturbations at the level of the rounding of the device's
    products: a symmetric 1e-15 of the largest entry of A, 1e-15 relative on F."""
    r = np.random.default_rng(seed)
    E = r.standard_normal(A.shape) * 1e-15 * np.abs(A).max()
    return A + 0.5 * (E + E.T), F * (1.0 + 1e-15 * r.standard_normal(F.shape))


def error_bound(A, F, tol):
    """||x - A^-1 F|| <= ||A^-1|| ||F - A x|| <= tol ||F|| / lambda_min(A) for any
    x with a relative residual below tol, relative to max |A^-1 F|."""
    x0 = np.linalg.solve(A, F)
    return tol * np.linalg.norm(F) / np.linalg.eigvalsh(A)[0] / np.abs(x0).max(), x0


@pytest.mark.parametrize("kind", [_lib.SHIFT_ADD, _lib.SHIFT_DIAG])
def test_minres_follows_the_oracle(gpu_ok, kind):
    """The device MINRES against the oracle's MINRES on the explicit S + shift
    and F of the same samples.  What rounding cannot move is held exactly:
    the true residual, and the distance to the LAPACK solve, which a relative
    residual below tol bounds by tol ||F|| / lambda_min.  With the additive
    shift (cond ~2e2 here) Lanczos keeps its orthogonality: the oracle takes
    19 products under every 1e-15 perturbation of S and F, so the count must
    match and the iterate agrees within 1e-9 (the perturbations move it by
    1.5e-12).  With eps diag(S) (cond ~4e5) orthogonality is lost and a
    rounding-level change of F alone moves the count by a few products
    (151..154 over eight perturbations, measured with the oracle on these
    inputs) and the iterate by ~1e-2, so the count is held to a window of
    8 products and the iterate to the residual bounds only."""
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
    assert np.linalg.norm(F - A @ x) <= 1.05 * tol * np.linalg.norm(F)
    bound, x0 = error_bound(A, F, tol)
    assert rel_err(x, x0) <= 1.05 * bound
    if kind == _lib.SHIFT_ADD:
        counts = {minres_oracle(*perturbed(A, F, s), tol, 10 * n)[1] for s in range(4)}
        assert counts == {itr} and abs(it - itr) <= 1
        if it == itr:
            assert rel_err(x, xr) <= 1e-9
    else:
        assert abs(it - itr) <= 8
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
    S + shift.  The oracle's product count is the same under every 1e-15
    perturbation of S and F for both shifts here (16 and 41), so the device
    must take it too (within one product).  As minresqlp.jl is written,
    reaching maxiter keeps the iterate, so the device's iterate after 10
    products is compared with the oracle's within 1e-9 (rounding-level
    perturbations move it by < 1e-12).  The test's relres is relative to
    ||A|| ||x|| + ||F||, so the final solution is held to that bound."""
    N, eps, tol = 6, 0.01, 1e-6
    keys = random_keys(N, 128 * 8, seed=35)
    A, F = explicit_system(N, 9, kind, eps, keys)
    n = len(F)
    xr, sc = minresqlp_oracle(A, F, tol, 10 * n)
    counts = {minresqlp_oracle(*perturbed(A, F, s), tol, 10 * n)[1].iters for s in range(4)}
    assert sc.flag == 1 and counts == {sc.iters}
    b = ndm(N, 1, 9, solver=_lib.SOLVER_MINRES_QLP, tol=tol)
    b.grad_from_samples(keys, 128, 8, want_S=False)
    x = b.precondition(eps, kind)
    it, relres = b.sr_solver_info()
    assert 0 < it and relres <= tol and abs(it - sc.iters) <= 1
    # relres bounds the true residual through the estimates of ||A|| and ||x||
    assert np.linalg.norm(F - A @ x) <= 1.01 * tol * (sc.anorm * np.linalg.norm(x) + np.linalg.norm(F))
    b.close()
    b = ndm(N, 1, 9, solver=_lib.SOLVER_MINRES_QLP, tol=1e-14, maxiter=10)
    b.grad_from_samples(keys, 128, 8, want_S=False)
    x10 = b.precondition(eps, kind)
    assert b.sr_solver_info()[0] == 10
    assert rel_err(x10, minresqlp_oracle(A, F, 1e-14, 10)[0]) <= 1e-9
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
    reference's default tolerances (atol = btol = sqrt(eps), conlim =
    1/sqrt(eps)): a converged stopping reason, and a step count and a true
    residual ||F - A x|| within what the oracle itself shows under 1e-15
    perturbations of S and F (LSQR's residual estimate drifts from the true
    residual over ~1000 steps, so the true residual is held to 3x the
    oracle's worst rather than to the stopping rule).  Measured with the
    oracle on these inputs: 36..37 steps with the additive shift, final
    iterates within 1.2e-6 of each other (the device's is held to 1e-4 of the
    oracle's); 979..991 steps with eps diag(S), iterates ~1e-2 apart (count
    window 30, residual only)."""
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
    resid = lambda y: np.linalg.norm(F - A @ y)
    worst = max([resid(xr)] + [resid(lsqr_oracle(*perturbed(A, F, s), 10 * n)[0]) for s in range(4)])
    assert resid(x) <= 3.0 * worst
    if kind == _lib.SHIFT_ADD:
        assert abs(it - itr) <= 2
        assert rel_err(x, xr) <= 1e-4
    else:
        assert abs(it - itr) <= 30
    b.close()
