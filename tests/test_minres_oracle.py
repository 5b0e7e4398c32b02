"""The MINRES oracle against what the mathematics fixes: the carried residual
norm is the true one at every step, the k-th iterate is the minimiser of
||b - A x|| over the k-dimensional Krylov space (scipy's MINRES after the same
number of products), exact termination in n steps, and the linear solve."""
import numpy as np
import pytest
from scipy.sparse.linalg import minres as scipy_minres

from oracle.minres_ref import givens, minres


def spd(n, cond, seed):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return (q * np.geomspace(1.0, cond, n)) @ q.T


def test_givens_convention():
    for f, g in [(3.0, 4.0), (-3.0, 4.0), (4.0, -3.0), (-4.0, -3.0), (0.0, 2.0), (2.0, 0.0)]:
        c, s, r = givens(f, g)
        assert np.allclose([c * f + s * g, -s * f + c * g], [r, 0.0], atol=1e-15)
        assert abs(c * c + s * s - 1.0) < 1e-15
        if abs(f) > abs(g):
            assert c > 0


@pytest.mark.parametrize("indefinite", [False, True])
def test_residual_estimate_is_the_true_residual(indefinite):
    n = 40
    A = spd(n, 1e3, 1)
    if indefinite:
        A = A - 0.5 * np.eye(n)   # MINRES needs symmetry only
    b = np.random.default_rng(2).standard_normal(n)
    hist = []
    minres(A, b, 1e-12, 10 * n, hist)
    assert len(hist) >= 5
    prev = np.inf
    for x, r in hist:
        true = np.linalg.norm(b - A @ x)
        assert abs(r - true) <= 1e-9 * np.linalg.norm(b)
        assert r <= prev * (1 + 1e-12)   # non-increasing
        prev = r


def test_iterates_match_scipy_minres():
    n = 30
    A = spd(n, 1e2, 3)
    b = np.random.default_rng(4).standard_normal(n)
    hist = []
    minres(A, b, 0.0, 12, hist)
    for k in (1, 2, 5, 12):
        xs, _ = scipy_minres(A, b, rtol=0.0, maxiter=k)
        assert np.linalg.norm(hist[k - 1][0] - xs) <= 1e-10 * np.linalg.norm(xs)


def test_exact_termination_and_solution():
    n = 6
    A = spd(n, 10.0, 5)
    b = np.random.default_rng(6).standard_normal(n)
    x, k, r, ok = minres(A, b, 1e-12, 100)
    assert ok and k <= n
    assert np.linalg.norm(x - np.linalg.solve(A, b)) <= 1e-10 * np.linalg.norm(x)
    A = spd(50, 1e4, 7)
    b = np.random.default_rng(8).standard_normal(50)
    x, k, r, ok = minres(A, b, 1e-13, 500)
    assert ok and r <= 1e-13 * np.linalg.norm(b)
    assert np.linalg.norm(x - np.linalg.solve(A, b)) <= 1e-8 * np.linalg.norm(x)


def test_maxiter_and_zero_rhs():
    A = spd(20, 1e3, 9)
    b = np.random.default_rng(10).standard_normal(20)
    x, k, r, ok = minres(A, b, 1e-14, 2)
    assert k == 2 and not ok and r > 1e-14 * np.linalg.norm(b)
    x, k, r, ok = minres(A, np.zeros(20), 1e-6, 10)
    assert k == 0 and ok and not np.any(x)


# ---- MINRES-QLP as the reference's minresqlp.jl runs it ---------------------
from oracle.minresqlp_ref import minresqlp, sym_givens


def test_sym_givens():
    for a, b in [(3.0, 4.0), (-3.0, 4.0), (4.0, -3.0), (-4.0, -3.0), (0.0, -2.0), (2.0, 0.0), (-2.0, 0.0)]:
        c, s, r = sym_givens(a, b)
        assert r >= 0 and abs(c * c + s * s - 1.0) < 1e-15
        assert np.allclose([c * a + s * b, s * a - c * b], [r, 0.0], atol=1e-15)


@pytest.mark.parametrize("cond", [1e2, 1e4])
def test_qlp_iterates_are_minres_iterates(cond):
    n, K = 30, 14
    A = spd(n, cond, 11)
    b = np.random.default_rng(12).standard_normal(n)
    h_ref, h_mr = [], []
    minresqlp(A, b, 0.0, K, "reference", h_ref)
    minresqlp(A, b, 0.0, K, "minres", h_mr)
    assert len(h_ref) == len(h_mr) == K
    for k in range(K):
        xs, _ = scipy_minres(A, b, rtol=0.0, maxiter=k + 1)
        scale = np.linalg.norm(xs)
        # two implementations of the Lanczos process part by rounding once
        # orthogonality is lost (after ~10 steps on this spectrum); the two
        # update branches of one Lanczos run agree to rounding throughout
        if k < 10:
            assert np.linalg.norm(h_ref[k][0] - xs) <= 1e-11 * scale
        assert np.linalg.norm(h_mr[k][0] - h_ref[k][0]) <= 1e-12 * scale
        x, st = h_ref[k]
        assert abs(st["xnorm"] - np.linalg.norm(x)) <= 1e-9 * scale
        assert abs(st["rnorm"] - np.linalg.norm(b - A @ x)) <= 1e-9 * np.linalg.norm(b)
        assert 0 < st["anorm"] <= np.linalg.norm(A, 2) * (1 + 1e-12)


def test_qlp_stopping_rule_and_solution():
    n = 50
    A = spd(n, 1e3, 13)
    b = np.random.default_rng(14).standard_normal(n)
    x, sc = minresqlp(A, b, 1e-10, 10 * n)
    assert sc.flag == 1 and sc.converged()
    assert sc.relres <= 1e-10
    assert abs(sc.relres - np.linalg.norm(b - A @ x) / (sc.anorm * sc.xnorm + np.linalg.norm(b))) < 1e-12
    assert np.linalg.norm(x - np.linalg.solve(A, b)) <= 1e-6 * np.linalg.norm(x)
    # relres is relative to ||A|| ||x|| + ||b||: a looser tolerance stops before ||r|| <= tol ||b||
    hist = []
    x6, sc6 = minresqlp(A, b, 1e-4, 10 * n, history=hist)
    assert sc6.flag == 1 and len(hist) == sc6.iters
    xm, km, _, _ = minres(A, b, 1e-4, 10 * n)
    assert sc6.iters <= km


def test_qlp_maxiter_reports_convergence():
    A = spd(20, 1e3, 15)
    b = np.random.default_rng(16).standard_normal(20)
    x, sc = minresqlp(A, b, 1e-14, 3)
    assert sc.iters == 3 and sc.flag == 8 and sc.converged() and np.any(x)


def test_qlp_eigenvector_rhs():
    A = np.diag(np.geomspace(1.0, 1e2, 10))
    e = np.zeros(10)
    e[3] = 2.0
    x, sc = minresqlp(A, e, 1e-8, 100)
    assert sc.flag == -1 and sc.converged()
    assert np.allclose(x, e / A[3, 3], rtol=1e-15, atol=0)


# ---- LSQR (IterativeSolvers.jl 0.8.1 lsqr, the published Paige-Saunders method)
from scipy.sparse.linalg import lsqr as scipy_lsqr

from oracle.lsqr_ref import SQRT_EPS, lsqr


@pytest.mark.parametrize("cond", [1e2, 1e3])
def test_lsqr_iterates_are_krylov_minimisers(cond):
    """The k-th LSQR iterate minimises ||b - A x|| over K_k(A^T A, A^T b):
    against that minimiser (orthonormal basis with full reorthogonalisation)
    and against scipy's LSQR while the recurrence keeps orthogonality (the
    normal equations have condition cond^2; both drift alike after that)."""
    n, K = 30, 8
    A = spd(n, cond, 18)
    b = np.random.default_rng(19).standard_normal(n)
    hist = []
    lsqr(A, b, K, atol=0.0, btol=0.0, conlim=0.0, history=hist)
    assert len(hist) == K
    M, q = A.T @ A, A.T @ b
    Q = np.zeros((n, 0))
    q = q / np.linalg.norm(q)
    for k in range(K):
        Q = np.column_stack([Q, q])
        z = M @ q
        for _ in range(2):
            z = z - Q @ (Q.T @ z)
        q = z / np.linalg.norm(z)
        xe = Q @ np.linalg.lstsq(A @ Q, b, rcond=None)[0]
        xs = scipy_lsqr(A, b, atol=0.0, btol=0.0, conlim=0.0, iter_lim=k + 1)[0]
        assert np.linalg.norm(hist[k][0] - xe) <= 1e-11 * np.linalg.norm(xe)
        assert np.linalg.norm(hist[k][0] - xs) <= 1e-11 * np.linalg.norm(xs)
        assert abs(hist[k][1]["rnorm"] - np.linalg.norm(b - A @ hist[k][0])) <= 1e-9 * np.linalg.norm(b)


@pytest.mark.parametrize("cond,seed", [(1e2, 20), (1e4, 21), (1e9, 22)])
def test_lsqr_stops_where_scipy_stops(cond, seed):
    """The reference's defaults (atol = btol = sqrt(eps), conlim = 1/sqrt(eps),
    maxiter = 10 n): the same stopping reason as scipy's LSQR, a step count
    within 5 % (rounding moves it by a few steps), and the solution."""
    n = 40
    A = spd(n, cond, seed)
    b = np.random.default_rng(seed + 100).standard_normal(n)
    x, itn, istop, ok = lsqr(A, b, 10 * n)
    xs, istop_s, itn_s = scipy_lsqr(A, b, atol=SQRT_EPS, btol=SQRT_EPS, conlim=1 / SQRT_EPS, iter_lim=10 * n)[:3]
    assert istop == istop_s and abs(itn - itn_s) <= 0.05 * itn_s
    assert np.linalg.norm(x - xs) <= 1e-4 * np.linalg.norm(xs)
    assert ok == (istop not in (3, 6, 7))
