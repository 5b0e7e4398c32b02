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
