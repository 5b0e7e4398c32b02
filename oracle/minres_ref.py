"""ORACLE (test infrastructure only) -- MINRES for the iterative SR solve.

The reference's `sr_minres` (src/Algorithms/SR/SRIterative.jl:101-108) calls
`minres(Sprecond, F; maxiter = 10 Np, tol = precision)` of IterativeSolvers.jl
(pinned 0.8.1 in the reference's Manifest.toml:153-157; the package is not part
of the reference tree).  That routine is the MINRES method of Paige and
Saunders (SIAM J. Numer. Anal. 12, 1975): Lanczos vectors v_k of the symmetric
matrix, the tridiagonal's active column reduced by Givens rotations, search
directions w = V R^-1 by a three-term recurrence, and the residual norm
carried by the rotated right-hand side, so no extra product is needed to test
convergence.  Started from x0 = 0 it stops when |rhs_2| <= tol ||b|| or after
maxiter products.  The rotations follow LAPACK dlartg's sign convention
(LinearAlgebra.givensAlgorithm).

Parity of this restatement is pinned in tests/test_minres_oracle.py against
what the mathematics fixes: the estimated residual equals ||b - A x_k|| at
every step, x_k equals scipy's MINRES iterate after the same number of
products, exact termination in n steps, and the solution of np.linalg.solve.
"""
import math

import numpy as np


def givens(f, g):
    """(c, s, r) with [c s; -s c] [f; g] = [r; 0]; c > 0 when |f| > |g|."""
    if g == 0.0:
        return 1.0, 0.0, f
    if f == 0.0:
        return 0.0, 1.0, g
    r = math.hypot(f, g)
    c, s = f / r, g / r
    if abs(f) > abs(g) and c < 0.0:
        c, s, r = -c, -s, -r
    return c, s, r


def minres(A, b, tol, maxiter, history=None):
    """Returns (x, products, resnorm, converged).  A: symmetric (n, n) or a
    callable v -> A v; history, if a list, receives (x_k, resnorm_k) after
    every product."""
    mv = A if callable(A) else (lambda v: A @ v)
    b = np.asarray(b, dtype=np.float64)
    n = b.shape[0]
    x = np.zeros(n)
    res0 = float(np.linalg.norm(b))
    tolerance = tol * res0
    resnorm = res0
    if res0 == 0.0:
        return x, 0, 0.0, True
    v_prev, v_curr = np.zeros(n), b / res0
    w_prev, w_curr = np.zeros(n), np.zeros(n)
    h1 = h2 = 0.0
    rhs1 = res0
    c_prev = s_prev = c_curr = s_curr = 0.0
    it = 1
    while not (it > maxiter or resnorm <= tolerance):
        v_next = mv(v_curr)
        if it > 1:
            v_next = v_next - h2 * v_prev
        h3 = float(v_curr @ v_next)
        v_next = v_next - h3 * v_curr
        h4 = float(np.linalg.norm(v_next))
        if it > 2:
            h1 = s_prev * h2
            h2 = c_prev * h2
        if it > 1:
            h2, h3 = c_curr * h2 + s_curr * h3, -s_curr * h2 + c_curr * h3
        c, s, h3 = givens(h3, h4)
        rhs2 = -s * rhs1
        rhs1 = c * rhs1
        w_next = v_curr.copy()
        if it > 1:
            w_next -= h2 * w_curr
        if it > 2:
            w_next -= h1 * w_prev
        w_next /= h3
        x = x + rhs1 * w_next
        v_next = v_next / h4 if h4 > 0.0 else v_next * 0.0
        v_prev, v_curr = v_curr, v_next
        w_prev, w_curr = w_curr, w_next
        c_prev, s_prev, c_curr, s_curr = c_curr, s_curr, c, s
        rhs1 = rhs2
        h2 = h4
        resnorm = abs(rhs2)
        if history is not None:
            history.append((x.copy(), resnorm))
        it += 1
    return x, it - 1, resnorm, resnorm <= tolerance
