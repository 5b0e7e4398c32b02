"""ORACLE (test infrastructure only) -- LSQR for the iterative SR solve.

The reference's `sr_lsq` (src/Algorithms/SR/SRIterative.jl:109-116) calls
`lsqr(Sprecond, F; maxiter = 10 Np)` of IterativeSolvers.jl (pinned 0.8.1 in
the reference's Manifest.toml:153-157; not part of the reference tree) with
no tolerances, so the package defaults apply: atol = btol = sqrt(eps),
conlim = 1 / sqrt(eps).  That routine is LSQR of Paige and Saunders (ACM TOMS
8, 1982) without damping: Golub-Kahan bidiagonalisation (two products per
step, A and A^T, both S here), one plane rotation per step, x and the search
direction w updated by a two-term recurrence, and the published stopping
tests on the estimates of ||r||, ||A^T r||, ||A||, cond(A) and ||x||:

    istop 1: ||r|| / ||b|| <= btol + atol ||A|| ||x|| / ||b||
    istop 2: ||A^T r|| / (||A|| ||r||) <= atol
    istop 3: 1 / cond(A) <= 1 / conlim
    istop 4-6: the same three below machine precision;  7: maxiter.

The package reports convergence unless istop is 3, 6 or 7.

Parity of this restatement is pinned in tests/test_minres_oracle.py against
scipy's LSQR (the same published algorithm): the iterates after every step,
and the stopping step and reason with the reference's defaults.
"""
import math

import numpy as np

SQRT_EPS = math.sqrt(np.finfo(np.float64).eps)


def lsqr(A, b, maxiter, atol=SQRT_EPS, btol=SQRT_EPS, conlim=1.0 / SQRT_EPS, history=None):
    """Returns (x, steps, istop, converged).  A: (n, n) or a callable v -> A v
    for a symmetric A (A^T u = A u)."""
    mv = A if callable(A) else (lambda v: A @ v)
    b = np.asarray(b, dtype=np.float64)
    n = b.shape[0]
    x = np.zeros(n)
    ctol = 1.0 / conlim if conlim > 0 else 0.0
    beta = float(np.linalg.norm(b))
    u = b / beta if beta > 0 else b.copy()
    v = mv(u) if beta > 0 else np.zeros(n)
    alfa = float(np.linalg.norm(v))
    if alfa > 0:
        v = v / alfa
    w = v.copy()
    if alfa * beta == 0.0:
        return x, 0, 0, True
    anorm = ddnorm = xxnorm = z = 0.0
    cs2, sn2 = -1.0, 0.0
    rhobar, phibar, bnorm = alfa, beta, beta
    itn = istop = 0
    while itn < maxiter:
        itn += 1
        u = mv(v) - alfa * u
        beta = float(np.linalg.norm(u))
        if beta > 0:
            u = u / beta
            anorm = math.sqrt(anorm * anorm + alfa * alfa + beta * beta)
            v = mv(u) - beta * v
            alfa = float(np.linalg.norm(v))
            if alfa > 0:
                v = v / alfa
        rho = math.hypot(rhobar, beta)
        cs, sn = rhobar / rho, beta / rho
        theta = sn * alfa
        rhobar = -cs * alfa
        phi = cs * phibar
        phibar = sn * phibar
        tau = sn * phi
        t1, t2 = phi / rho, -theta / rho
        ddnorm += float(w @ w) / (rho * rho)
        x = x + t1 * w
        w = v + t2 * w
        delta = sn2 * rho
        gambar = -cs2 * rho
        rhs = phi - delta * z
        zbar = rhs / gambar
        xnorm = math.sqrt(xxnorm + zbar * zbar)
        gamma = math.hypot(gambar, theta)
        cs2, sn2 = gambar / gamma, theta / gamma
        z = rhs / gamma
        xxnorm += z * z
        acond = anorm * math.sqrt(ddnorm)
        rnorm = abs(phibar)
        arnorm = alfa * abs(tau)
        test1 = rnorm / bnorm
        test2 = arnorm / (anorm * rnorm) if anorm * rnorm > 0 else 0.0
        test3 = 1.0 / acond if acond > 0 else math.inf
        t1n = test1 / (1 + anorm * xnorm / bnorm)
        rtol = btol + atol * anorm * xnorm / bnorm
        if itn >= maxiter:
            istop = 7
        if 1 + test3 <= 1:
            istop = 6
        if 1 + test2 <= 1:
            istop = 5
        if 1 + t1n <= 1:
            istop = 4
        if test3 <= ctol:
            istop = 3
        if test2 <= atol:
            istop = 2
        if test1 <= rtol:
            istop = 1
        if history is not None:
            history.append((x.copy(), dict(rnorm=rnorm, xnorm=xnorm, anorm=anorm, istop=istop)))
        if istop:
            break
    return x, itn, istop, istop not in (3, 6, 7)
