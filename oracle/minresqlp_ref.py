"""ORACLE (test infrastructure only) -- the reference's MINRES-QLP (`sr_qlp`).

Restates `minresqlp` of src/External/IterativeSolvers/minresqlp.jl (the
reference's own module; called from src/Algorithms/SR/SRIterative.jl:92-100
with maxiter = 10 Np and tol = precision) for the unpreconditioned, unshifted,
real symmetric case, with the behaviour of that file as written:

  * the Lanczos recurrence in Paige-Saunders form (minresqlp.jl:163-199):
    v = r3 / beta, r3 = A v - (beta / beta_prev) r1, alfa = r3 . v,
    r3 -= (alfa / beta) r2, betan = ||r3||;
  * the left rotation Q_k and the right rotations P_{k-2,k}, P_{k-1,k} with
    SymGivens (minresqlp.jl:202-240, 484-509), the xnorm recurrence
    (:243-261), the next right rotation and the norm estimates (:330-352);
  * the solution update: the file's MINRES-branch test reads
    `flag != flag0` (minresqlp.jl:265), which is false while no stopping flag
    is set, so every step takes the QLP branch (:277-318);
  * gmin: the update at minresqlp.jl:344 assigns a local, so the stored gmin
    keeps its first-step value and Acond = Anorm / gmin_1;
  * stopping (:356-393): flags 8 (maxiter), 7 (Acond >= 1e15), 6 (xnorm >=
    1e7), 5 (Anorm xnorm eps >= beta1), 4 / 3 (relAres / relres below eps),
    2 (relAres <= tol), 1 (relres <= tol), later ones overriding earlier ones;
    flags 2, 4, 6 and 7 take back the step count and the last rnorm / relres /
    Acond (not x); `converged` is resnorm <= tol or any flag set (:154), so a
    run ended by maxiter also reports convergence.

`branch="minres"` replaces the QLP-branch update by the MINRES-branch one of
the same file (minresqlp.jl:266-275: w = (v - epln w_{k-2} - dlta w_{k-1}) /
gama, x += tau w), which is what the device runs; in exact arithmetic the two
give the same iterates for a nonsingular matrix, and the tests pin that
numerically.

Parity of this restatement is pinned in tests/test_minres_oracle.py: its
iterates equal scipy's MINRES iterates, xnorm equals ||x_k||, rnorm equals
||b - A x_k||, both branches agree, and the solution of np.linalg.solve.
"""
import math

import numpy as np

TRAN_COND = 10e6
MAX_XNORM = 10e6
ACOND_LIM = 10e14


def sym_givens(a, b):
    """(c, s, r) with r >= 0 and [c s; s -c] [a; b] = [r; 0]
    (minresqlp.jl:484-509)."""
    if b == 0:
        c = 1.0 if a == 0 else math.copysign(1.0, a)
        return c, 0.0, abs(a)
    if a == 0:
        return 0.0, math.copysign(1.0, b), abs(b)
    if abs(b) > abs(a):
        t = a / b
        s = math.copysign(1.0, b) / math.sqrt(1 + t * t)
        c = s * t
        return c, s, b / s
    t = b / a
    c = math.copysign(1.0, a) / math.sqrt(1 + t * t)
    s = c * t
    return c, s, a / c


class Scalars:
    """The scalar recurrences of one MINRES-QLP run, advanced one Lanczos
    step at a time from (alfa, betan); no vectors."""

    def __init__(self, beta1, tol, maxiter):
        self.tol, self.maxiter = tol, maxiter
        self.flag0 = self.flag = -2
        self.iters = self.qlpiter = 0
        self.beta = self.tau = self.taul = 0.0
        self.phi = self.beta1 = self.betan = beta1
        self.gmin = 0.0
        self.cs, self.sn, self.cr1, self.sr1, self.cr2, self.sr2 = -1.0, 0.0, -1.0, 0.0, -1.0, 0.0
        self.dltan = self.eplnn = 0.0
        self.gama = self.gamal = self.gamal2 = self.gama_qlp = self.gamal_qlp = 0.0
        self.eta = self.etal = self.etal2 = self.vepln = self.veplnl = self.veplnl2 = self.vepln_qlp = 0.0
        self.ul3 = self.ul2 = self.ul = self.u = self.ul_qlp = self.u_qlp = 0.0
        self.rnorm = beta1
        self.xnorm = self.arnorm = self.relares = 0.0
        self.xl2norm = self.axnorm = self.anorm = 0.0
        self.acond = 1.0
        self.relres = beta1 / (beta1 + 1e-50)
        self.gminl = 0.0
        self.resnorm = 123.0

    def converged(self):
        return self.resnorm <= self.tol or self.flag != self.flag0

    def done(self, iteration):
        return iteration > self.maxiter or self.converged()

    def begin(self):
        """Start of a step: returns (beta_prev, beta) for the Lanczos vectors."""
        self.iters += 1
        betal = self.beta
        self.beta = self.betan
        return betal, self.beta

    def rotate(self, betal, alfa, betan):
        """Everything between the Lanczos step and the vector update; returns
        the MINRES-branch coefficients (epln, dlta_qlp, gama_tmp, tau)."""
        self.betan = betan
        self.pnorm = math.sqrt(betal * betal + alfa * alfa + betan * betan)
        dbar = self.dltan
        dlta = self.cs * dbar + self.sn * alfa
        epln = self.eplnn
        gbar = self.sn * dbar - self.cs * alfa
        self.eplnn = self.sn * betan
        self.dltan = -self.cs * betan
        dlta_qlp = dlta
        self.gamal3 = self.gamal2
        self.gamal2 = self.gamal
        self.gamal = self.gama
        self.cs, self.sn, self.gama = sym_givens(gbar, betan)
        gama_tmp = self.gama
        self.taul2 = self.taul
        self.taul = self.tau
        self.tau = self.cs * self.phi
        self.axnorm = math.sqrt(self.axnorm ** 2 + self.tau ** 2)
        self.phi = self.sn * self.phi
        if self.iters > 2:
            self.veplnl2 = self.veplnl
            self.etal2 = self.etal
            self.etal = self.eta
            d = self.sr2 * self.vepln - self.cr2 * dlta
            self.veplnl = self.cr2 * self.vepln + self.sr2 * dlta
            dlta = d
            self.eta = self.sr2 * self.gama
            self.gama = -self.cr2 * self.gama
        if self.iters > 1:
            self.cr1, self.sr1, self.gamal = sym_givens(self.gamal, dlta)
            self.vepln = self.sr1 * self.gama
            self.gama = -self.cr1 * self.gama
        ul4 = self.ul3
        self.ul3 = self.ul2
        if self.iters > 2:
            self.ul2 = (self.taul2 - self.etal2 * ul4 - self.veplnl2 * self.ul3) / self.gamal2
        if self.iters > 1:
            self.ul = (self.taul - self.etal * self.ul3 - self.veplnl * self.ul2) / self.gamal
        xn = math.sqrt(self.xl2norm ** 2 + self.ul2 ** 2 + self.ul ** 2)
        if abs(self.gama) > np.finfo(np.float64).tiny and xn < MAX_XNORM:
            self.u = (self.tau - self.eta * self.ul2 - self.vepln * self.ul) / self.gama
            if math.sqrt(xn * xn + self.u * self.u) > MAX_XNORM:
                self.u = 0.0
                self.flag = 6
        else:
            self.u = 0.0
            self.flag = 9
        self.xl2norm = math.sqrt(self.xl2norm ** 2 + self.ul2 ** 2)
        self.xnorm = math.sqrt(self.xl2norm ** 2 + self.ul ** 2 + self.u ** 2)
        self.gbar = gbar
        return epln, dlta_qlp, gama_tmp, self.tau

    def minres_branch(self):
        """The file's own test for the MINRES-branch update (as written)."""
        return self.acond < TRAN_COND and self.flag != self.flag0 and self.qlpiter == 0

    def finish(self):
        """Next right rotation, norm estimates and the stopping flags."""
        gamal_tmp = self.gamal
        self.cr2, self.sr2, self.gamal = sym_givens(self.gamal, self.eplnn)
        self.gamal_qlp = gamal_tmp
        self.vepln_qlp = self.vepln
        self.gama_qlp = self.gama
        self.ul_qlp = self.ul
        self.u_qlp = self.u
        abs_gama = abs(self.gama)
        self.anorm = max(self.anorm, self.pnorm, self.gamal, abs_gama)
        if self.iters == 1:
            self.gmin = self.gama
            self.gminl = self.gmin
        elif self.iters > 1:
            self.gminl = self.gmin   # the new minimum goes to a local (minresqlp.jl:344)
        acondl = self.acond
        self.acond = self.anorm / self.gmin
        rnorml = self.rnorm
        relresl = self.relres
        if self.flag != 9:
            self.rnorm = self.phi
        self.relres = self.rnorm / (self.anorm * self.xnorm + self.beta1)
        rootl = math.sqrt(self.gbar ** 2 + self.dltan ** 2)
        self.arnorm = rnorml * rootl
        self.relares = rootl / self.anorm
        epsx = self.anorm * self.xnorm * np.finfo(np.float64).eps
        if self.flag == self.flag0 or self.flag == 9:
            t1, t2 = 1 + self.relres, 1 + self.relares
            if self.iters >= self.maxiter:
                self.flag = 8
            if self.acond >= ACOND_LIM:
                self.flag = 7
            if self.xnorm >= MAX_XNORM:
                self.flag = 6
            if epsx >= self.beta1:
                self.flag = 5
            if t2 <= 1:
                self.flag = 4
            if t1 <= 1:
                self.flag = 3
            if self.relares <= self.tol:
                self.flag = 2
            if self.relres <= self.tol:
                self.flag = 1
        if self.flag in (2, 4, 6, 7):
            self.iters -= 1
            self.acond = acondl
            self.rnorm = rnorml
            self.relres = relresl
        self.resnorm = min(self.relres, self.relares)


def minresqlp(A, b, tol, maxiter, branch="reference", history=None):
    """Returns (x, scalars).  branch "reference": the vector updates the file
    runs (QLP branch); "minres": the MINRES-branch update.  history, if a
    list, receives (x_k, scalars snapshot dict) after every step."""
    mv = A if callable(A) else (lambda v: A @ v)
    b = np.asarray(b, dtype=np.float64)
    n = b.shape[0]
    x = np.zeros(n)
    r1, r2, r3 = np.zeros(n), b.copy(), b.copy()
    sc = Scalars(float(np.linalg.norm(r2)), tol, maxiter)
    w, wl, xl2, wl2 = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    iteration = 1
    while not sc.done(iteration):
        betal, beta = sc.begin()
        v = r3 / beta
        r3 = mv(v)
        if sc.iters > 1:
            r3 = r3 - (beta / betal) * r1
        alfa = float(r3 @ v)
        r3 = r3 - (alfa / beta) * r2
        r1, r2 = r2, r3.copy()
        betan = float(np.linalg.norm(r3))
        if sc.iters == 1 and betan == 0.0:
            # b is an eigenvector: x = b / alfa, flag -1 (the file's intent at
            # minresqlp.jl:187-188; its test of an undefined field would throw)
            x = b / alfa
            sc.flag = -1
            break
        epln, dlta_qlp, gama_tmp, tau = sc.rotate(betal, alfa, betan)
        if branch == "minres" or sc.minres_branch():
            wl2, wl = wl, w
            w = (v - epln * wl2 - dlta_qlp * wl) / gama_tmp
            if sc.xnorm < MAX_XNORM:
                x = x + tau * w
            else:
                sc.flag = 6
        else:
            sc.qlpiter += 1
            if sc.qlpiter == 1 and sc.iters > 1:
                raise NotImplementedError("MINRES -> QLP transfer after step 1 (not reached as written)")
            cr1, sr1, cr2, sr2 = sc.cr1, sc.sr1, sc.cr2, sc.sr2
            if sc.iters == 1:
                wl2, wl, w = wl, v * sr1, -v * cr1
            elif sc.iters == 2:
                wl2, wl, w = wl, w * cr1 + v * sr1, w * sr1 - v * cr1
            else:
                wl2_old, w_old = wl, w
                w2 = wl2_old * sr2 - v * cr2
                wl2 = wl2_old * cr2 + v * sr2
                wl = w_old * cr1 + w2 * sr1
                w = w_old * sr1 - w2 * cr1
            xl2 = xl2 + wl2 * sc.ul2
            x = xl2 + wl * sc.ul + w * sc.u
        sc.finish()
        if history is not None:
            history.append((x.copy(), dict(rnorm=sc.rnorm, xnorm=sc.xnorm, relres=sc.relres, anorm=sc.anorm,
                                           iters=sc.iters, flag=sc.flag)))
        iteration = sc.iters
    return x, sc
