"""CPU restatement of the reference's nonlinear solve, for tests.

TEST INFRASTRUCTURE ONLY: imported by tests/; the product never imports it.

The reference solves R(U) = 0 with PETSc SNESNEWTONLS (src/snes_init.c:18-22:
atol 1e-10, rtol = stol = 1e-20, one Newton step per SNESSolve, BT line
search of cubic order) inside the loop of src/snes_solve.c:24-33, which calls
SNESSolve until it stops for a reason other than the iteration limit and
declares stagnation when |F_old - F_new| / F_old < 1e-6 (:30).  The linear
solve there is GMRES/ASM to PETSc's default rtol 1e-5; here it is a direct
sparse LU (scipy), so the iterates are those of exact Newton.

The line search is the backtracking method of Dennis & Schnabel, "Numerical
Methods for Unconstrained Optimization and Nonlinear Equations" (1983),
algorithm A6.3.1 (quadratic first step, cubic after), on f = |F|^2 / 2 with
alpha = 1e-4, steps clamped to [0.1, 0.5] of the previous one, which is the
method PETSc's SNESLINESEARCHBT implements.

The tangent has its Dirichlet rows and columns zeroed (MatZeroRowsColumns
without a right-hand-side fix, src/snesiga_Tangent.c:270), so the first step,
which moves the Dirichlet dofs onto g, ignores their coupling into the
interior; convergence is quadratic once U = g on those rows.

Pins (tests/test_oracle_newton.py): the line search backtracks on a planted
overshoot (atan), accepts full steps where Newton is well behaved, and returns
the exact minimiser on a quadratic (cubic model with a = 0); Newton on the
oracle's residual and tangent reaches R(U) < 1e-12 with quadratic contraction.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import oracle as O


def line_search_bt(f_of, fnorm, slope, alpha=1e-4, max_its=40, minlambda=1e-12):
    """Backtracking along the Newton direction.  f_of(lam) returns |F(U - lam Y)|;
    slope is the directional derivative of |F|^2/2 (negative).  Returns
    (lam, gnorm, ok)."""
    f0 = 0.5 * fnorm * fnorm
    lam = 1.0
    g = f_of(lam)
    if np.isfinite(g) and 0.5 * g * g <= f0 + lam * alpha * slope:
        return lam, g, True
    lt = -slope / (g * g - fnorm * fnorm - 2.0 * lam * slope) if np.isfinite(g) else 0.1 * lam
    lprev, gprev = lam, g
    lam = 0.5 * lam if lt > 0.5 * lam else (0.1 * lam if lt <= 0.1 * lam else lt)
    g = f_of(lam)
    if np.isfinite(g) and 0.5 * g * g < f0 + lam * alpha * slope:
        return lam, g, True
    for _ in range(max_its):
        if lam < minlambda:
            break
        t1 = 0.5 * (g * g - fnorm * fnorm) - lam * slope
        t2 = 0.5 * (gprev * gprev - fnorm * fnorm) - lprev * slope
        a = (t1 / (lam * lam) - t2 / (lprev * lprev)) / (lam - lprev)
        b = (-lprev * t1 / (lam * lam) + lam * t2 / (lprev * lprev)) / (lam - lprev)
        d = max(b * b - 3.0 * a * slope, 0.0)
        lt = -slope / (2.0 * b) if a == 0.0 else (-b + np.sqrt(d)) / (3.0 * a)
        if not (np.isfinite(lt) and np.isfinite(g)):
            lt = 0.1 * lam
        lprev, gprev = lam, g
        lam = 0.5 * lam if lt > 0.5 * lam else (0.1 * lam if lt <= 0.1 * lam else lt)
        g = f_of(lam)
        if np.isfinite(g) and 0.5 * g * g < f0 + lam * alpha * slope:
            return lam, g, True
    return lam, g, False


def newton(pb, U_hist, atol=1e-10, rtol=0.0, max_newton=50, stagnation=True, residual=None, tangent=None):
    """Newton with backtracking from slice 0 of U_hist.  Returns (U_hist,
    info) with info = {converged, newton_its, fnorm_hist}.  converged: 1 by
    tolerance, 0 iteration limit, -1 stagnation, -3 line search failure."""
    residual = residual or (lambda Uh: O.residual(pb, Uh))
    if tangent is None:
        rowptr, col = O.pattern(pb)

        def tangent(Uh):
            return sp.csr_matrix((O.tangent(pb, Uh, rowptr, col)[2], col, rowptr))
    Uh = np.array(U_hist, dtype=np.float64)
    n = O.ndof(pb)
    R = residual(Uh)
    fnorm = np.linalg.norm(R)
    hist = [fnorm]
    fold, fnorm0, conv, its = 999.0, fnorm, 0, 0

    def done(f):
        return (atol > 0 and f < atol) or (rtol > 0 and f < rtol * fnorm0)

    for it in range(max_newton):
        if done(fnorm):
            conv = 1
            break
        K = tangent(Uh)
        Y = spla.spsolve(K.tocsc(), R)
        slope = float(R @ (K @ Y))
        slope = -abs(slope) if slope != 0.0 else -1.0
        U0 = Uh[:n].copy()

        last = {}

        def f_of(lam):
            Uh[:n] = U0 - lam * Y
            last[lam] = residual(Uh)
            return np.linalg.norm(last[lam])

        lam, g, ok = line_search_bt(f_of, fnorm, slope)
        its = it + 1
        if not ok:
            Uh[:n] = U0
            conv = -3
            break
        Uh[:n] = U0 - lam * Y
        R = last[lam]
        fnorm = g
        hist.append(fnorm)
        if stagnation and abs(fold - fnorm) / fold < 1e-6 and not done(fnorm):
            conv = -1
            break
        fold = fnorm
        if done(fnorm):
            conv = 1
            break
    return Uh, {"converged": conv, "newton_its": its, "fnorm_hist": hist}
