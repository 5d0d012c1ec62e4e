"""BiCGStab with the reference's call semantics (oracle; test infrastructure only).

Algorithm: Barrett et al., "Templates for the Solution of Linear Systems"
(SIAM 1994), sec. 2.3.8, BiCGSTAB with M = I — the algorithm of SciPy's
``bicgstab`` that the reference calls at second_projection.py:230 as
``bicgstab(lap, rhs_inner, tol=ud.tol, maxiter=ud.max_iterations,
callback=counter)`` (SciPy 1.7.3, requirements.txt:11).

Reference semantics reproduced:
  * x0 = 0 every solve;
  * legacy tolerance (atol=None): if ||b - A x0|| <= tol (an absolute test)
    return x0 at once; otherwise stop when the recursive residual satisfies
    ||s|| <= tol*||b|| (early half-step exit, x += alpha p) or
    ||r|| <= tol*||b||;
  * breakdown tests |rho| < eps^2 and |omega| < eps^2;
  * ``info``: 0 converged, maxiter if not converged, <0 breakdown (the
    reference discards it; here it is returned).
``niter`` counts completed full iterations (the callback count); a half-step
exit is reported through ``half_step``.
"""
from dataclasses import dataclass

import numpy as np


@dataclass
class SolveResult:
    x: np.ndarray
    info: int
    niter: int
    half_step: bool
    resid: float       # last recursive residual norm
    bnrm2: float
    history: list      # recursive residual norm after each iteration


def bicgstab(matvec, b, tol=1e-8, maxiter=6000, x0=None, record=False, on_iter=None):
    b = np.asarray(b, dtype=np.float64)
    n = b.shape[0]
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    r = b - matvec(x) if x.any() else b.copy()
    bnrm2 = float(np.linalg.norm(b))
    # legacy early exit on the absolute initial residual (SciPy 1.7.3 _get_atol)
    if float(np.linalg.norm(r)) <= tol:
        return SolveResult(x, 0, 0, False, float(np.linalg.norm(r)), bnrm2, [])
    atol = tol * bnrm2
    eps2 = np.finfo(np.float64).eps ** 2
    rtld = r.copy()
    p = v = None
    rho1 = alpha = omega = None
    hist = []
    for it in range(1, maxiter + 1):
        rho = float(np.dot(rtld, r))
        if abs(rho) < eps2:
            return SolveResult(x, -10, it - 1, False, float(np.linalg.norm(r)), bnrm2, hist)
        if it > 1:
            beta = (rho / rho1) * (alpha / omega)
            p = p - omega * v
            p = beta * p
            p = p + r
        else:
            p = r.copy()
        v = matvec(p)
        rv = float(np.dot(rtld, v))
        if rv == 0.0:
            return SolveResult(x, -11, it - 1, False, float(np.linalg.norm(r)), bnrm2, hist)
        alpha = rho / rv
        s = r - alpha * v
        snrm = float(np.linalg.norm(s))
        if snrm <= atol:
            x = x + alpha * p
            return SolveResult(x, 0, it - 1, True, snrm, bnrm2, hist)
        t = matvec(s)
        omega = float(np.dot(t, s)) / float(np.dot(t, t))
        x = x + alpha * p
        x = x + omega * s
        r = s - omega * t
        rnrm = float(np.linalg.norm(r))
        if record:
            hist.append(rnrm)
        if on_iter is not None:
            on_iter(it, x, r, p, s)
        if rnrm <= atol:
            return SolveResult(x, 0, it, False, rnrm, bnrm2, hist)
        if abs(omega) < eps2:
            return SolveResult(x, -11, it, False, rnrm, bnrm2, hist)
        rho1 = rho
    return SolveResult(x, maxiter, maxiter, False, float(np.linalg.norm(r)), bnrm2, hist)


def true_rel_resid(matvec, b, x):
    """||b - A x|| / ||b|| recomputed from scratch in FP64."""
    b = np.asarray(b, dtype=np.float64)
    nb = float(np.linalg.norm(b))
    return float(np.linalg.norm(b - matvec(np.asarray(x, dtype=np.float64)))) / (nb if nb > 0 else 1.0)
