"""TEST INFRASTRUCTURE ONLY: independent exact discrete solver.

Solves the 5-point Dirichlet problem  -(D_xx + D_yy)_h u = f  on the interior
of an (nx, ny) node grid with boundary values taken from ``u`` itself, by the
DST-I eigen-decomposition of the 1-D second-difference matrices (the textbook
fast Poisson solver).  This is the discrete solution every converged multigrid
run must reach, whatever its smoother or cycle.
"""
import numpy as np
from scipy.fft import dstn, idstn


def _eig(n, h):
    k = np.arange(1, n + 1)
    return (2.0 - 2.0 * np.cos(np.pi * k / (n + 1))) / (h * h)


def boundary_rhs(g, hx, hy):
    """B g: contributions of the boundary values of g to the interior rows."""
    b = np.zeros((g.shape[0] - 2, g.shape[1] - 2))
    b[0, :] += g[0, 1:-1] / hx**2
    b[-1, :] += g[-1, 1:-1] / hx**2
    b[:, 0] += g[1:-1, 0] / hy**2
    b[:, -1] += g[1:-1, -1] / hy**2
    return b


def _solve_interior(b, hx, hy):
    lam = _eig(b.shape[0], hx)[:, None] + _eig(b.shape[1], hy)[None, :]
    return idstn(dstn(b, type=1) / lam, type=1)


def dst_solve(f, g, hx, hy):
    """Exact discrete solution.  f: PDE right-hand side (u_xx+u_yy = -f) at all
    nodes (None = Laplace); g: array whose boundary entries are the Dirichlet
    values.  Returns the full grid (boundary copied from g)."""
    g = np.asarray(g, dtype=np.float64)
    b = boundary_rhs(g, hx, hy)
    if f is not None:
        b = b + np.asarray(f, dtype=np.float64)[1:-1, 1:-1]
    u = g.copy()
    u[1:-1, 1:-1] = _solve_interior(b, hx, hy)
    return u


def dst_adjoint(ubar, hx, hy):
    """Gradient of L through the exact solve, given dL/du (interior nodes).
    Returns (fbar, gbar) on the full grid: fbar = A^{-1} ubar (A symmetric),
    gbar = B^T fbar on boundary nodes (corners 0)."""
    ubar = np.asarray(ubar, dtype=np.float64)
    lam = np.zeros_like(ubar)
    lam[1:-1, 1:-1] = _solve_interior(ubar[1:-1, 1:-1], hx, hy)
    gbar = np.zeros_like(ubar)
    gbar[0, 1:-1] = lam[1, 1:-1] / hx**2
    gbar[-1, 1:-1] = lam[-2, 1:-1] / hx**2
    gbar[1:-1, 0] = lam[1:-1, 1] / hy**2
    gbar[1:-1, -1] = lam[1:-1, -2] / hy**2
    return lam, gbar
