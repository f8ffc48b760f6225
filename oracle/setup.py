"""TEST INFRASTRUCTURE ONLY: CPU restatement of the reference's set-up and
post-processing steps around the solve, the checker of the GPU kernels of
``vlp2d_amd/csrc/mg_setup.hpp``.  Arrays are row-major (i, j) with
``j`` contiguous, as the reference's flattened GPU arrays.

Node coordinates are the products ``x = hx * i``, ``y = hy * j``: the reference
accumulates ``x += hx * cLoop`` per host thread (UtilsBorders.cs:125-136,
UtilsDiff.cs:53-75), so its coordinates depend on its thread count; the
product is the value that accumulation approximates.
"""
import numpy as np


def boundary(nx, ny, hx, hy, bottom=None, top=None, left=None, right=None, border=None):
    """Dirichlet data: bottom (j = 0) / top (j = ny-1) edges first, from the
    edge functions of x or border(x, 0) / border(x, yMax); then left (i = 0) /
    right (i = nx-1) from the edge functions of y or border(0, y) /
    border(xMax, y), overwriting the corners (UtilsBorders.cs:115-144 and
    210-239, in the order of VLPRectangleModel.cs:541-542).  Returns the grid
    (interior 0) and the min / max over every value written."""
    u = np.zeros((nx, ny))
    x = hx * np.arange(nx, dtype=np.float64)
    y = hy * np.arange(ny, dtype=np.float64)
    ymax, xmax = hy * (ny - 1), hx * (nx - 1)
    vals = []
    b = bottom(x) if bottom else border(x, np.zeros_like(x))
    t = top(x) if top else border(x, np.full_like(x, ymax))
    u[:, 0], u[:, -1] = b, t
    vals += [b, t]
    lft = left(y) if left else border(np.zeros_like(y), y)
    rgt = right(y) if right else border(np.full_like(y, xmax), y)
    u[0, :], u[-1, :] = lft, rgt
    vals += [lft, rgt]
    allv = np.concatenate([np.asarray(v, dtype=np.float64).ravel() for v in vals])
    return u, float(np.nanmin(allv)), float(np.nanmax(allv))


def init_guess(u, kind, value=0.0):
    """Initial guess of the interior from the boundary values of u
    (UtilsII.cs): 0 = value everywhere (Mean, :15-19), 1 = mean of the node's
    four edge values (ArithmeticMean, :21-33), 2 = mean of the linear
    interpolations along x and y (Linear, :35-75), 3 = the two weighted by the
    distances to the nearer edges (WeightLinear, :77-105).  The weights are FP32
    quotients (float)i / upperX converted to double, as there."""
    u = u.copy()
    nx, ny = u.shape
    ux, uy = nx - 1, ny - 1
    i = np.arange(1, nx - 1)[:, None]
    j = np.arange(1, ny - 1)[None, :]
    if kind == 0:
        u[1:-1, 1:-1] = value
        return u
    r0, r1 = u[0, 1:-1][None, :], u[-1, 1:-1][None, :]
    c0, c1 = u[1:-1, 0][:, None], u[1:-1, -1][:, None]
    if kind == 1:
        u[1:-1, 1:-1] = (r0 + r1 + c0 + c1) / 4.0
        return u
    sx = (i.astype(np.float32) / np.float32(ux)).astype(np.float64)
    sy = (j.astype(np.float32) / np.float32(uy)).astype(np.float64)
    lx = r0 * (1.0 - sx) + r1 * sx
    ly = c0 * (1.0 - sy) + c1 * sy
    if kind == 2:
        u[1:-1, 1:-1] = (lx + ly) / 2.0
        return u
    lngx = np.where(ux > 2 * i, i, ux - i)
    lngy = np.where(uy > 2 * j, j, uy - j)
    w = (lngx.astype(np.float32) / (lngx + lngy).astype(np.float32)).astype(np.float64)
    u[1:-1, 1:-1] = lx * (1.0 - w) + ly * w
    return u


def points_min_max(u):
    """min / max over all nodes, NaN skipped (UtilsDiff.cs:19-40)."""
    return float(np.nanmin(u)), float(np.nanmax(u))


def difference(u, analytic, hx, hy):
    """u - analytic(x, y) on nodes i, j >= 1 (row 0 and column 0 left 0), NaN
    u copied; min / max over the differences, NaN skipped
    (UtilsDiff.cs:42-88)."""
    nx, ny = u.shape
    x = hx * np.arange(nx, dtype=np.float64)[:, None]
    y = hy * np.arange(ny, dtype=np.float64)[None, :]
    d = np.zeros_like(u)
    a = analytic(x[1:], y[:, 1:])
    v = u[1:, 1:]
    d[1:, 1:] = np.where(np.isnan(v), v, v - a)
    return d, float(np.nanmin(d[1:, 1:])), float(np.nanmax(d[1:, 1:]))
