"""Seeded synthetic problems for the benchmark and test configurations.

Random numbers come from SplitMix64 (Steele, Lea, Flood 2014; the generator
published by Vigna): state += 0x9E3779B97F4A7C15, output = mix(state).
Uniforms use the top 53 bits, normals use Box-Muller.  Every configuration is
a pure function of its seed, so any consumer regenerates identical inputs.

Grids: node (i, j) sits at x = hx*i, y = hy*j, arrays are (nx, ny) with j
contiguous (the reference layout, MultiGridSchemeCU.cs:41-45).  ``f`` is the
PDE right-hand side of u_xx + u_yy = -f; ``g`` is a full (nx, ny) array whose
boundary entries are the Dirichlet values (interior entries 0).
"""
from dataclasses import dataclass, field

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    """Vectorised SplitMix64.  ``next(n)`` returns the next n 64-bit outputs."""

    def __init__(self, seed):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def next(self, n=1):
        with np.errstate(over="ignore"):
            k = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + _GOLDEN * k
            self.state = z[-1] if n > 0 else self.state
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            return z ^ (z >> np.uint64(31))

    def uniform(self, n=1):
        return (self.next(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def integers(self, lo, hi, n=1):
        """Uniform integers in [lo, hi]."""
        return lo + np.floor(self.uniform(n) * (hi - lo + 1)).astype(np.int64)

    def normal(self, n=1):
        u1 = self.uniform(n)
        u2 = self.uniform(n)
        return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


@dataclass
class Problem:
    name: str
    seg_x: int
    seg_y: int
    hx: float
    hy: float
    f: np.ndarray = None          # None = Laplace
    g: np.ndarray = None          # boundary values in a full array
    exact: np.ndarray = None      # closed-form solution when known
    meta: dict = field(default_factory=dict)

    @property
    def shape(self):
        return (self.seg_x + 1, self.seg_y + 1)

    @property
    def unknowns(self):
        return (self.seg_x - 1) * (self.seg_y - 1)

    def initial(self, kind="zero"):
        """Boundary + initial guess: 'zero' interior, or the reference 'mean'
        ((min+max)/2 of the boundary values, VLPRectangleModel.cs:540-546)."""
        u = np.zeros(self.shape)
        u[0, :], u[-1, :], u[:, 0], u[:, -1] = self.g[0, :], self.g[-1, :], self.g[:, 0], self.g[:, -1]
        if kind == "mean":
            b = np.concatenate([u[0, :], u[-1, :], u[:, 0], u[:, -1]])
            u[1:-1, 1:-1] = (b.min() + b.max()) / 2.0
        return u


def _coords(seg_x, seg_y, hx, hy):
    return hx * np.arange(seg_x + 1, dtype=np.float64), hy * np.arange(seg_y + 1, dtype=np.float64)


def _edges(shape, left, right, bottom, top):
    """Full array with boundary values; bottom/top first, then left/right
    overwrite the corners (VLPRectangleModel.cs:541-542)."""
    g = np.zeros(shape)
    g[:, 0], g[:, -1] = bottom, top
    g[0, :], g[-1, :] = left, right
    return g


def harmonic_quadratic(seg=256):
    """Config 1: Laplace on the unit square, g = x^2 - y^2 (harmonic, and the
    5-point operator is exact on quadratics, so u = g everywhere)."""
    h = 1.0 / seg
    x, y = _coords(seg, seg, h, h)
    ex = x[:, None] ** 2 - y[None, :] ** 2
    g = _edges(ex.shape, ex[0, :], ex[-1, :], ex[:, 0], ex[:, -1])
    return Problem("laplace_x2_minus_y2", seg, seg, h, h, None, g, ex)


def laplace1(seg=256):
    """The reference's Laplace1*.txt problems: g = u = x^2 - y^2 + 10xy + 10x -
    10y + 1 on the unit square (VLP2D_Data/InputData/Laplace1.txt:10-11)."""
    h = 1.0 / seg
    x, y = _coords(seg, seg, h, h)
    X, Y = x[:, None], y[None, :]
    ex = X**2 - Y**2 + 10 * X * Y + 10 * X - 10 * Y + 1
    g = _edges(ex.shape, ex[0, :], ex[-1, :], ex[:, 0], ex[:, -1])
    return Problem("laplace1", seg, seg, h, h, None, g, ex)


def poisson1():
    """The reference's Poisson1.txt: [0,e]x[0,pi], 128 x 256 segments (unequal
    steps), f = -4, u = (x - e/2)^2 + (y - pi/2)^2 (Poisson1.txt:1-7)."""
    seg_x, seg_y = 128, 256
    hx, hy = np.e / seg_x, np.pi / seg_y
    x, y = _coords(seg_x, seg_y, hx, hy)
    X, Y = x[:, None], y[None, :]
    ex = (X - np.e / 2) ** 2 + (Y - np.pi / 2) ** 2
    f = np.full(ex.shape, -4.0)
    g = _edges(ex.shape, ex[0, :], ex[-1, :], ex[:, 0], ex[:, -1])
    return Problem("poisson1", seg_x, seg_y, hx, hy, f, g, ex)


def sine_modes(seed=2, seg_x=4096, seg_y=2048):
    """Config 2: [0,2]x[0,1], h = 1/2048 in x and y.  f = sum of 16 modes
    a sin(k pi x/2) sin(l pi y), k,l ~ U{1..32}, a ~ N(0,1).  g = 0 on the
    bottom/top edges; left/right edges are cubics c0 + c1 y + c2 y^2 + c3 y^3
    with c ~ N(0,1)."""
    rng = SplitMix64(seed)
    hx, hy = 2.0 / seg_x, 1.0 / seg_y
    x, y = _coords(seg_x, seg_y, hx, hy)
    k = rng.integers(1, 32, 16)
    l = rng.integers(1, 32, 16)
    a = rng.normal(16)
    sx = np.sin(np.pi * np.outer(k, x) / 2.0)       # (16, nx)
    sy = np.sin(np.pi * np.outer(l, y))             # (16, ny)
    f = (sx * a[:, None]).T @ sy
    cl, cr = rng.normal(4), rng.normal(4)
    left = cl[0] + cl[1] * y + cl[2] * y**2 + cl[3] * y**3
    right = cr[0] + cr[1] * y + cr[2] * y**2 + cr[3] * y**3
    g = _edges(f.shape, left, right, 0.0, 0.0)
    return Problem("sine_modes_2x1", seg_x, seg_y, hx, hy, np.ascontiguousarray(f), g,
                   meta={"k": k.tolist(), "l": l.tolist(), "a": a.tolist(), "left": cl.tolist(),
                         "right": cr.tolist(), "seed": seed})


def gaussian_bumps(seed=3, seg=8192, ncenters=64, boundary="sin3x_cosh3y"):
    """Config 3 (and 5): unit square, f = sum_c A_c exp(-|x - x_c|^2 / (2 w_c^2)),
    x_c ~ U(0,1)^2, w_c ~ U(0.01, 0.05), A_c ~ N(0, sigma=100).  The bumps are
    separable, so f is assembled as one (nx, n) @ (n, ny) product.
    boundary: 'sin3x_cosh3y' (g = sin(3x) cosh(3y)) or 'smoothed_normal'
    (N(0,1) per edge node, smoothed by the 9-tap binomial filter)."""
    rng = SplitMix64(seed)
    h = 1.0 / seg
    x, y = _coords(seg, seg, h, h)
    xc = rng.uniform(ncenters)
    yc = rng.uniform(ncenters)
    w = 0.01 + 0.04 * rng.uniform(ncenters)
    A = 100.0 * rng.normal(ncenters)
    ex = np.exp(-((x[None, :] - xc[:, None]) ** 2) / (2 * w[:, None] ** 2))  # (n, nx)
    ey = np.exp(-((y[None, :] - yc[:, None]) ** 2) / (2 * w[:, None] ** 2))  # (n, ny)
    f = np.ascontiguousarray((ex * A[:, None]).T @ ey)
    shape = (seg + 1, seg + 1)
    if boundary == "sin3x_cosh3y":
        g = _edges(shape, np.sin(0.0) * np.cosh(3 * y), np.sin(3 * x[-1]) * np.cosh(3 * y),
                   np.sin(3 * x) * np.cosh(0.0), np.sin(3 * x) * np.cosh(3 * y[-1]))
    else:
        taps = np.array([1, 8, 28, 56, 70, 56, 28, 8, 1], dtype=np.float64) / 256.0
        e = [np.convolve(rng.normal(seg + 1), taps, mode="same") for _ in range(4)]
        g = _edges(shape, e[0], e[1], e[2], e[3])
    return Problem(f"gaussian_bumps_{seg + 1}", seg, seg, h, h, f, g,
                   meta={"seed": seed, "centers": ncenters, "boundary": boundary})


def sine_modes_unit(seed=4, seg=1024, nmodes=32):
    """Config 4 f_true: 32 modes a sin(k pi x) sin(l pi y) on the unit square,
    k,l ~ U{1..32}, a ~ N(0,1)."""
    rng = SplitMix64(seed)
    h = 1.0 / seg
    x, y = _coords(seg, seg, h, h)
    k = rng.integers(1, 32, nmodes)
    l = rng.integers(1, 32, nmodes)
    a = rng.normal(nmodes)
    f = (np.sin(np.pi * np.outer(k, x)) * a[:, None]).T @ np.sin(np.pi * np.outer(l, y))
    return np.ascontiguousarray(f), rng


def config(n, seg=None):
    """Problem of BASELINE config n (1-based).  ``seg`` overrides the size
    for reduced replicas (same formulas)."""
    if n == 1:
        return harmonic_quadratic(seg or 256)
    if n == 2:
        return sine_modes(2, seg or 4096, (seg or 4096) // 2)
    if n == 3:
        return gaussian_bumps(3, seg or 8192, 64, "sin3x_cosh3y")
    if n == 5:
        return gaussian_bumps(5, seg or 16384, 256, "smoothed_normal")
    raise ValueError(f"config {n} has no single-problem generator")
