"""Pins of the oracle's adjoint (transposed V-cycle) and of its
double-double arithmetic (the reference's DD128 path,
VLP/DD128_QD256/DD128.cs:52-203)."""
import mpmath
import numpy as np
import pytest

from oracle import adjoint_cycles, dst_adjoint, dst_solve, forward_cycles, Oracle
from oracle.oracle import dd_op


def _rand(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape)


@pytest.mark.parametrize("seg_x,seg_y,hx,hy,K", [(8, 8, 1 / 8, 1 / 8, 1), (16, 8, 0.05, 0.11, 2), (32, 32, 1 / 32, 1 / 32, 3)])
def test_adjoint_dot_product(seg_x, seg_y, hx, hy, K):
    # Phi^K is linear in (u0, f, g) (zero inputs give zero output), so
    # <ubar, Phi(v)> must equal <Phi^T(ubar), v> for every direction v.
    shape = (seg_x + 1, seg_y + 1)
    ubar = _rand(shape, 1)
    ubar[0, :] = ubar[-1, :] = ubar[:, 0] = ubar[:, -1] = 0
    fb, gb, u0b = adjoint_cycles(seg_x, seg_y, hx, hy, ubar, K)
    for s in range(3):
        vf = _rand(shape, 10 + s)
        vu = _rand(shape, 20 + s)  # interior: initial guess, boundary: g
        y = forward_cycles(seg_x, seg_y, hx, hy, vf, vu, K)
        lhs = np.sum(ubar * y)
        inner = np.zeros(shape, bool)
        inner[1:-1, 1:-1] = True
        rhs = np.sum(fb[inner] * vf[inner]) + np.sum(gb[~inner] * vu[~inner]) + np.sum(u0b[inner] * vu[inner])
        assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + np.abs(ubar).sum())


def test_adjoint_matches_brute_force_jacobian():
    seg_x, seg_y, hx, hy, K = 4, 8, 0.2, 0.1, 2
    shape = (seg_x + 1, seg_y + 1)
    n = shape[0] * shape[1]
    ubar = _rand(shape, 3)
    fb, gb, u0b = adjoint_cycles(seg_x, seg_y, hx, hy, ubar, K)
    # column k of the Jacobian w.r.t. f: Phi(e_k in f, u = 0)
    jf = np.zeros(n)
    ju = np.zeros(n)
    zero = np.zeros(shape)
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        jf[k] = np.sum(ubar * forward_cycles(seg_x, seg_y, hx, hy, e.reshape(shape), zero, K))
        ju[k] = np.sum(ubar * forward_cycles(seg_x, seg_y, hx, hy, zero, e.reshape(shape), K))
    inner = np.zeros(shape, bool)
    inner[1:-1, 1:-1] = True
    np.testing.assert_allclose(fb[inner], jf.reshape(shape)[inner], rtol=1e-11, atol=1e-13)
    np.testing.assert_allclose(gb[~inner], ju.reshape(shape)[~inner], rtol=1e-11, atol=1e-13)
    np.testing.assert_allclose(u0b[inner], ju.reshape(shape)[inner], rtol=1e-11, atol=1e-13)


def test_truncated_adjoint_tends_to_exact_adjoint():
    seg, h = 32, 1 / 32
    shape = (seg + 1, seg + 1)
    ubar = _rand(shape, 4)
    ubar[0, :] = ubar[-1, :] = ubar[:, 0] = ubar[:, -1] = 0
    lam, gex = dst_adjoint(ubar, h, h)
    fb, gb, _ = adjoint_cycles(seg, seg, h, h, ubar, 40)
    assert np.abs(fb - lam).max() < 1e-10 * np.abs(lam).max()
    assert np.abs(gb - gex).max() < 1e-10 * np.abs(gex).max()


def test_dst_adjoint_finite_differences():
    seg, h = 16, 1 / 16
    shape = (seg + 1, seg + 1)
    rng = np.random.default_rng(7)
    f0, g0 = rng.standard_normal(shape), rng.standard_normal(shape)
    uobs = rng.standard_normal(shape)

    def loss(f, g):
        u = dst_solve(f, g, h, h)
        return 0.5 * np.sum((u - uobs)[1:-1, 1:-1] ** 2)

    u = dst_solve(f0, g0, h, h)
    ubar = np.zeros(shape)
    ubar[1:-1, 1:-1] = (u - uobs)[1:-1, 1:-1]
    fb, gb = dst_adjoint(ubar, h, h)
    for s in range(4):
        df, dg = rng.standard_normal(shape), rng.standard_normal(shape)
        eps = 1e-3
        fd = (loss(f0 + eps * df, g0 + eps * dg) - loss(f0 - eps * df, g0 - eps * dg)) / (2 * eps)
        an = np.sum(fb[1:-1, 1:-1] * df[1:-1, 1:-1]) + np.sum(gb * dg)
        assert abs(fd - an) <= 1e-8 * abs(an)


# -------------------------------------------------------------- double-double
def _mp(hi, lo):
    return mpmath.mpf(hi) + mpmath.mpf(lo)


def test_dd_arithmetic_against_mpmath():
    mpmath.mp.prec = 200
    rng = np.random.default_rng(11)
    worst = 0.0
    for _ in range(400):
        a = (float(rng.standard_normal() * 10.0 ** rng.integers(-5, 5)), 0.0)
        b = (float(rng.standard_normal() * 10.0 ** rng.integers(-5, 5)), 0.0)
        a = (a[0], a[0] * 2.0**-60 * rng.standard_normal())
        b = (b[0], b[0] * 2.0**-60 * rng.standard_normal())
        for op in "+-*/":
            hi, lo = dd_op(op, a, b)
            x, y = _mp(*a), _mp(*b)
            exact = {"+": x + y, "-": x - y, "*": x * y, "/": x / y}[op]
            if exact == 0:
                continue
            rel = abs((_mp(hi, lo) - exact) / exact)
            # |x+y| may cancel: bound the error relative to the operands too
            scale = max(abs(x), abs(y)) if op in "+-" else abs(exact)
            worst = max(worst, float(abs(_mp(hi, lo) - exact) / scale))
            assert abs(lo) <= abs(hi) * 2.0**-52 or hi == 0
    assert worst < 1e-30


def test_dd_multigrid_reaches_exact_discrete_solution():
    # 17 x 17 Poisson with rational data: exact solution by mpmath LU at 60 digits
    seg, h = 16, 1 / 16
    rng = np.random.default_rng(12)
    f = np.round(rng.standard_normal((seg + 1, seg + 1)) * 8) / 8
    g = np.zeros_like(f)
    g[0, :] = 1.0
    g[:, -1] = 0.5
    o = Oracle(seg, seg, h, h, f, "dd")
    o.set_u(g)
    for _ in range(45):
        o.vcycle(3, 3, 3)
    hi, lo = o.get_u(with_lo=True)
    mpmath.mp.dps = 60
    n = seg - 1
    idx = {(i, j): k for k, (i, j) in enumerate((i, j) for i in range(1, seg) for j in range(1, seg))}
    A = mpmath.zeros(n * n, n * n)
    b = mpmath.zeros(n * n, 1)
    ih2 = mpmath.mpf(1) / mpmath.mpf(h) ** 2
    for (i, j), k in idx.items():
        A[k, k] = 4 * ih2
        b[k] = mpmath.mpf(f[i, j])
        for di, dj in ((-1, 0), (1, 0), (0, -1), (0, 1)):
            q = (i + di, j + dj)
            if q in idx:
                A[k, idx[q]] = -ih2
            else:
                b[k] += mpmath.mpf(g[q]) * ih2
    x = mpmath.lu_solve(A, b)
    err = max(abs(_mp(hi[i, j], lo[i, j]) - x[k]) for (i, j), k in idx.items())
    assert err < 1e-27
