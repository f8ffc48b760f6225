"""GPU checks of the benchmark configurations that are not bench lines:
the adjoint through an exact solve (1025^2), the adjoint through K truncated
V-cycles at 16385^2 (dot-product identity and central differences; peak HBM
and times reported), and the DD128 cross-check of the mixed-precision solver
on a 513^2 replica."""
import numpy as np
import pytest

from oracle import Oracle, dst_adjoint, dst_solve
from vlp2d_amd import MultiGridSchemeHIP, Options, problems

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _solve(seg, h, f, g, rtol=3e-12, mixed=False):
    s = MultiGridSchemeHIP(seg, seg, h, h, f, options=Options.performance(mixed))
    u0 = np.zeros((seg + 1, seg + 1))
    u0[0, :], u0[-1, :], u0[:, 0], u0[:, -1] = g[0, :], g[-1, :], g[:, 0], g[:, -1]
    s.un0[:] = u0
    s.init_after_boundaries_and_initial_iteration_inited()
    st = s.solve(rtol, 100, raise_on_notconv=False)
    assert st["rel_residual"] <= 10 * rtol, st  # FP64 stagnates near 1e-12 for smooth data (oracle: 1.05e-12 at 1025^2)
    u = s.get_array().copy()
    s.cleanup()
    return u, st


def test_config4_exact_adjoint_finite_differences():
    """Config 4: 1025^2, L = 0.5 ||u - u_obs||^2 over interior nodes with
    u_obs = S(f_true) + N(0, 1e-3), evaluated at f = 0, g = 0.  The adjoint
    gradient (one adjoint multigrid solve) matches central finite differences
    of GPU solves along 64 seeded directions to 1e-6 relative, and the CPU
    FP64 gradient (DST-I solve of A lambda = dL/du, then B^T lambda) to 1e-9."""
    seg, h = 1024, 1.0 / 1024
    f_true, rng = problems.sine_modes_unit(seed=4, seg=seg)
    zero = np.zeros_like(f_true)
    _say("config4: forward u_obs")
    u_true, _ = _solve(seg, h, f_true, zero)
    u_obs = u_true.copy()
    u_obs[1:-1, 1:-1] += 1e-3 * rng.normal((seg - 1) ** 2).reshape(seg - 1, seg - 1)
    f0, g0 = zero.copy(), zero.copy()
    u0, _ = _solve(seg, h, f0, g0)
    ubar = np.zeros_like(u0)
    ubar[1:-1, 1:-1] = (u0 - u_obs)[1:-1, 1:-1]

    def loss(f, g):
        u, st = _solve(seg, h, f, g)
        assert st["rel_residual"] <= 1e-11
        return 0.5 * np.sum((u - u_obs)[1:-1, 1:-1] ** 2)

    _say("config4: adjoint")
    s = MultiGridSchemeHIP(seg, seg, h, h, f0, options=Options.performance())
    fbar, gbar = s.adjoint_exact(ubar, 1e-11, 100)  # FP64 floor of this system: 4.3e-12 (oracle)
    s.cleanup()
    lam, gex = dst_adjoint(ubar, h, h)
    assert np.abs(fbar - lam).max() <= 1e-9 * np.abs(lam).max()
    assert np.abs(gbar - gex).max() <= 1e-9 * np.abs(gex).max()
    assert gbar[0, 0] == gbar[0, -1] == gbar[-1, 0] == gbar[-1, -1] == 0.0
    dirs = problems.SplitMix64(404)
    worst = 0.0
    for k in range(64):
        df = dirs.normal((seg + 1) ** 2).reshape(seg + 1, seg + 1)
        dg = dirs.normal((seg + 1) ** 2).reshape(seg + 1, seg + 1)
        df[0, :] = df[-1, :] = df[:, 0] = df[:, -1] = 0.0
        dg[1:-1, 1:-1] = 0.0
        eps = 1.0  # L is quadratic in (f, g): central differences are exact for any step
        fd = (loss(f0 + eps * df, g0 + eps * dg) - loss(f0 - eps * df, g0 - eps * dg)) / (2 * eps)
        an = np.sum(fbar * df) + np.sum(gbar * dg)
        worst = max(worst, abs(fd - an) / abs(an))
    _say(f"config4: worst relative FD error over 64 directions {worst:.2e}")
    assert worst <= 1e-6


def _dot(a, b):
    """<a, b> over interior nodes, products in FP64, summed in extended
    precision by row blocks (the FP64 rounding of a 2.7e8-term sum would
    otherwise reach 1e-11 relative for terms of random sign)."""
    acc = np.longdouble(0)
    for i0 in range(1, a.shape[0] - 1, 2048):
        i1 = min(i0 + 2048, a.shape[0] - 1)
        acc += (a[i0:i1, 1:-1] * b[i0:i1, 1:-1]).sum(dtype=np.longdouble)
    return float(acc)


def _dot_all(a, b):
    """<a, b> over all nodes (boundary gradients), extended-precision sum."""
    return float((a * b).sum(dtype=np.longdouble))


@pytest.mark.timeout(900)
def test_config5_truncated_adjoint_16385():
    """Config 5: 16385^2, 256 Gaussian bumps, boundary N(0,1) smoothed by the
    9-tap binomial filter.  (a) FP32 V(2,2) cycles in FP64 refinement reach
    rel. residual 1e-10.  (b) The gradient (fbar, gbar) of
    L = 0.5 ||u_K - u_obs||^2 through K = 6 FP64 V(2,2) cycles, from the
    transposed cycles (no forward state stored), satisfies the dot-product
    identity <ubar, J d> = <fbar, df> + <gbar, dg> for 8 seeded directions,
    where J d = Phi^K(u0 = 0; df, dg) is the map's linear part run from a zero
    interior, to 1e-10 relative.  (c) Central differences of L at the data
    point along the same directions, L(x + d) - L(x - d) = 2 <ubar, J d> for
    the quadratic L, evaluated as <ubar, Phi(x + d) - Phi(x - d)> / 2 (the
    same difference without the cancellation of the O(1e8) quadratic term),
    agree with the adjoint to 1e-6 relative.  The peak device memory of the
    handle and the forward and transposed times are reported."""
    import time
    p = problems.config(5)
    seg, h = p.seg_x, p.hx
    _say("config5: mixed solve")
    u, st = _solve(seg, h, p.f, p.g, rtol=1e-10, mixed=True)
    assert st["rel_residual"] <= 1e-10
    _say(f"config5: mixed {st['cycles']} steps, {st['seconds'] * 1e3:.1f} ms, {st['munknowns_per_s']:.0f} Munknowns/s")
    rng = problems.SplitMix64(55)
    u_obs = u + 1e-3 * rng.normal(u.size).reshape(u.shape)
    del u
    K = 6

    def forward(f, g, interior=None):
        s = MultiGridSchemeHIP(seg, seg, h, h, f, options=Options.performance())
        x = np.zeros_like(u_obs) if interior is None else interior.copy()
        x[0, :], x[-1, :], x[:, 0], x[:, -1] = g[0, :], g[-1, :], g[:, 0], g[:, -1]
        s.un0[:] = x
        s.init_after_boundaries_and_initial_iteration_inited()
        t0 = time.perf_counter()
        s.forward_cycles(K)
        dt = time.perf_counter() - t0
        return s, dt

    s, t_fwd = forward(p.f, p.g)
    uK = s.get_array().copy()
    ubar = np.zeros_like(uK)
    ubar[1:-1, 1:-1] = (uK - u_obs)[1:-1, 1:-1]
    del uK
    t0 = time.perf_counter()
    fbar, gbar, _ = s.adjoint_cycles(ubar)
    t_adj = time.perf_counter() - t0
    info = s.info()
    s.cleanup()
    _say(f"config5: K={K} forward {t_fwd * 1e3:.1f} ms, transposed sweep with transfers {t_adj * 1e3:.1f} ms, "
         f"peak device memory {info['device_peak_bytes'] / 2**30:.2f} GiB, {info['checkpoints']} cycles recorded")
    assert info["checkpoints"] == K

    def run(f, g):
        sf, _ = forward(f, g)
        v = sf.get_array().copy()
        sf.cleanup()
        return v

    # seeded directions (NumPy's PCG64: a 2.7e8-node field in about a second)
    dirs = np.random.default_rng(505)
    worst_dot = worst_fd = 0.0
    for k in range(8):
        df = np.zeros_like(p.f)
        df[1:-1, 1:-1] = dirs.standard_normal((seg - 1, seg - 1))
        dg = np.zeros_like(p.f)
        dg[0, :], dg[-1, :] = dirs.standard_normal(seg + 1), dirs.standard_normal(seg + 1)
        dg[1:-1, 0], dg[1:-1, -1] = dirs.standard_normal(seg - 1), dirs.standard_normal(seg - 1)
        an = _dot(fbar, df) + _dot_all(gbar, dg)
        # (b) linear part from a zero interior
        jd = run(df, dg)
        lin = _dot(ubar, jd)
        del jd
        rel_dot = abs(lin - an) / abs(an)
        # (c) central difference at the data point, step 1 (exact for the quadratic L up to rounding)
        up = run(p.f + df, p.g + dg)
        dn = run(p.f - df, p.g - dg)
        up -= dn
        del dn
        fd = 0.5 * _dot(ubar, up)
        del up
        rel_fd = abs(fd - an) / abs(an)
        worst_dot, worst_fd = max(worst_dot, rel_dot), max(worst_fd, rel_fd)
        _say(f"config5: direction {k}: adjoint {an:.15e} linear {lin:.15e} (rel {rel_dot:.2e}) "
             f"central difference {fd:.15e} (rel {rel_fd:.2e})")
    _say(f"config5: worst dot-product identity error {worst_dot:.2e}, worst central-difference error {worst_fd:.2e}")
    assert worst_dot <= 1e-10
    assert worst_fd <= 1e-6


def test_dd128_crosscheck_513():
    """Mixed-precision GPU solution of the 513^2 replica of config 3 against
    the oracle's double-double multigrid (the reference's DD128 CPU path),
    converged to its own floor; FP64 GPU multigrid and the DST-I solve on the
    same inputs sit at double-precision rounding from it."""
    p = problems.config(3, 512)
    o = Oracle(p.seg_x, p.seg_y, p.hx, p.hy, p.f, "dd")
    o.set_u(p.initial("zero"))
    for _ in range(60):
        if o.vcycle(3, 3, 3) < 1e-26:
            break
    hi, lo = o.get_u(with_lo=True)
    scale = np.abs(hi).max()
    um, stm = _solve(p.seg_x, p.hx, p.f, p.g, rtol=1e-10, mixed=True)
    ud, std = _solve(p.seg_x, p.hx, p.f, p.g, rtol=3e-12, mixed=False)
    e_mixed = np.abs((um - hi) - lo).max() / scale
    e_fp64 = np.abs((ud - hi) - lo).max() / scale
    e_dst = np.abs((dst_solve(p.f, p.g, p.hx, p.hy) - hi) - lo).max() / scale
    _say(f"dd128 513^2: mixed rel err {e_mixed:.2e} ({stm['cycles']} steps), fp64 {e_fp64:.2e}, dst {e_dst:.2e}")
    assert e_mixed <= 1e-8
    assert e_fp64 <= 1e-10
    assert e_dst <= 1e-12
