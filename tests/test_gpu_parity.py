"""GPU parity: the HIP multigrid (through the C ABI) against the oracle on the
same seeded inputs.  The library evaluates the reference point formulas in the
same order without FMA contraction, so iterates are expected to agree bit for
bit; the bar written in each assert is the one the comparison must meet."""
import numpy as np
import pytest

from oracle import Oracle, adjoint_cycles, dst_adjoint, dst_solve, mixed_solve, plain_solve
from vlp2d_amd import MultiGridSchemeHIP, Options, VLP2DError, problems
from vlp2d_amd import _lib

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _random(seg_x, seg_y, seed, laplace=False):
    rng = np.random.default_rng(seed)
    f = None if laplace else rng.standard_normal((seg_x + 1, seg_y + 1))
    g = np.zeros((seg_x + 1, seg_y + 1))
    g[0, :], g[-1, :] = rng.standard_normal(seg_y + 1), rng.standard_normal(seg_y + 1)
    g[:, 0], g[:, -1] = rng.standard_normal(seg_x + 1), rng.standard_normal(seg_x + 1)
    return f, g


def _mean_init(g):
    u = g.copy()
    b = np.concatenate([g[0, :], g[-1, :], g[:, 0], g[:, -1]])
    u[1:-1, 1:-1] = (b.min() + b.max()) / 2
    return u


PARITY_CASES = [
    # seg_x, seg_y, hx, hy, laplace
    (4, 4, 0.25, 0.25, False),
    (8, 16, 1 / 8, 1 / 16, False),      # unequal steps
    (16, 16, 1 / 16, 1 / 16, True),     # Laplace kernels
    (64, 32, 1 / 64, 1 / 64, False),
    (128, 256, np.e / 128, np.pi / 256, False),  # Poisson1 shape, unequal steps
    (256, 256, 1 / 256, 1 / 256, True),
    (512, 128, 1 / 512, 1 / 512, False),
]


@pytest.mark.parametrize("seg_x,seg_y,hx,hy,laplace", PARITY_CASES)
def test_parity_mode_cycles_bitwise(seg_x, seg_y, hx, hy, laplace):
    """Reference parity mode: V(3,3), half weighting, red-first RB-GS, 3
    coarsest sweeps, Mean initial guess; every iterate and the update norm
    equal the oracle's after each of 6 cycles."""
    _say(f"parity {seg_x}x{seg_y} laplace={laplace}")
    f, g = _random(seg_x, seg_y, seg_x * 7 + seg_y, laplace)
    u0 = _mean_init(g)
    o = Oracle(seg_x, seg_y, hx, hy, f)
    o.set_u(u0)
    s = MultiGridSchemeHIP(seg_x, seg_y, hx, hy, f, eps=1e-13)
    s.un0[:] = u0
    s.init_after_boundaries_and_initial_iteration_inited()
    for k in range(6):
        m_ref = o.vcycle(3, 3, 3)
        s.do_iteration(k)
        ug, ur = s.get_array(), o.get_u()
        assert np.array_equal(ug, ur), (k, np.abs(ug - ur).max())
        assert s.last_metric == m_ref
    s.cleanup()


KERNEL_PATHS = {
    "default": {},
    "fused_all_levels": {"coarse_nodes": -1},        # tiles down to the coarsest level
    "one_workgroup_all": {"coarse_nodes": 1 << 30},  # whole cycle in k_coarse (too big for LDS: L2)
    "small_levels_in_l2": {"flags": _lib.FLAG_COARSE_L2},  # k_coarse instead of k_coarse_lds
    "small_levels_lds_65": {"coarse_nodes": 65 * 65},   # k_coarse_lds over > 64 KB of LDS in FP64
    "store_presmooth": {"flags": _lib.FLAG_STORE_PRESMOOTH},  # ascent reloads the pre-smoothed level
    "stream_small": {"flags": _lib.FLAG_STREAM_SMALL},  # streaming ascent on levels <= 1025^2 (not k_lt_up)
    "no_cycle_fusion": {"flags": _lib.FLAG_NO_CYCLE_FUSION},  # two-cycle steps: level 0 of each cycle apart
    "no_step_fold": {"flags": _lib.FLAG_NO_STEP_FOLD},  # each step's first descent in its own launch
    "no_lt_pair": {"flags": _lib.FLAG_NO_LT_PAIR},  # FP32: one launch per LDS-tiled ascent level
    "one_op_per_launch": {"flags": _lib.FLAG_SIMPLE},
    "no_graph": {"flags": _lib.FLAG_NO_GRAPH},
    "graph_solve": {"flags": _lib.FLAG_GRAPH_SOLVE},  # the solve loop replays captured graphs
    "separate_refine": {"flags": _lib.FLAG_SEPARATE_REFINE},
}


@pytest.mark.parametrize("path", sorted(KERNEL_PATHS))
@pytest.mark.parametrize("seg_x,seg_y,hx,hy,laplace", [(256, 256, 1 / 256, 1 / 256, True),
                                                       (128, 256, np.e / 128, np.pi / 256, False),
                                                       (512, 64, 1 / 512, 1 / 512, False)])
def test_kernel_paths_bitwise(path, seg_x, seg_y, hx, hy, laplace):
    """Every kernel path (streaming tiles, one-workgroup small levels in LDS
    or in L2, one operator per launch, graph or direct launches) gives the
    oracle's iterate bit for bit, in parity mode (V(3,3) + update norm) and
    in the mixed V(2,2) solve with one or (at most) two FP32 cycles per
    refinement step."""
    _say(f"paths {path} {seg_x}x{seg_y}")
    f, g = _random(seg_x, seg_y, 17 + seg_x, laplace)
    u0 = _mean_init(g)
    o = Oracle(seg_x, seg_y, hx, hy, f)
    o.set_u(u0)
    s = MultiGridSchemeHIP(seg_x, seg_y, hx, hy, f, eps=1e-13, options=Options(**KERNEL_PATHS[path]))
    s.un0[:] = u0
    s.init_after_boundaries_and_initial_iteration_inited()
    for k in range(3):
        m_ref = o.vcycle(3, 3, 3)
        s.do_iteration(k)
        assert np.array_equal(s.get_array(), o.get_u()), (k, np.abs(s.get_array() - o.get_u()).max())
        assert s.last_metric == m_ref
    s.cleanup()
    z = np.zeros_like(g) if laplace else f
    for inner in (1, 2):
        ur, kr, _ = mixed_solve(seg_x, seg_y, hx, hy, z, g, 2, 2, 3, 1e-10, 40, inner=inner)
        s = MultiGridSchemeHIP(seg_x, seg_y, hx, hy, z,
                               options=Options.performance(True, max_inner=inner, **KERNEL_PATHS[path]))
        s.un0[:] = g
        s.init_after_boundaries_and_initial_iteration_inited()
        st = s.solve(1e-10, 40)
        assert st["cycles"] == kr, (inner, st["cycles"], kr)
        assert np.array_equal(s.get_array(), ur), (inner, np.abs(s.get_array() - ur).max())
        # a second solve from the restored state (replaying the captured graphs where the path has them)
        s.save_iterate()
        s.restore_iterate()
        s.un0[:] = g
        s.init_after_boundaries_and_initial_iteration_inited()
        st = s.solve(1e-10, 40)
        assert st["cycles"] == kr and np.array_equal(s.get_array(), ur)
        s.cleanup()


@pytest.mark.parametrize("nu2", [1, 2, 3])
@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("seg_x,seg_y,hx,hy,laplace", [(256, 256, 1 / 256, 1 / 256, True),
                                                       (128, 256, np.e / 128, np.pi / 256, False),
                                                       (512, 64, 1 / 512, 1 / 512, False)])
def test_recomputed_presmoothing_bitwise(seg_x, seg_y, hx, hy, laplace, mixed, nu2):
    """V(1,nu2): the descent of a zero-initialised level does not store its
    pre-smoothed iterate and the ascent recomputes it (streaming levels), or
    stores it for the LDS-tiled ascent (levels <= 1025^2); mixed solves also
    with two FP32 cycles per step, whose level-0 kernels between the cycles
    run as one launch (ascent continued into the next descent), and with the
    next step's first descent folded into each step's refinement ascent
    (one, two or three FP32 cycles per step: with three only a step the
    schedule shortens to one or two cycles hands a descent on).  Same
    cycles and the same iterate, bit for bit, as the oracle, as the path
    that stores and reloads every level, as the streaming ascent on every
    level, as separate level-0 launches and as each step's first descent in
    a launch of its own.  V(1,1) is left out on the 128 x 256 grid with
    steps e/128, pi/256 ((hx/hy)^2 = 3): point red-black smoothing diverges
    there (the oracle's residual grows by 1.76 per cycle)."""
    if nu2 == 1 and hx != hy:
        pytest.skip("V(1,1) diverges on the anisotropic grid (oracle included)")
    f, g = _random(seg_x, seg_y, 5 + seg_y, laplace)
    z = np.zeros_like(g) if laplace else f
    for inner in ((1, 2, 3) if mixed else (1,)):
        _say(f"presmooth {seg_x}x{seg_y} mixed={mixed} nu=(1,{nu2}) inner={inner}")
        if mixed:
            ur, kr, _ = mixed_solve(seg_x, seg_y, hx, hy, z, g, 1, nu2, 3, 1e-10, 60, inner=inner)
        else:
            ur, kr, _ = plain_solve(seg_x, seg_y, hx, hy, z, g, 1, nu2, 3, 1e-10, 60)
        for flags in (0, _lib.FLAG_STORE_PRESMOOTH, _lib.FLAG_STREAM_SMALL, _lib.FLAG_NO_CYCLE_FUSION,
                      _lib.FLAG_NO_STEP_FOLD, _lib.FLAG_NO_LT_PAIR):
            s = MultiGridSchemeHIP(seg_x, seg_y, hx, hy, z,
                                   options=Options.performance(mixed, nu1=1, nu2=nu2, flags=flags, max_inner=inner))
            s.un0[:] = g
            s.init_after_boundaries_and_initial_iteration_inited()
            st = s.solve(1e-10, 60)
            assert st["cycles"] == kr
            assert np.array_equal(s.get_array(), ur), (flags, inner, np.abs(s.get_array() - ur).max())
            s.cleanup()


@pytest.mark.parametrize("rtol", [1e-5, 1e-7, 1e-10])
def test_speculative_steps_match_sequential_schedule(rtol):
    """Mixed solves launch the next refinement step before the current one's
    result is known.  With rtol 1e-5 the step launched ahead (two FP32
    cycles) turns out to need one, and is dropped and relaunched; the steps
    accepted and the iterate equal the oracle's sequential schedule bit for
    bit at every tolerance, and equal the profiled (sequential) GPU loop and
    the loop that replays captured graphs."""
    p = problems.config(3, 256)
    _say(f"speculation rtol={rtol}")
    u0 = p.initial("zero")
    ur, kr, hr = mixed_solve(p.seg_x, p.seg_y, p.hx, p.hy, p.f, u0, 1, 3, 3, rtol, 60, inner=2)
    for flags in (0, _lib.FLAG_PROFILE, _lib.FLAG_GRAPH_SOLVE):
        s = MultiGridSchemeHIP(p.seg_x, p.seg_y, p.hx, p.hy, p.f,
                               options=Options.performance(True, nu1=1, nu2=3, max_inner=2, flags=flags))
        s.un0[:] = u0
        s.init_after_boundaries_and_initial_iteration_inited()
        st = s.solve(rtol, 60)
        assert st["cycles"] == kr and st["rel_residual"] <= rtol, (flags, st, kr)
        assert np.array_equal(s.get_array(), ur), (flags, np.abs(s.get_array() - ur).max())
        s.cleanup()


def test_speculative_step_dropped_when_already_solved():
    """f = 0 and g = 0 from a zero guess: the initial residual is 0, the
    step launched ahead is dropped, no step is counted and the iterate stays
    exactly zero."""
    z = np.zeros((65, 65))
    s = MultiGridSchemeHIP(64, 64, 1 / 64, 1 / 64, z, options=Options.performance(True, nu1=1, nu2=3, max_inner=2))
    s.un0[:] = z
    s.init_after_boundaries_and_initial_iteration_inited()
    st = s.solve(1e-10, 20)
    assert st["cycles"] == 0 and st["rel_residual"] == 0.0
    assert np.array_equal(s.get_array(), z) and not np.signbit(s.get_array()).any()
    s.cleanup()


# the fused streaming kernels and one operator per launch
SMOOTHER_PATHS = {"fused": 0, "one_op_per_launch": _lib.FLAG_SIMPLE}


@pytest.mark.parametrize("path", list(SMOOTHER_PATHS))
@pytest.mark.parametrize("omega", [0.8, 2 / 3])
@pytest.mark.parametrize("seg_x,seg_y,hx,hy,laplace", [(64, 32, 1 / 64, 1 / 64, False),
                                                       (128, 256, np.e / 128, np.pi / 256, False),
                                                       (128, 128, 1 / 128, 1 / 128, True)])
def test_weighted_jacobi_bitwise(seg_x, seg_y, hx, hy, laplace, omega, path):
    """Weighted-Jacobi smoother (an addition; the reference smooths with
    red-black Gauss-Seidel only): parity-mode cycles with the update norm,
    and FP64 and mixed V(2,2) solves to 1e-10, bit for bit the oracle's."""
    _say(f"jacobi {seg_x}x{seg_y} omega={omega:.3f} {path}")
    fl = SMOOTHER_PATHS[path]
    f, g = _random(seg_x, seg_y, 11 + seg_x, laplace)
    u0 = _mean_init(g)
    o = Oracle(seg_x, seg_y, hx, hy, f)
    o.set_smoother(1, omega)
    o.set_u(u0)
    s = MultiGridSchemeHIP(seg_x, seg_y, hx, hy, f, eps=1e-13,
                           options=Options(smoother=_lib.SMOOTH_WJACOBI, omega=omega, flags=fl))
    s.un0[:] = u0
    s.init_after_boundaries_and_initial_iteration_inited()
    for k in range(4):
        m_ref = o.vcycle(3, 3, 3)
        s.do_iteration(k)
        assert np.array_equal(s.get_array(), o.get_u()), (k, np.abs(s.get_array() - o.get_u()).max())
        assert s.last_metric == m_ref
    s.cleanup()
    z = np.zeros_like(g) if laplace else f
    for mixed in (False, True):
        solve = mixed_solve if mixed else plain_solve
        # (point Jacobi smooths the anisotropic grid, (hx/hy)^2 = 3, slowly: 109 cycles at omega = 2/3)
        ur, kr, _ = solve(seg_x, seg_y, hx, hy, z, g, 2, 2, 3, 1e-10, 150, smoother=1, omega=omega)
        s = MultiGridSchemeHIP(seg_x, seg_y, hx, hy, z,
                               options=Options.performance(mixed, smoother=_lib.SMOOTH_WJACOBI, omega=omega, flags=fl))
        s.un0[:] = g
        s.init_after_boundaries_and_initial_iteration_inited()
        st = s.solve(1e-10, 150)
        assert st["cycles"] == kr, (mixed, st["cycles"], kr)
        assert np.array_equal(s.get_array(), ur), (mixed, np.abs(s.get_array() - ur).max())
        s.cleanup()


@pytest.mark.parametrize("path", list(SMOOTHER_PATHS))
@pytest.mark.parametrize("smoother", [_lib.SMOOTH_RBGS, _lib.SMOOTH_WJACOBI])
@pytest.mark.parametrize("seg_x,seg_y,hx,hy,laplace", [(64, 32, 1 / 64, 1 / 64, False),
                                                       (256, 128, np.e / 256, np.pi / 128, False),
                                                       (128, 128, 1 / 128, 1 / 128, True)])
def test_full_weighting_bitwise(seg_x, seg_y, hx, hy, laplace, smoother, path):
    """Full-weighting restriction (an addition; the reference restricts by
    half weighting): parity-mode cycles with the update norm, and FP64 and
    mixed V(2,2) solves to 1e-10, bit for bit the oracle's."""
    _say(f"full weighting {seg_x}x{seg_y} smoother={smoother} {path}")
    fl = SMOOTHER_PATHS[path]
    f, g = _random(seg_x, seg_y, 23 + seg_x, laplace)
    u0 = _mean_init(g)
    o = Oracle(seg_x, seg_y, hx, hy, f)
    o.set_restriction(1)
    o.set_smoother(smoother)
    o.set_u(u0)
    s = MultiGridSchemeHIP(seg_x, seg_y, hx, hy, f, eps=1e-13,
                           options=Options(restriction=_lib.RESTRICT_FULL, smoother=smoother, flags=fl))
    s.un0[:] = u0
    s.init_after_boundaries_and_initial_iteration_inited()
    for k in range(4):
        m_ref = o.vcycle(3, 3, 3)
        s.do_iteration(k)
        assert np.array_equal(s.get_array(), o.get_u()), (k, np.abs(s.get_array() - o.get_u()).max())
        assert s.last_metric == m_ref
    s.cleanup()
    z = np.zeros_like(g) if laplace else f
    for mixed in (False, True):
        solve = mixed_solve if mixed else plain_solve
        ur, kr, _ = solve(seg_x, seg_y, hx, hy, z, g, 2, 2, 3, 1e-10, 150, smoother=smoother, restriction=1)
        s = MultiGridSchemeHIP(seg_x, seg_y, hx, hy, z,
                               options=Options.performance(mixed, smoother=smoother, restriction=_lib.RESTRICT_FULL,
                                                           flags=fl))
        s.un0[:] = g
        s.init_after_boundaries_and_initial_iteration_inited()
        st = s.solve(1e-10, 150)
        assert st["cycles"] == kr, (mixed, st["cycles"], kr)
        assert np.array_equal(s.get_array(), ur), (mixed, np.abs(s.get_array() - ur).max())
        s.cleanup()


@pytest.mark.parametrize("smoother,restriction", [(_lib.SMOOTH_WJACOBI, _lib.RESTRICT_HALF),
                                                  (_lib.SMOOTH_RBGS, _lib.RESTRICT_FULL),
                                                  (_lib.SMOOTH_WJACOBI, _lib.RESTRICT_FULL)])
@pytest.mark.parametrize("nu", [(1, 3), (2, 2), (3, 1)])
def test_fused_smoother_restriction_large(smoother, restriction, nu):
    """Weighted Jacobi and full weighting in the fused streaming kernels on a
    grid of many tiles per level (2048 x 1024 segments, equal steps):
    FP64 and mixed (one and two FP32 cycles per refinement step) solves to
    1e-10, cycle counts and iterates bit for bit the oracle's."""
    seg_x, seg_y = 2048, 1024
    hx = hy = 1 / 2048
    _say(f"fused {seg_x}x{seg_y} smoother={smoother} restriction={restriction} V{nu}")
    f, g = _random(seg_x, seg_y, 5 + 3 * smoother + restriction, False)
    for mixed, inner in ((False, 1), (True, 1), (True, 2)):
        solve = mixed_solve if mixed else plain_solve
        kw = {"inner": inner} if mixed else {}
        ur, kr, _ = solve(seg_x, seg_y, hx, hy, f, g, nu[0], nu[1], 3, 1e-10, 200, smoother=smoother,
                          restriction=restriction, **kw)
        s = MultiGridSchemeHIP(seg_x, seg_y, hx, hy, f,
                               options=Options.performance(mixed, nu1=nu[0], nu2=nu[1], smoother=smoother,
                                                           restriction=restriction, max_inner=inner))
        s.un0[:] = g
        s.init_after_boundaries_and_initial_iteration_inited()
        st = s.solve(1e-10, 200)
        assert st["cycles"] == kr, (mixed, inner, st["cycles"], kr)
        assert np.array_equal(s.get_array(), ur), (mixed, inner, np.abs(s.get_array() - ur).max())
        s.cleanup()


@pytest.mark.parametrize("restriction", [_lib.RESTRICT_HALF, _lib.RESTRICT_FULL])
@pytest.mark.parametrize("seg_x,seg_y,hx,hy,laplace", [(64, 32, 1 / 64, 1 / 64, False),
                                                       (256, 128, np.e / 256, np.pi / 128, False),
                                                       (128, 128, 1 / 128, 1 / 128, True)])
def test_fmg_bitwise(seg_x, seg_y, hx, hy, laplace, restriction):
    """FMG iterations (the reference's unused FMG(), MultiGridScheme.cs:117-132):
    parity-mode iterations with the update norm, and an FP64 solve to 1e-10,
    bit for bit the oracle's."""
    _say(f"fmg {seg_x}x{seg_y} restriction={restriction}")
    f, g = _random(seg_x, seg_y, 31 + seg_x, laplace)
    u0 = _mean_init(g)
    o = Oracle(seg_x, seg_y, hx, hy, f)
    o.set_cycle(1)
    o.set_restriction(restriction)
    o.set_u(u0)
    s = MultiGridSchemeHIP(seg_x, seg_y, hx, hy, f, eps=1e-13,
                           options=Options(cycle=_lib.CYCLE_FMG, restriction=restriction))
    s.un0[:] = u0
    s.init_after_boundaries_and_initial_iteration_inited()
    for k in range(3):
        m_ref = o.vcycle(3, 3, 3)
        s.do_iteration(k)
        assert np.array_equal(s.get_array(), o.get_u()), (k, np.abs(s.get_array() - o.get_u()).max())
        assert s.last_metric == m_ref
    s.cleanup()
    z = np.zeros_like(g) if laplace else f
    ur, kr, _ = plain_solve(seg_x, seg_y, hx, hy, z, g, 2, 2, 3, 1e-10, 60, restriction=restriction, cycle=1)
    s = MultiGridSchemeHIP(seg_x, seg_y, hx, hy, z,
                           options=Options.performance(cycle=_lib.CYCLE_FMG, restriction=restriction))
    s.un0[:] = g
    s.init_after_boundaries_and_initial_iteration_inited()
    st = s.solve(1e-10, 60)
    assert st["cycles"] == kr and st["vcycles"] == 2 * kr, (st["cycles"], kr)
    assert np.array_equal(s.get_array(), ur), np.abs(s.get_array() - ur).max()
    s.cleanup()


def test_config1_reference_stop_rule_cycle_count():
    """Config 1 (257^2 Laplace, g = x^2 - y^2) through the reference loop:
    cycles until the update norm <= eps, same count as the oracle, and the
    discrete solution equals the harmonic quadratic."""
    p = problems.config(1)
    _say("config1 257^2 reference loop")
    o = Oracle(p.seg_x, p.seg_y, p.hx, p.hy, None)
    o.set_u(p.initial("mean"))
    s = MultiGridSchemeHIP(p.seg_x, p.seg_y, p.hx, p.hy, None, eps=1e-13)
    s.un0[:] = p.initial("mean")
    s.init_after_boundaries_and_initial_iteration_inited()
    n_ref = n_gpu = 0
    while n_ref < 40 and o.vcycle(3, 3, 3) > 1e-13:
        n_ref += 1
    while n_gpu < 40 and s.do_iteration(n_gpu) >= 1e-13:
        n_gpu += 1
    assert n_gpu == n_ref
    assert np.array_equal(s.get_array(), o.get_u())
    assert np.abs(s.get_array() - p.exact).max() < 1e-12
    s.cleanup()


@pytest.mark.parametrize("nu", [(2, 2), (1, 3), (3, 1)])
@pytest.mark.parametrize("mixed,inner", [(False, 1), (True, 1), (True, 2), (True, 3)])
@pytest.mark.parametrize("cfg,seg", [(2, 512), (3, 256), (1, 128), (3, 1024)])
def test_performance_mode_matches_oracle(cfg, seg, mixed, inner, nu):
    """V(nu1,nu2) to relative residual 1e-10 from a zero interior (V(2,2), and
    the benchmark's V(1,3)), FP64 or FP32 cycles in FP64 refinement (at most
    `inner` FP32 cycles per refinement step, chosen by the predicted
    contraction): same number of steps as the oracle and the same iterate
    (bitwise), which is the DST-I solution to solver accuracy."""
    p = problems.config(cfg, seg)
    _say(f"perf cfg{cfg} seg{seg} mixed={mixed} inner={inner} nu={nu}")
    u0 = p.initial("zero")
    if mixed:
        ur, kr, hr = mixed_solve(p.seg_x, p.seg_y, p.hx, p.hy, p.f, u0, nu[0], nu[1], 3, 1e-10, 60, inner=inner)
    else:
        ur, kr, hr = plain_solve(p.seg_x, p.seg_y, p.hx, p.hy, p.f, u0, nu[0], nu[1], 3, 1e-10, 60)
    s = MultiGridSchemeHIP(p.seg_x, p.seg_y, p.hx, p.hy, p.f,
                           options=Options.performance(mixed, nu1=nu[0], nu2=nu[1], max_inner=inner))
    s.un0[:] = u0
    s.init_after_boundaries_and_initial_iteration_inited()
    st = s.solve(1e-10, 60)
    assert st["cycles"] == kr
    assert st["rel_residual"] <= 1e-10
    assert abs(st["rel_residual"] - hr[-1]) <= 1e-6 * hr[-1]
    ug = s.get_array()
    assert np.array_equal(ug, ur), np.abs(ug - ur).max()
    ex = dst_solve(p.f, p.g, p.hx, p.hy)
    assert np.abs(ug - ex).max() <= 1e-8 * np.abs(ex).max()
    s.cleanup()


@pytest.mark.parametrize("mixed,inner,nu", [(False, 1, (2, 2)), (True, 2, (2, 2)), (False, 1, (1, 3))])
@pytest.mark.parametrize("cfg,seg", [(2, None), (1, 256)])
def test_full_size_configs_to_rtol_match_oracle(cfg, seg, mixed, inner, nu):
    """Config 2 at its full size (4097 x 2049: [0,2] x [0,1], sine-mode f,
    cubic side boundaries) and config 1 at its full size (257^2 Laplace,
    g = x^2 - y^2) solved to relative residual 1e-10 from a zero interior with
    the configs' V(2,2) cycle (and the benchmark's V(1,3)): the same number of
    cycles or refinement steps as the oracle (MultiGridScheme.cs:63-71 looped
    to the residual stop) and the same iterate, bit for bit."""
    p = problems.config(cfg, seg)
    _say(f"full-size cfg{cfg} {p.seg_x + 1}x{p.seg_y + 1} mixed={mixed} nu={nu}")
    u0 = p.initial("zero")
    if mixed:
        ur, kr, hr = mixed_solve(p.seg_x, p.seg_y, p.hx, p.hy, p.f, u0, nu[0], nu[1], 3, 1e-10, 60, inner=inner)
    else:
        ur, kr, hr = plain_solve(p.seg_x, p.seg_y, p.hx, p.hy, p.f, u0, nu[0], nu[1], 3, 1e-10, 60)
    s = MultiGridSchemeHIP(p.seg_x, p.seg_y, p.hx, p.hy, p.f,
                           options=Options.performance(mixed, nu1=nu[0], nu2=nu[1], max_inner=inner))
    s.un0[:] = u0
    s.init_after_boundaries_and_initial_iteration_inited()
    st = s.solve(1e-10, 60)
    ug = s.get_array().copy()
    s.cleanup()
    _say(f"  {st['cycles']} steps ({kr} oracle), rel {st['rel_residual']:.3e}")
    assert st["cycles"] == kr
    assert st["rel_residual"] <= 1e-10
    assert np.array_equal(ug, ur), np.abs(ug - ur).max()
    if cfg == 1:  # the discrete solution of the harmonic quadratic is the quadratic itself
        assert np.abs(ug - p.exact).max() <= 1e-9


def test_config2_full_size_parity():
    """Config 2 at its full size (4097 x 2049, coarsest 5 x 3): FP64 GPU vs
    the oracle's double multigrid at equal cycle counts (V(3,3), Mean init).
    Bar from the specification: max|du| <= 1e-12 max|u|; bitwise expected."""
    p = problems.config(2)
    _say("config2 4097x2049 parity")
    u0 = p.initial("mean")
    o = Oracle(p.seg_x, p.seg_y, p.hx, p.hy, p.f)
    o.set_u(u0)
    s = MultiGridSchemeHIP(p.seg_x, p.seg_y, p.hx, p.hy, p.f, eps=1e-13)
    s.un0[:] = u0
    s.init_after_boundaries_and_initial_iteration_inited()
    for k in range(4):
        mr = o.vcycle(3, 3, 3)
        s.do_iteration(k)
        assert s.last_metric == mr
    ug, ur = s.get_array(), o.get_u()
    assert np.abs(ug - ur).max() <= 1e-12 * np.abs(ur).max()
    assert np.array_equal(ug, ur)
    s.cleanup()


def test_mixed_full_size_8193():
    """Config 3 at full size (8193^2): FP32 V(2,2) cycles inside FP64
    refinement reach relative residual 1e-10 (size-independent property),
    and agree with the plain FP64 GPU multigrid solution to the accuracy the
    residual bound implies."""
    p = problems.config(3)
    _say("config3 8193^2 mixed vs fp64")
    out = {}
    for mixed in (True, False):
        s = MultiGridSchemeHIP(p.seg_x, p.seg_y, p.hx, p.hy, p.f, options=Options.performance(mixed))
        s.un0[:] = p.initial("zero")
        s.init_after_boundaries_and_initial_iteration_inited()
        st = s.solve(1e-10, 40)
        assert st["rel_residual"] <= 1e-10
        out[mixed] = (s.get_array().copy(), st)
        _say(f"  mixed={mixed} cycles={st['cycles']} t={st['seconds']:.4f}s")
        s.cleanup()
    d = np.abs(out[True][0] - out[False][0]).max()
    assert d <= 1e-7 * np.abs(out[False][0]).max()


@pytest.mark.parametrize("seg,K", [(4096, 2), (8192, 2), (16384, 1)])
def test_forward_cycles_bitwise_full_sizes(seg, K):
    """FP64 V(2,2) cycles at the benchmark sizes (config 3 and config 5
    shapes) through the fused kernels: the iterate after K cycles is the
    oracle's bit for bit (the map the adjoint transposes)."""
    from oracle import forward_cycles
    p = problems.config(3, seg)
    _say(f"forward cycles {seg + 1}^2 K={K}")
    u0 = p.initial("zero")
    ur = forward_cycles(p.seg_x, p.seg_y, p.hx, p.hy, p.f, u0, K, 2, 2, 3)
    s = MultiGridSchemeHIP(p.seg_x, p.seg_y, p.hx, p.hy, p.f, options=Options.performance())
    s.un0[:] = u0
    s.init_after_boundaries_and_initial_iteration_inited()
    s.forward_cycles(K)
    ug = s.get_array()
    d = np.abs(ug - ur)
    bad = np.argwhere(d > 0)
    assert bad.shape[0] == 0, (bad.shape[0], bad[:10].tolist(), float(d.max()))
    s.cleanup()


# ------------------------------------------------------------------ adjoint
@pytest.mark.parametrize("seg_x,seg_y,hx,hy,K,nu", [(16, 16, 1 / 16, 1 / 16, 2, (2, 2)),
                                                   (64, 32, 0.01, 0.02, 3, (2, 2)),
                                                   (1024, 1024, 1 / 1024, 1 / 1024, 2, (2, 2)),
                                                   (256, 512, 1 / 256, 1 / 512, 2, (1, 3)),
                                                   (512, 128, 0.002, 0.007, 3, (3, 1)),
                                                   (128, 128, 1 / 128, 1 / 128, 2, (1, 1)),
                                                   (256, 256, 1 / 256, 1 / 256, 1, (3, 3))])
def test_adjoint_cycles_match_oracle(seg_x, seg_y, hx, hy, K, nu):
    """The transposed sweep (the streaming k_tdown/k_tup pair on every level
    above the coarsest, one operator per launch there) against the oracle's
    transposed cycles, bit for bit, for every sweep count the streaming
    kernels take."""
    _say(f"adjoint cycles {seg_x}x{seg_y} K={K} nu={nu}")
    f, g = _random(seg_x, seg_y, 99)
    rng = np.random.default_rng(5)
    ubar = rng.standard_normal(g.shape)
    s = MultiGridSchemeHIP(seg_x, seg_y, hx, hy, f, options=Options.performance(nu1=nu[0], nu2=nu[1]))
    s.un0[:] = g
    s.init_after_boundaries_and_initial_iteration_inited()
    s.forward_cycles(K)
    fb, gb, u0b = s.adjoint_cycles(ubar)
    rf, rg, ru = adjoint_cycles(seg_x, seg_y, hx, hy, ubar, K, nu[0], nu[1], 3)
    assert np.array_equal(fb, rf) and np.array_equal(gb, rg) and np.array_equal(u0b, ru)
    s.cleanup()


def test_adjoint_cycles_streaming_matches_one_operator_path():
    """The streaming transposed cycle and the one-operator path
    (VLP2D_FLAG_SIMPLE) give the same bits at 2049 x 1025."""
    seg_x, seg_y, hx, hy, K = 2048, 1024, 1 / 2048, 1 / 1024, 2
    _say(f"adjoint cycles streaming vs one-operator {seg_x}x{seg_y}")
    f, g = _random(seg_x, seg_y, 7)
    ubar = np.random.default_rng(6).standard_normal(g.shape)
    out = []
    for flags in (0, _lib.FLAG_SIMPLE):
        s = MultiGridSchemeHIP(seg_x, seg_y, hx, hy, f, options=Options.performance(flags=flags))
        s.un0[:] = g
        s.init_after_boundaries_and_initial_iteration_inited()
        s.forward_cycles(K)
        out.append(s.adjoint_cycles(ubar))
        s.cleanup()
    for a, b in zip(*out):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("kw", [{}, dict(smoother=_lib.SMOOTH_WJACOBI), dict(restriction=_lib.RESTRICT_FULL),
                                dict(cycle=_lib.CYCLE_FMG),
                                dict(cycle=_lib.CYCLE_FMG, restriction=_lib.RESTRICT_FULL, smoother=_lib.SMOOTH_WJACOBI)])
def test_adjoint_exact_matches_dst_gradient(kw):
    """The exact adjoint solves A lambda = dL/du with the handle's own cycle
    (any smoother, restriction or FMG): the gradient is the DST solution's."""
    seg, h = 128, 1 / 128
    _say(f"adjoint exact 129^2 {kw}")
    rng = np.random.default_rng(8)
    ubar = np.zeros((seg + 1, seg + 1))
    ubar[1:-1, 1:-1] = rng.standard_normal((seg - 1, seg - 1))
    s = MultiGridSchemeHIP(seg, seg, h, h, np.zeros((seg + 1, seg + 1)), options=Options.performance(**kw))
    fb, gb = s.adjoint_exact(ubar, 1e-12, 60)
    lam, gex = dst_adjoint(ubar, h, h)
    assert np.abs(fb - lam).max() <= 1e-9 * np.abs(lam).max()
    assert np.abs(gb - gex).max() <= 1e-9 * np.abs(gex).max()
    s.cleanup()


@pytest.mark.parametrize("seg_x,seg_y", [(4, 2048), (2048, 4), (4, 8192)])
def test_extreme_aspect_ratios_bitwise(seg_x, seg_y):
    """Two-level hierarchies of extreme shape: one column strip or one row
    tile at level 0, and a coarsest level of 3 x 1025 (LDS above 64 KB), 1025
    x 3 or 3 x 4097 nodes (too large for LDS: the global-memory coarse
    kernel).  Parity-mode cycles, and 4 FP64 cycles and 3 mixed steps at
    rtol 0, bit for bit the oracle's."""
    _say(f"aspect {seg_x}x{seg_y}")
    hx, hy = 1 / seg_x, 1 / seg_y
    f, g = _random(seg_x, seg_y, 51 + seg_x)
    u0 = _mean_init(g)
    o = Oracle(seg_x, seg_y, hx, hy, f)
    o.set_u(u0)
    s = MultiGridSchemeHIP(seg_x, seg_y, hx, hy, f, eps=1e-13)
    s.un0[:] = u0
    s.init_after_boundaries_and_initial_iteration_inited()
    for k in range(3):
        m_ref = o.vcycle(3, 3, 3)
        s.do_iteration(k)
        assert np.array_equal(s.get_array(), o.get_u()), (k, np.abs(s.get_array() - o.get_u()).max())
        assert s.last_metric == m_ref
    s.cleanup()
    for mixed, n in ((False, 4), (True, 3)):
        if mixed:
            ur, kr, _ = mixed_solve(seg_x, seg_y, hx, hy, f, g, 2, 2, 3, 0.0, n, inner=2)
        else:
            ur, kr, _ = plain_solve(seg_x, seg_y, hx, hy, f, g, 2, 2, 3, 0.0, n)
        s = MultiGridSchemeHIP(seg_x, seg_y, hx, hy, f, options=Options.performance(mixed, max_inner=2))
        s.un0[:] = g
        s.init_after_boundaries_and_initial_iteration_inited()
        st = s.solve(0.0, n, raise_on_notconv=False)
        assert st["rc"] == _lib.E_NOTCONV and st["cycles"] == kr == n
        assert np.array_equal(s.get_array(), ur), (mixed, np.abs(s.get_array() - ur).max())
        s.cleanup()


@pytest.mark.parametrize("mixed", [False, True])
def test_divergent_solve_reports_nonfinite(mixed):
    """Point Gauss-Seidel V(2,2) diverges on a strongly anisotropic grid
    ((hy/hx)^2 = 16; the oracle diverges as well, FP64 by about 4x per cycle,
    the mixed steps faster until FP32 overflows at step 30): the solve ends
    with VLP2D_E_NONFINITE once the residual is no longer finite, rather than
    running on or returning garbage."""
    _say(f"divergent solve mixed={mixed}")
    f, g = _random(256, 128, 43)
    s = MultiGridSchemeHIP(256, 128, 1 / 256, 2 / 128, f, options=Options.performance(mixed, max_inner=2))
    s.un0[:] = g
    s.init_after_boundaries_and_initial_iteration_inited()
    with pytest.raises(VLP2DError) as e:
        s.solve(1e-10, 1000)
    assert e.value.code == _lib.E_NONFINITE
    s.cleanup()
    # the reference's own loop (one doIteration per call, update-norm metric)
    # on the same diverging problem ends the same way once the update
    # overflows to Inf or NaN
    s = MultiGridSchemeHIP(256, 128, 1 / 256, 2 / 128, f, options=Options.performance(mixed))
    s.un0[:] = g
    s.init_after_boundaries_and_initial_iteration_inited()
    with pytest.raises(VLP2DError) as e:
        for it in range(3000):
            s.do_iteration(it)
    assert e.value.code == _lib.E_NONFINITE
    _say(f"divergent cycles mixed={mixed}: non-finite update after {it + 1} cycles")
    s.cleanup()


# ------------------------------------------------------------ independent handles
def _handle_cases():
    """(name, seg_x, seg_y, hx, hy, f, g, mixed, parity_cycles) with the
    oracle's result of each: a parity-mode run (doIteration, graph replay),
    an FP64 solve and a two-cycle mixed solve of different shapes."""
    cases = []
    for name, sx, sy, seed, mixed, parity in (("parity", 64, 128, 41, False, 3), ("fp64", 128, 256, 42, False, 0),
                                              ("mixed", 256, 128, 43, True, 0)):
        f, g = _random(sx, sy, seed)
        hx = hy = 1 / 256  # equal steps: point Gauss-Seidel V(2,2) diverges on strongly anisotropic grids
        if parity:
            o = Oracle(sx, sy, hx, hy, f)
            o.set_u(_mean_init(g))
            for _ in range(parity):
                o.vcycle(3, 3, 3)
            ref = o.get_u()
            o.close()
        elif mixed:
            ref, _, _ = mixed_solve(sx, sy, hx, hy, f, g, 2, 2, 3, 1e-10, 40, inner=2)
        else:
            ref, _, _ = plain_solve(sx, sy, hx, hy, f, g, 2, 2, 3, 1e-10, 40)
        cases.append((name, sx, sy, hx, hy, f, g, mixed, parity, ref))
    return cases


def _run_case(case):
    name, sx, sy, hx, hy, f, g, mixed, parity, ref = case
    if parity:
        s = MultiGridSchemeHIP(sx, sy, hx, hy, f, eps=1e-13)
        s.un0[:] = _mean_init(g)
        s.init_after_boundaries_and_initial_iteration_inited()
        for k in range(parity):
            s.do_iteration(k)
    else:
        s = MultiGridSchemeHIP(sx, sy, hx, hy, f, options=Options.performance(mixed, max_inner=2 if mixed else 1))
        s.un0[:] = g
        s.init_after_boundaries_and_initial_iteration_inited()
        s.solve(1e-10, 40)
    u = s.get_array().copy()
    s.cleanup()
    return np.array_equal(u, ref)


def test_independent_handles_interleaved_and_threaded():
    """Handles share no state: three live handles with their calls
    interleaved, then two host threads each running its own handles at the
    same time (one in graph-replayed parity mode), all bit for bit the
    oracle's results."""
    import threading
    cases = _handle_cases()
    _say("handles interleaved")
    hs = []
    for name, sx, sy, hx, hy, f, g, mixed, parity, ref in cases:
        opts = Options() if parity else Options.performance(mixed, max_inner=2 if mixed else 1)
        s = MultiGridSchemeHIP(sx, sy, hx, hy, f, eps=1e-13, options=opts)
        s.un0[:] = _mean_init(g) if parity else g
        s.init_after_boundaries_and_initial_iteration_inited()
        hs.append(s)
    for k in range(3):  # parity cycles between the other handles' calls
        hs[0].do_iteration(k)
        if k == 0:
            hs[1].solve(1e-10, 40)
        if k == 1:
            hs[2].solve(1e-10, 40)
    for s, case in zip(hs, cases):
        assert np.array_equal(s.get_array(), case[-1]), case[0]
        s.cleanup()
    _say("handles on two threads")
    ok = {}

    def worker(tid):
        try:
            ok[tid] = [_run_case(cases[(tid + r) % len(cases)]) for r in range(4)]
        except Exception as e:  # reported by the main thread
            ok[tid] = [repr(e)]

    ts = [threading.Thread(target=worker, args=(t,)) for t in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=100)
    assert not any(t.is_alive() for t in ts), "worker thread did not finish"
    assert ok == {0: [True] * 4, 1: [True] * 4}, ok


@pytest.mark.parametrize("mixed", [False, True])
def test_cancel_from_another_thread_ends_a_running_solve(mixed):
    """vlp2d_mg_cancel from another thread while vlp2d_mg_solve runs (rtol 0,
    so only the cancel can end it before max_cycles): the solve returns
    VLP2D_E_CANCELLED at its next step boundary, and after a new upload the
    handle solves normally."""
    import threading
    import time
    p = problems.config(3, 4096)
    _say(f"cancel from a thread, 4097^2 mixed={mixed}")
    s = MultiGridSchemeHIP(p.seg_x, p.seg_y, p.hx, p.hy, p.f, options=Options.performance(mixed, max_inner=2))
    s.un0[:] = p.initial("zero")
    s.init_after_boundaries_and_initial_iteration_inited()
    s.solve(1e-10, 50)  # warm: programs and scratch set up
    s.init_after_boundaries_and_initial_iteration_inited()
    out = {}

    def run():
        t0 = time.perf_counter()
        try:
            out["st"] = s.solve(0.0, 20000, raise_on_notconv=False)
        except VLP2DError as e:
            out["code"] = e.code
        out["t"] = time.perf_counter() - t0

    th = threading.Thread(target=run)
    th.start()
    time.sleep(0.3)
    t_cancel = time.perf_counter()
    s.cancel_iterations()
    th.join(timeout=60)
    assert not th.is_alive()
    assert out.get("code") == _lib.E_CANCELLED, out
    assert 0.25 < out["t"] < 0.3 + (time.perf_counter() - t_cancel) + 0.5, out
    s.init_after_boundaries_and_initial_iteration_inited()  # upload clears the request
    st = s.solve(1e-10, 50)
    assert st["rel_residual"] <= 1e-10 and st["cycles"] <= 8
    s.cleanup()


# ------------------------------------------------------------ error handling
def test_errors_are_reported_not_faulted():
    f, g = _random(16, 16, 3)
    s = MultiGridSchemeHIP(16, 16, 1 / 16, 1 / 16, f, options=Options.performance())
    bad = g.copy()
    bad[0, 3] = np.inf
    with pytest.raises(VLP2DError) as e:
        _lib.check(s.lib, s.h, s.lib.vlp2d_mg_upload(s.h, _lib.ptr(bad)))
    assert e.value.code == _lib.E_NONFINITE
    s.un0[:] = g
    s.init_after_boundaries_and_initial_iteration_inited()
    st = s.solve(1e-14, 1, raise_on_notconv=False)
    assert st["rc"] == _lib.E_NOTCONV and st["cycles"] == 1
    s.cancel_iterations()
    with pytest.raises(VLP2DError) as e:
        s.solve(1e-14, 50)
    assert e.value.code == _lib.E_CANCELLED
    s.init_after_boundaries_and_initial_iteration_inited()  # upload clears the request
    assert s.solve(1e-10, 50)["rel_residual"] <= 1e-10
    with pytest.raises(VLP2DError) as e:
        s.adjoint_cycles(g)
    assert e.value.code == _lib.E_STATE
    s.cleanup()
    # the truncated-cycle adjoint transposes the Gauss-Seidel, half-weighting cycle only
    for kw in (dict(smoother=_lib.SMOOTH_WJACOBI), dict(restriction=_lib.RESTRICT_FULL), dict(cycle=_lib.CYCLE_FMG)):
        s = MultiGridSchemeHIP(16, 16, 1 / 16, 1 / 16, f, options=Options.performance(**kw))
        s.un0[:] = g
        s.init_after_boundaries_and_initial_iteration_inited()
        with pytest.raises(VLP2DError) as e:
            s.forward_cycles(2)
        assert e.value.code == _lib.E_STATE
        s.cleanup()
