"""GPU set-up and post-processing (vlp2d_amd/csrc/mg_setup.hpp through the C
ABI) against the oracle's restatement (oracle/setup.py), and MultiGrid runs
of the reference's own input files through the C++ host layer against the
analytic solutions and iteration counts the files state."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import setup as osetup
from vlp2d_amd import MultiGridSchemeHIP, Options, _lib, host_lib

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INPUTS = os.path.join(ROOT, "tests", "golden", "inputdata")


def _say(msg):
    print(msg, flush=True)


POLY = "x*x - y*y + 10*x*y + 10*x - 10*y + 1"


def poly(x, y):
    return x * x - y * y + 10 * x * y + 10 * x - 10 * y + 1


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_boundary_and_initial_guess_bitwise(kind):
    """Boundary sampled on the GPU, then each initial guess: the grid equals
    the oracle's bit for bit (polynomial data, FP32 weights as the reference),
    and min / max of the boundary equal the oracle's."""
    sx, sy, hx, hy = 64, 128, 1 / 64, 0.5 / 128
    _say(f"setup kind={kind}")
    s = MultiGridSchemeHIP(sx, sy, hx, hy, None)
    mn, mx = s.set_boundary_expr(boundary=POLY)
    u, omn, omx = osetup.boundary(sx + 1, sy + 1, hx, hy, border=poly)
    assert (mn, mx) == (omn, omx)
    s.init_guess(kind, (mn + mx) / 2 if kind == 0 else 0.0)
    ref = osetup.init_guess(u, kind, (omn + omx) / 2)
    got = s.download()
    assert np.array_equal(got, ref), np.abs(got - ref).max()
    s.cleanup()


def test_edge_functions_and_corner_order():
    sx, sy, h = 32, 32, 1 / 32
    s = MultiGridSchemeHIP(sx, sy, h, h, None)
    mn, mx = s.set_boundary_expr(bottom="10 + x", top="20 + x", left="-1 - y", right="100 + y")
    u, omn, omx = osetup.boundary(sx + 1, sy + 1, h, h, bottom=lambda x: 10 + x, top=lambda x: 20 + x,
                                  left=lambda y: -1 - y, right=lambda y: 100 + y)
    got = s.download()
    assert np.array_equal(got[0, :], u[0, :]) and np.array_equal(got[-1, :], u[-1, :])
    assert np.array_equal(got[:, 0], u[:, 0]) and np.array_equal(got[:, -1], u[:, -1])
    assert (mn, mx) == (omn, omx)
    with pytest.raises(_lib.VLP2DError):
        s.set_boundary_expr(bottom="x", top="x")          # neither all edges nor the boundary function
    with pytest.raises(_lib.VLP2DError):
        s.set_boundary_expr(boundary="x +")               # syntax error
    s.cleanup()


@pytest.mark.parametrize("fexpr,fpy,exact", [
    ("x*x*y - 3*y + 1", lambda x, y: x * x * y - 3 * y + 1, True),
    ("sin(pi(x))*cosh(y) + exp(x*y)", lambda x, y: np.sin(np.pi * x) * np.cosh(y) + np.exp(x * y), False),
])
def test_rhs_sampled_on_gpu_matches_host_sampling(fexpr, fpy, exact):
    """f sampled on the GPU from its expression gives the same solve as f
    sampled on the host: bit for bit for + - * / expressions; within rounding
    where the GPU's libm may differ from the host's by an ulp."""
    sx = sy = 256
    h = 1 / 256
    out = []
    for use_expr in (True, False):
        if use_expr:
            s = MultiGridSchemeHIP(sx, sy, h, h, f_expr=fexpr, options=Options.performance())
        else:
            s = MultiGridSchemeHIP(sx, sy, h, h, lambda x, y: fpy(x, y), options=Options.performance())
        s.set_boundary_expr(boundary=POLY)
        s.init_guess(_lib.INIT_MEAN)
        st = s.solve(1e-10, 40)
        out.append((s.download(), st["cycles"]))
        s.cleanup()
    (ua, ka), (ub, kb) = out
    assert ka == kb
    if exact:
        assert np.array_equal(ua, ub)
    else:
        assert np.abs(ua - ub).max() <= 1e-13 * np.abs(ub).max()


def test_nonfinite_rhs_expression_is_rejected():
    with pytest.raises(_lib.VLP2DError) as e:
        MultiGridSchemeHIP(16, 16, 1 / 16, 1 / 16, f_expr="1/(x - 0.5)")   # x = 0.5 is a node: 1/0
    assert e.value.code == _lib.E_NONFINITE


def test_points_min_max_and_difference():
    """After a solve: min / max over all nodes and u - analytic (Laplace with
    a discrete-harmonic boundary, so the difference is the solver error)
    equal the oracle's, array included."""
    sx = sy = 512
    h = 1 / 512
    s = MultiGridSchemeHIP(sx, sy, h, h, None, options=Options.performance())
    s.set_boundary_expr(boundary=POLY)
    s.init_guess(_lib.INIT_LINEAR)
    s.solve(1e-10, 40)
    u = s.download()
    assert s.points_min_max_device() == osetup.points_min_max(u)
    dmin, dmax, d = s.difference(POLY, with_array=True)
    od, omin, omax = osetup.difference(u, poly, h, h)
    assert np.array_equal(d, od)
    assert (dmin, dmax) == (omin, omax)
    assert max(abs(dmin), abs(dmax)) <= 1e-8
    s.cleanup()


@pytest.mark.parametrize("mixed", [False, True])
def test_maximum_size_32769_on_device(mixed):
    """The largest grid the ABI accepts (32768 x 32768 segments, 1.07e9
    nodes): boundary, initial guess, solve and the deviation from the analytic
    solution all on the GPU (no host copy of the grid).  The boundary is a
    discrete-harmonic polynomial, so the discrete solution is the polynomial
    and the deviation is the solver's error.  The mean initial guess puts the
    initial residual in the boundary layer (||r0|| ~ 3e12), far above the FP64
    rounding floor of the 5-point residual at h = 2^-15, so rel. residual 1e-12
    is reachable (measured: FP64 8 cycles, deviation 1e-9; mixed 6 steps,
    8e-11; 34 and 47 GB on the device)."""
    seg = 32768
    _say(f"maximum size 32769^2 mixed={mixed}")
    s = MultiGridSchemeHIP(seg, seg, 1 / seg, 1 / seg, None,
                           options=Options.performance(mixed, nu1=1, nu2=3, max_inner=2))
    s.set_boundary_expr(boundary=POLY)
    s.init_guess(_lib.INIT_MEAN)
    st = s.solve(1e-12, 40)
    dmin, dmax = s.difference(POLY)
    inf = s.info()
    _say(f"  cycles {st['cycles']} V-cycles {st['vcycles']} rel {st['rel_residual']:.2e} "
         f"{st['seconds'] * 1e3:.1f} ms deviation [{dmin:.2e}, {dmax:.2e}] peak {inf['device_peak_bytes'] / 1e9:.1f} GB")
    assert st["rel_residual"] <= 1e-12 and st["cycles"] <= (10 if mixed else 12)
    assert max(abs(dmin), abs(dmax)) <= 1e-8
    assert (inf["nx"], inf["ny"]) == (32769, 32769) and inf["device_peak_bytes"] < 64e9
    s.cleanup()


@pytest.mark.parametrize("name,kinds", [("Laplace1", [0, 1, 2, 3]), ("Poisson1", [0, 1, 2, 3]),
                                         ("Laplace1_01K", [0, 1, 2, 3])])
def test_reference_input_files(name, kinds):
    """MultiGrid runs of the reference's problems (input files in
    tests/golden/inputdata) through the C++ host layer, in the reference's
    parity mode: segments rounded to powers of two, f and the boundary
    sampled on the GPU, each initial guess, doIteration until the update norm
    is below epsilonDouble.  The number of cycles is exactly the MultiGrid
    iteration count the file states for that initial guess (the reference's
    own runs), and the solution equals the analytic one (both problems have
    quadratic solutions, which the 5-point scheme reproduces exactly) to the
    solver's tolerance."""
    inp = host_lib.parse_input(os.path.join(INPUTS, name + ".txt"))
    for kind in kinds:
        r = host_lib.run_input(inp, kind, 0, None, 100)
        _say(f"{name} init={kind}: {r.seg_x}x{r.seg_y} cycles={r.cycles} elapsed_iters={r.elapsed_iters} "
             f"budget={inp.mg_iters[kind]} metric={r.last_metric:.2e} diff=[{r.dmin:.2e}, {r.dmax:.2e}] "
             f"solve={r.solve_seconds * 1e3:.2f} ms")
        assert r.converged
        assert r.cycles == inp.mg_iters[kind]
        assert max(abs(r.dmin), abs(r.dmax)) <= 1e-9 * max(abs(r.vmin), abs(r.vmax))
    if name == "Laplace1":
        assert (r.seg_x, r.seg_y) == (256, 256)


def test_solve_cli_on_poisson1():
    """The vlp2d_solve executable runs an input file end to end."""
    out = subprocess.run([host_lib.SOLVE_PATH, os.path.join(INPUTS, "Poisson1.txt"), "--init", "wlinear"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rec = json.loads(out.stdout)
    _say(out.stdout.strip())
    assert rec["converged"] and rec["segments"] == [128, 256] and rec["cycles"] <= 10


@pytest.mark.parametrize("args", [["--fmg"], ["--restriction", "full"], ["--smoother", "jacobi:0.8"],
                                  ["--fmg", "--restriction", "full", "--smoother", "jacobi"]])
def test_solve_cli_scheme_options(args):
    """The additions (FMG iterations, full weighting, weighted Jacobi) through
    the executable: the reference's update-norm loop converges (oracle
    counts from the mean guess: FMG 5, full weighting 11, Jacobi 23, all
    three 11, against the reference's 9) and the solution equals the
    quadratic analytic one to the solver's tolerance."""
    out = subprocess.run([host_lib.SOLVE_PATH, os.path.join(INPUTS, "Poisson1.txt"), "--init", "wlinear",
                          "--max-cycles", "100"] + args, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rec = json.loads(out.stdout)
    _say(" ".join(args) + " " + out.stdout.strip())
    assert rec["converged"] and rec["segments"] == [128, 256] and rec["cycles"] <= 40
    assert max(abs(rec["diff_min"]), abs(rec["diff_max"])) <= 1e-9 * max(abs(rec["min"]), abs(rec["max"]))
