"""Expressions and input files on the CPU: the compiler / host evaluator
behind the library's expression entry points (vlp2d_expr_eval) against
Python's arithmetic, its grammar rules (the reference's operator table,
Substitutor/OperationsRegistry.cs:64-84), and the host layer's parser on the
reference's own input files (tests/golden/inputdata)."""
import ctypes
import math
import os

import numpy as np
import pytest

from vlp2d_amd import _lib
from vlp2d_amd._lib import load_library
from vlp2d_amd import host_lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INPUTS = os.path.join(ROOT, "tests", "golden", "inputdata")


def ev(expr, x, y):
    lib = load_library()
    x = np.ascontiguousarray(np.atleast_1d(np.asarray(x, dtype=np.float64)))
    y = np.ascontiguousarray(np.broadcast_to(np.asarray(y, dtype=np.float64), x.shape).copy())
    out = np.empty_like(x)
    rc = lib.vlp2d_expr_eval(expr.encode(), _lib.ptr(x), _lib.ptr(y), x.size, _lib.ptr(out))
    if rc != 0:
        raise _lib.VLP2DError(rc, lib.vlp2d_mg_last_error(None).decode())
    return out


PTS = np.random.default_rng(3).uniform(-2.0, 2.0, (2, 64))


@pytest.mark.parametrize("expr,py", [
    ("x^2 - y^2 + 10 * x * y + 10 * x - 10 * y + 1", lambda x, y: x * x - y * y + 10 * x * y + 10 * x - 10 * y + 1),
    ("(x-exp(1)/2)^2 + (y-pi(1)/2)^2", lambda x, y: (x - math.e / 2) * (x - math.e / 2) + (y - math.pi / 2) * (y - math.pi / 2)),
    ("-4", lambda x, y: -4.0 + 0 * x),
    ("-x^2", lambda x, y: -(x * x)),
    ("x*y/(1+x*x)", lambda x, y: x * y / (1 + x * x)),
    ("x^3", lambda x, y: (x * x) * x),      # a^n: square-and-multiply, x^3 = (x*x)*x
    ("x^4", lambda x, y: (x * x) * (x * x)),
    ("x^-2", lambda x, y: 1.0 / (x * x)),
    ("2^-1 * x", lambda x, y: 0.5 * x),
    ("x ? y : 2", lambda x, y: np.where(x != 0, y, 2.0)),
])
def test_expr_matches_python_exactly(expr, py):
    """+ - * / and integer powers are evaluated in the written order and match
    the same sequence of IEEE operations bit for bit."""
    x, y = PTS
    assert np.array_equal(ev(expr, x, y), py(x, y))


@pytest.mark.parametrize("expr,py", [
    ("sin(pi(x))*sinh(pi(y))/sinh(pi(1))", lambda x, y: np.sin(np.pi * x) * np.sinh(np.pi * y) / np.sinh(np.pi)),
    ("exp(-x*x)*cos(3*y) + sqrt(abs(x)) - ln(1+y*y) + log10(2+x)",
     lambda x, y: np.exp(-x * x) * np.cos(3 * y) + np.sqrt(np.abs(x)) - np.log(1 + y * y) + np.log10(2 + x)),
    ("tan(x/4) + atan(y) + asin(x/3) + acos(y/3) + tanh(x) + cosh(y)",
     lambda x, y: np.tan(x / 4) + np.arctan(y) + np.arcsin(x / 3) + np.arccos(y / 3) + np.tanh(x) + np.cosh(y)),
    ("abs(x)^1.5", lambda x, y: np.abs(x) ** 1.5),
])
def test_expr_functions_close_to_numpy(expr, py):
    x, y = PTS
    r, e = ev(expr, x, y), py(x, y)
    assert np.allclose(r, e, rtol=4e-16 * 8, atol=1e-15)


def test_operator_table_of_the_reference():
    """^ binds tighter than unary minus and is left-associative (priority 1,
    LeftAssociated); unary minus precedes * and /; pi(a) = pi * a."""
    assert ev("2^3^2", 0, 0)[0] == 64.0                # (2^3)^2, not 2^(3^2)
    assert ev("-2^2", 0, 0)[0] == -4.0                 # -(2^2)
    assert ev("2*-3", 0, 0)[0] == -6.0
    assert ev("1 - 2 - 3", 0, 0)[0] == -4.0            # left-associative
    assert ev("8 / 2 / 2", 0, 0)[0] == 2.0
    assert ev("pi(2)", 0, 0)[0] == 2 * math.pi
    assert ev("0 ? 1 : 0 ? 2 : 3", 0, 0)[0] == 3.0     # right-associative conditional
    assert ev("1.5E-3 + .25", 0, 0)[0] == 1.5e-3 + 0.25


@pytest.mark.parametrize("bad", ["x +", "(x", "foo(1)", "x $ y", "sin x", ""])
def test_expr_errors_are_reported(bad):
    with pytest.raises(_lib.VLP2DError) as e:
        ev(bad, 0, 0)
    assert e.value.code == _lib.E_ARG


def test_input_files_parse():
    """Every reference input file parses; the values are those written in it
    (xMax / yMax substituted into the function texts as InputParser.cs:107-114
    does)."""
    names = sorted(f for f in os.listdir(INPUTS) if f.endswith(".txt"))
    assert len(names) == 6
    for f in names:
        p = host_lib.parse_input(os.path.join(INPUTS, f))
        assert p.name.decode() == f[:-4]
        assert p.seg_x > 0 and p.seg_y > 0 and p.f_boundary and p.f_analytic
    p = host_lib.parse_input(os.path.join(INPUTS, "Poisson1.txt"))
    assert (p.x_max, p.y_max, p.seg_x, p.seg_y) == (b"exp(1)", b"pi(1)", 128, 256)
    assert p.f_rhs == b"-4"
    assert p.f_boundary == b"(x-exp(1)/2)^2 + (y-pi(1)/2)^2"
    assert p.eps_double == 5e-13
    assert list(p.mg_iters) == [9, 9, 9, 9]
    q = host_lib.parse_input(os.path.join(INPUTS, "Laplace1.txt"))
    assert (q.seg_x, q.seg_y, q.f_rhs) == (200, 200, b"")
    assert q.eps_double == 5e-14 and list(q.mg_iters) == [9, 9, 8, 7]
    assert q.f_left == b"" and q.f_boundary == b"x^2 - y^2 + 10 * x * y + 10 * x - 10 * y + 1"
    r = host_lib.parse_input(os.path.join(INPUTS, "Laplace1_16K.txt"))
    assert (r.seg_x, r.seg_y) == (16384, 16384) and list(r.mg_iters) == [5000] * 4


def test_input_file_errors(tmp_path):
    f = tmp_path / "bad.txt"
    f.write_text("xMax=1\nsegmentsX=abc\nfuncRHS=1\n")
    with pytest.raises(_lib.VLP2DError) as e:
        host_lib.parse_input(str(f))
    msg = str(e.value)
    assert "yMax" in msg and "segmentsY" in msg and "segmentsX is not an integer" in msg and "boundary" in msg
    g = tmp_path / "edges.txt"
    g.write_text("xMax=1\nyMax=2\nsegmentsX=8\nsegmentsY=8\nfuncLeft=y\nfuncRight=1+y\nfuncTop=x\nfuncBottom=0\n")
    p = host_lib.parse_input(str(g))
    assert p.f_boundary == b"" and p.f_left == b"y" and p.eps_double == 1e-13


def test_pow2_sizing_of_the_multigrid_method():
    """StepsRecalculator (Model/CPU/StepsRecalculator.cs:26-30) rounds the
    MultiGrid segment counts up to powers of two: Laplace1's 200 -> 256."""
    import subprocess
    assert os.path.exists(host_lib.SOLVE_PATH)
    # the usage line needs no GPU
    r = subprocess.run([host_lib.SOLVE_PATH], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.parametrize("args,msg", [(["--smoother", "sor"], "unknown smoother"),
                                      (["--smoother", "jacobi:0"], "omega"),
                                      (["--smoother", "jacobi:1.5"], "omega"),
                                      (["--smoother", "jacobi:0.8x"], "omega"),
                                      (["--restriction", "quarter"], "unknown restriction"),
                                      (["--cycle", "2"], "nu1,nu2"),
                                      (["--fmg", "--bogus"], "unknown option")])
def test_solve_cli_rejects_bad_options_before_touching_a_device(args, msg):
    import subprocess
    r = subprocess.run([host_lib.SOLVE_PATH, os.path.join(INPUTS, "Poisson1.txt")] + args,
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and msg in r.stderr, r.stderr
