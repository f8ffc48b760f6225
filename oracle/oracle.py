"""ctypes wrapper over the oracle library (oracle/lib/libvlp2d_oracle.so).

TEST INFRASTRUCTURE ONLY.  The C++ restatement follows
VLP/Model/CPU/MultiGridScheme.cs:23-186, MultiGridSlidingIteration.cs:29-70,
VLP/Common/GridIterator.cs:170-222 and UtilsOpLap.cs:15-20 (see the header of
oracle/mg_oracle.cpp for the line-by-line map).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libvlp2d_oracle.so")
_lib = None

PREC = {"double": 0, "float": 1, "dd": 2}
_dp = ctypes.POINTER(ctypes.c_double)


def build():
    """Compile the oracle with its Makefile (g++, OpenMP, no FMA contraction)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    lib.orc_create.restype = ctypes.c_void_p
    lib.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                               _dp, ctypes.c_int]
    lib.orc_destroy.argtypes = [ctypes.c_void_p]
    lib.orc_levels.argtypes = [ctypes.c_void_p]
    lib.orc_set_u.argtypes = [ctypes.c_void_p, _dp]
    lib.orc_get_u.argtypes = [ctypes.c_void_p, _dp, _dp]
    lib.orc_vcycle.restype = ctypes.c_double
    lib.orc_vcycle.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 4
    lib.orc_residual_norm2.restype = ctypes.c_double
    lib.orc_residual_norm2.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.orc_plain_solve.restype = ctypes.c_int
    lib.orc_plain_solve.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_int, _dp, ctypes.c_int]
    lib.orc_mixed_solve.restype = ctypes.c_int
    lib.orc_mixed_solve.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int]
    lib.orc_set_restriction.restype = None
    lib.orc_set_restriction.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.orc_set_cycle.restype = None
    lib.orc_set_cycle.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.orc_set_smoother.restype = None
    lib.orc_set_smoother.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
    lib.orc_forward_cycles.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.orc_adjoint_cycles.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp,
                                       ctypes.c_int]
    lib.orc_op_restrict.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int]
    lib.orc_op_prolong.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
    lib.orc_op_residual.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp]
    lib.orc_op_smooth.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_double]
    lib.orc_dd_op.argtypes = [ctypes.c_int, _dp, _dp, _dp]
    _lib = lib
    return lib


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


class Oracle:
    """Step-by-step oracle scheme, same call pattern as the reference IScheme:
    construct with (segX, segY, hx, hy, f), set the level-0 iterate (boundary +
    initial guess), then one V-cycle per ``do_iteration``."""

    def __init__(self, seg_x, seg_y, hx, hy, f=None, prec="double", nthreads=0):
        self.lib = load()
        self.nx, self.ny = seg_x + 1, seg_y + 1
        fa = None if f is None else np.ascontiguousarray(f, dtype=np.float64).reshape(self.nx, self.ny)
        self._f = fa
        self.nthreads = nthreads
        self.h = self.lib.orc_create(PREC[prec], seg_x, seg_y, hx, hy, _ptr(fa), nthreads)
        self.levels = self.lib.orc_levels(self.h)

    def set_u(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(self.nx, self.ny)
        self.lib.orc_set_u(self.h, _ptr(u))

    def get_u(self, with_lo=False):
        hi = np.zeros((self.nx, self.ny))
        lo = np.zeros((self.nx, self.ny)) if with_lo else None
        self.lib.orc_get_u(self.h, _ptr(hi), _ptr(lo))
        return (hi, lo) if with_lo else hi

    def set_smoother(self, kind, omega=0.8):
        """kind 0: red-black Gauss-Seidel (the reference's); 1: weighted
        Jacobi u += omega (g - u) with the same point formula g."""
        self.lib.orc_set_smoother(self.h, int(kind), float(omega))

    def set_restriction(self, kind):
        """kind 0: half weighting (the reference's); 1: full weighting."""
        self.lib.orc_set_restriction(self.h, int(kind))

    def set_cycle(self, kind):
        """kind 0: V-cycle (the reference's); 1: FMG, the reference's unused
        FMG() driver (MultiGridScheme.cs:117-132)."""
        self.lib.orc_set_cycle(self.h, int(kind))

    def vcycle(self, nu1=3, nu2=3, nuc=3):
        return self.lib.orc_vcycle(self.h, nu1, nu2, nuc, self.nthreads)

    def residual_norm(self):
        return float(np.sqrt(self.lib.orc_residual_norm2(self.h, self.nthreads)))

    def plain_solve(self, nu1=2, nu2=2, nuc=3, rtol=1e-10, max_cycles=100):
        hist = np.zeros(max_cycles + 1)
        k = self.lib.orc_plain_solve(self.h, nu1, nu2, nuc, rtol, max_cycles, _ptr(hist), self.nthreads)
        return k, hist[: k + 1]

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def plain_solve(seg_x, seg_y, hx, hy, f, u, nu1=2, nu2=2, nuc=3, rtol=1e-10, max_cycles=100, nthreads=0,
                smoother=0, omega=0.8, restriction=0, cycle=0):
    o = Oracle(seg_x, seg_y, hx, hy, f, "double", nthreads)
    o.set_smoother(smoother, omega)
    o.set_restriction(restriction)
    o.set_cycle(cycle)
    o.set_u(u)
    k, hist = o.plain_solve(nu1, nu2, nuc, rtol, max_cycles)
    out = o.get_u()
    o.close()
    return out, k, hist


def mixed_solve(seg_x, seg_y, hx, hy, f, u, nu1=2, nu2=2, nuc=3, rtol=1e-10, max_outer=100, nthreads=0, inner=1,
                smoother=0, omega=0.8, restriction=0):
    lib = load()
    u = np.array(u, dtype=np.float64, order="C").reshape(seg_x + 1, seg_y + 1)
    fa = None if f is None else np.ascontiguousarray(f, dtype=np.float64)
    hist = np.zeros(max_outer + 1)
    k = lib.orc_mixed_solve(seg_x, seg_y, hx, hy, _ptr(fa), _ptr(u), nu1, nu2, nuc, rtol, max_outer, _ptr(hist),
                            nthreads, inner, smoother, omega, restriction)
    return u, k, hist[: k + 1]


def forward_cycles(seg_x, seg_y, hx, hy, f, u0, K, nu1=2, nu2=2, nuc=3, nthreads=0):
    lib = load()
    u = np.array(u0, dtype=np.float64, order="C").reshape(seg_x + 1, seg_y + 1)
    fa = None if f is None else np.ascontiguousarray(f, dtype=np.float64)
    lib.orc_forward_cycles(seg_x, seg_y, hx, hy, _ptr(fa), _ptr(u), K, nu1, nu2, nuc, nthreads)
    return u


def adjoint_cycles(seg_x, seg_y, hx, hy, ubar, K, nu1=2, nu2=2, nuc=3, nthreads=0):
    lib = load()
    shape = (seg_x + 1, seg_y + 1)
    ub = np.ascontiguousarray(ubar, dtype=np.float64).reshape(shape)
    fb, gb, u0b = np.zeros(shape), np.zeros(shape), np.zeros(shape)
    lib.orc_adjoint_cycles(seg_x, seg_y, hx, hy, K, nu1, nu2, nuc, _ptr(ub), _ptr(fb), _ptr(gb), _ptr(u0b),
                           nthreads)
    return fb, gb, u0b


# ---- single operators (FP64) -------------------------------------------------
def op_restrict(seg_x, seg_y, res_fine, full=False):
    lib = load()
    r = np.ascontiguousarray(res_fine, dtype=np.float64)
    out = np.zeros((seg_x // 2 + 1, seg_y // 2 + 1))
    lib.orc_op_restrict(seg_x, seg_y, _ptr(r), _ptr(out), 1 if full else 0)
    return out


def op_prolong(seg_x, seg_y, u_coarse, u_fine):
    lib = load()
    uc = np.ascontiguousarray(u_coarse, dtype=np.float64)
    uf = np.array(u_fine, dtype=np.float64, order="C")
    lib.orc_op_prolong(seg_x, seg_y, _ptr(uc), _ptr(uf))
    return uf


def op_residual(seg_x, seg_y, hx, hy, f, u):
    lib = load()
    fa = None if f is None else np.ascontiguousarray(f, dtype=np.float64)
    ua = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros_like(ua)
    lib.orc_op_residual(seg_x, seg_y, hx, hy, _ptr(fa), _ptr(ua), _ptr(out))
    return out


def op_smooth(seg_x, seg_y, hx, hy, f, u, sweeps=1, smoother=0, omega=0.8):
    lib = load()
    fa = None if f is None else np.ascontiguousarray(f, dtype=np.float64)
    ua = np.array(u, dtype=np.float64, order="C")
    lib.orc_op_smooth(seg_x, seg_y, hx, hy, _ptr(fa), _ptr(ua), sweeps, smoother, omega)
    return ua


def dd_op(op, a, b):
    """op in '+-*/'; a, b = (hi, lo).  Returns (hi, lo)."""
    lib = load()
    x = np.array(a, dtype=np.float64)
    y = np.array(b, dtype=np.float64)
    out = np.zeros(2)
    lib.orc_dd_op("+-*/".index(op), _ptr(x), _ptr(y), _ptr(out))
    return float(out[0]), float(out[1])
