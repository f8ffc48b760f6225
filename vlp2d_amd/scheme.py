"""Host mirror of the reference solver-method interface for the HIP multigrid.

``MultiGridSchemeHIP`` exposes the methods of the reference's
``IScheme<T> = ISchemeInit<T> + ISchemeCalculate<T>``
(VisualLaplacePoisson2D/Model/CPU/IScheme.cs:16-45) as implemented for GPU
schemes by ``Iterative1DScheme<T>`` (Model/CUDA/Iterative1DScheme.cs:13-70)
and ``MultiGridSchemeCU<T>`` (Model/CUDA/MultiGridSchemeCU.cs), with Python
names.  It owns a host mirror of the level-0 array (``un0``: boundary values
plus initial guess, i*ny + j layout) that the model-side utilities fill before
``init_after_boundaries_and_initial_iteration_inited`` uploads it, exactly as
the reference's model does (VLPRectangleModel.cs:535-577).  All numerical
work goes through the C ABI (include/vlp2d_mg.h).
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import VLP2DError, check, load_library, ptr


@dataclass
class Options:
    """Cycle shape and precision.  Defaults are the reference's: V(3,3) with 3
    sweeps on the coarsest level (MultiGridScheme.cs:21,97,104,112), FP64."""
    nu1: int = 3
    nu2: int = 3
    nu_coarse: int = 3
    precision: int = _lib.PREC_FP64
    flags: int = 0
    coarse_nodes: int = 0
    max_inner: int = 1  # mixed: FP32 V-cycles per FP64 refinement step, at most
    smoother: int = _lib.SMOOTH_RBGS  # or SMOOTH_WJACOBI (one operator per launch)
    omega: float = 0.8  # weighted-Jacobi weight
    restriction: int = _lib.RESTRICT_HALF  # or RESTRICT_FULL (one operator per launch)
    cycle: int = _lib.CYCLE_V  # or CYCLE_FMG (FP64, one operator per launch)

    @staticmethod
    def performance(mixed=False, **kw):
        """V(2,2) cycles unless nu1/nu2 are given (relative-residual stop),
        FP64 or FP32-in-FP64."""
        base = dict(nu1=2, nu2=2, nu_coarse=3, precision=_lib.PREC_MIXED if mixed else _lib.PREC_FP64)
        base.update(kw)
        return Options(**base)

    def to_c(self):
        return _lib.Opts(self.nu1, self.nu2, self.nu_coarse, self.precision, self.flags, self.coarse_nodes,
                         self.max_inner, self.smoother, self.omega, self.restriction,
                         self.cycle)


def _sample_rhs(f_ksi, nx, ny, hx, hy):
    if f_ksi is None:
        return None
    if callable(f_ksi):
        x = hx * np.arange(nx, dtype=np.float64)
        y = hy * np.arange(ny, dtype=np.float64)
        f = np.asarray(f_ksi(x[:, None], y[None, :]), dtype=np.float64)
        return np.ascontiguousarray(np.broadcast_to(f, (nx, ny)))
    f = np.ascontiguousarray(f_ksi, dtype=np.float64)
    if f.shape != (nx, ny):
        raise ValueError(f"f must have shape {(nx, ny)}, got {f.shape}")
    return f


class MultiGridSchemeHIP:
    """Drop-in HIP multigrid scheme.

    Constructor arguments follow ``MultiGridSchemeCU<T>(segX, segY, hx, hy,
    fKsi, eps, device)`` (MultiGridSchemeCU.cs:32).  ``f_ksi`` is the PDE
    right-hand side f of u_xx + u_yy = -f, either a vectorised callable f(x, y)
    sampled at x = hx*i, y = hy*j, an (nx, ny) array, or None (Laplace).
    """

    def __init__(self, seg_x, seg_y, step_x, step_y, f_ksi=None, eps=1e-13, device=0, options=None, f_expr=None):
        """``f_expr`` (a text in the input files' syntax, vlp2d_mg.h) instead of
        ``f_ksi`` samples f on the GPU."""
        self.lib = load_library()
        self.seg_x, self.seg_y = int(seg_x), int(seg_y)
        self.dim_x, self.dim_y = self.seg_x + 1, self.seg_y + 1
        self.step_x, self.step_y = float(step_x), float(step_y)
        self.eps = float(eps)
        self.options = options or Options()
        opts = self.options.to_c()
        h = ctypes.c_void_p()
        if f_expr is not None:
            if f_ksi is not None:
                raise ValueError("give f_ksi or f_expr, not both")
            rc = self.lib.vlp2d_mg_create_expr(self.seg_x, self.seg_y, self.step_x, self.step_y, f_expr.encode(),
                                               self.eps, int(device), ctypes.byref(opts), ctypes.byref(h))
        else:
            f = _sample_rhs(f_ksi, self.dim_x, self.dim_y, self.step_x, self.step_y)
            rc = self.lib.vlp2d_mg_create(self.seg_x, self.seg_y, self.step_x, self.step_y, ptr(f), self.eps,
                                          int(device), ctypes.byref(opts), ctypes.byref(h))
        if rc != _lib.OK:
            raise VLP2DError(rc, self.lib.vlp2d_mg_last_error(None).decode())
        self.h = h
        self.un0 = np.zeros((self.dim_x, self.dim_y), dtype=np.float64)
        self._un0_changed = False
        self.last_metric = None

    # ------------------------------------------------- device-side set-up
    def set_boundary_expr(self, bottom="", top="", left="", right="", boundary=""):
        """Samples the boundary on the GPU into the iterate (vlp2d_mg_set_boundary_expr);
        returns (min, max) of the values written.  The host mirror ``un0`` is
        not involved; ``get_array`` downloads the device grid."""
        lo, hi = ctypes.c_double(), ctypes.c_double()
        check(self.lib, self.h, self.lib.vlp2d_mg_set_boundary_expr(
            self.h, bottom.encode(), top.encode(), left.encode(), right.encode(), boundary.encode(),
            ctypes.byref(lo), ctypes.byref(hi)))
        self._un0_changed = False
        return lo.value, hi.value

    def init_guess(self, kind, value=float("nan")):
        """Initial guess of the interior on the GPU (vlp2d_mg_init_guess, _lib.INIT_*)."""
        check(self.lib, self.h, self.lib.vlp2d_mg_init_guess(self.h, int(kind), float(value)))
        self._un0_changed = False

    def points_min_max_device(self):
        lo, hi = ctypes.c_double(), ctypes.c_double()
        check(self.lib, self.h, self.lib.vlp2d_mg_points_min_max(self.h, ctypes.byref(lo), ctypes.byref(hi)))
        return lo.value, hi.value

    def difference(self, analytic, with_array=False):
        """(min, max[, diff]) of u - analytic on nodes i, j >= 1 (vlp2d_mg_difference)."""
        lo, hi = ctypes.c_double(), ctypes.c_double()
        d = np.zeros((self.dim_x, self.dim_y)) if with_array else None
        check(self.lib, self.h, self.lib.vlp2d_mg_difference(self.h, analytic.encode(), ctypes.byref(lo),
                                                             ctypes.byref(hi), ptr(d)))
        return (lo.value, hi.value, d) if with_array else (lo.value, hi.value)

    def download(self):
        """The device iterate (nx, ny), without touching the host mirror."""
        u = np.empty((self.dim_x, self.dim_y))
        check(self.lib, self.h, self.lib.vlp2d_mg_download(self.h, ptr(u)))
        return u

    # ----------------------------------------------------------- ISchemeInit
    def get_array_dimensions(self):
        return self.dim_x, self.dim_y

    def init_top_bottom_borders(self, delta_x, delta_y, func_bottom=None, func_top=None, func_border=None,
                                val_min=np.inf, val_max=-np.inf):
        """Bottom (j = 0) and top (j = ny-1) edges for all i (UtilsBorders.cs:115-144).
        Coordinates are x = delta_x * i (the reference accumulates x += h*cores,
        which makes its rounding depend on the core count).  Returns (min, max)."""
        x = delta_x * np.arange(self.dim_x, dtype=np.float64)
        y_max = delta_y * (self.dim_y - 1)
        bottom = func_bottom(x) if func_bottom is not None else func_border(x, 0.0 * x)
        top = func_top(x) if func_top is not None else func_border(x, 0.0 * x + y_max)
        self.un0[:, 0] = bottom
        self.un0[:, -1] = top
        return (min(val_min, self.un0[:, 0].min(), self.un0[:, -1].min()),
                max(val_max, self.un0[:, 0].max(), self.un0[:, -1].max()))

    def init_left_right_borders(self, delta_x, delta_y, func_left=None, func_right=None, func_border=None,
                                val_min=np.inf, val_max=-np.inf):
        """Left (i = 0) and right (i = nx-1) edges for all j, overwriting the
        corners written by the top/bottom pass (UtilsBorders.cs:210-239,
        VLPRectangleModel.cs:541-542).  Returns (min, max)."""
        y = delta_y * np.arange(self.dim_y, dtype=np.float64)
        x_max = delta_x * (self.dim_x - 1)
        left = func_left(y) if func_left is not None else func_border(0.0 * y, y)
        right = func_right(y) if func_right is not None else func_border(0.0 * y + x_max, y)
        self.un0[0, :] = left
        self.un0[-1, :] = right
        return (min(val_min, self.un0[0, :].min(), self.un0[-1, :].min()),
                max(val_max, self.un0[0, :].max(), self.un0[-1, :].max()))

    def set_boundary(self, g):
        """Copy the boundary entries of a full (nx, ny) array into un0."""
        g = np.asarray(g, dtype=np.float64)
        self.un0[0, :], self.un0[-1, :] = g[0, :], g[-1, :]
        self.un0[:, 0], self.un0[:, -1] = g[:, 0], g[:, -1]

    def points_min_max(self):
        a = self.get_array()
        return float(a.min()), float(a.max())

    def init_initial_iteration_mean(self, val):
        """Interior := val (UtilsII.cs:15-19); the model passes (min+max)/2 of
        the boundary values (VLPRectangleModel.cs:546)."""
        self.un0[1:-1, 1:-1] = val

    def init_initial_iteration_arithmetic_mean(self):
        """Interior := mean of the four boundary values on its row and column
        (UtilsII.cs:28-33)."""
        u = self.un0
        u[1:-1, 1:-1] = (u[0, 1:-1][None, :] + u[-1, 1:-1][None, :] + u[1:-1, 0][:, None]
                         + u[1:-1, -1][:, None]) / 4.0

    def _lin_xy(self):
        u = self.un0
        ux, uy = self.dim_x - 1, self.dim_y - 1
        i = np.arange(1, ux, dtype=np.float32)
        j = np.arange(1, uy, dtype=np.float32)
        sx = (i / np.float32(ux)).astype(np.float64)[:, None]  # (float)i / upperX, UtilsII.cs:56-61
        sy = (j / np.float32(uy)).astype(np.float64)[None, :]
        lx = u[0, 1:-1][None, :] * (1.0 - sx) + u[-1, 1:-1][None, :] * sx
        ly = u[1:-1, 0][:, None] * (1.0 - sy) + u[1:-1, -1][:, None] * sy
        return lx, ly

    def init_initial_iteration_linear_interpolation(self):
        """Average of the x- and y-linear interpolants of the boundary
        (UtilsII.cs:70-73)."""
        lx, ly = self._lin_xy()
        self.un0[1:-1, 1:-1] = (lx + ly) / 2.0

    def init_initial_iteration_weight_linear_interpolation(self):
        """Distance-weighted blend of the two interpolants (UtilsII.cs:90-102)."""
        lx, ly = self._lin_xy()
        cx, cy = self.dim_x - 1, self.dim_y - 1
        i = np.arange(1, cx)
        j = np.arange(1, cy)
        lng_x = np.where(cx > 2 * i, i, cx - i)[:, None]
        lng_y = np.where(cy > 2 * j, j, cy - j)[None, :]
        w = (lng_x.astype(np.float32) / (lng_x + lng_y).astype(np.float32)).astype(np.float64)
        self.un0[1:-1, 1:-1] = lx * (1.0 - w) + ly * w

    def calculate_difference(self, func_analytic):
        """u - u_exact at every node and its (min, max) (UtilsDiff.cs:42-88)."""
        a = self.get_array()
        x = self.step_x * np.arange(self.dim_x)[:, None]
        y = self.step_y * np.arange(self.dim_y)[None, :]
        d = a - func_analytic(x, y)
        return d, float(np.nanmin(d)), float(np.nanmax(d))

    def iterations_kind(self):
        return "unknown"

    def get_elapsed_info(self):
        return None

    # ------------------------------------------------------ ISchemeCalculate
    def init_after_boundaries_and_initial_iteration_inited(self):
        """Upload un0 (MultiGridSchemeCU.cs:263-266)."""
        check(self.lib, self.h, self.lib.vlp2d_mg_upload(self.h, ptr(self.un0)))
        self._un0_changed = False

    def do_iteration(self, it=0):
        """One V-cycle.  Like MultiGridSchemeCU.doIteration (:140-146) returns
        eps + eps while the update norm exceeds eps, else 0; the update norm
        itself is kept in ``last_metric``."""
        m = ctypes.c_double()
        check(self.lib, self.h, self.lib.vlp2d_mg_cycle(self.h, ctypes.byref(m)))
        self._un0_changed = True
        self.last_metric = m.value
        return self.eps + self.eps if m.value > self.eps else 0.0

    def get_array(self):
        """Device iterate -> un0 (MultiGridSchemeCU.cs:252-261)."""
        if self._un0_changed:
            check(self.lib, self.h, self.lib.vlp2d_mg_download(self.h, ptr(self.un0)))
            self._un0_changed = False
        return self.un0

    def cancel_iterations(self):
        self.lib.vlp2d_mg_cancel(self.h)

    def max_iterations(self):
        return 0

    def should_report_progress(self):
        return True

    def cleanup(self):
        if getattr(self, "h", None):
            self.lib.vlp2d_mg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.cleanup()
        except Exception:
            pass

    # ------------------------------------------------ additions of this build
    def residual_norm(self):
        v = ctypes.c_double()
        check(self.lib, self.h, self.lib.vlp2d_mg_residual_norm(self.h, ctypes.byref(v)))
        return v.value

    def solve(self, rtol=1e-10, max_cycles=100, raise_on_notconv=True):
        """Cycles until ||r|| <= rtol*||r0|| on the device.  Returns stats dict."""
        st = _lib.Stats()
        rc = self.lib.vlp2d_mg_solve(self.h, float(rtol), int(max_cycles), ctypes.byref(st))
        self._un0_changed = True
        if rc != _lib.OK and (raise_on_notconv or rc != _lib.E_NOTCONV):
            check(self.lib, self.h, rc)
        return {"cycles": st.cycles, "levels": st.levels, "r0": st.r0, "rel_residual": st.rel_residual,
                "seconds": st.seconds, "munknowns_per_s": st.munknowns_per_s, "vcycles": st.vcycles, "rc": rc}

    def adjoint_exact(self, dL_du, rtol=1e-12, max_cycles=100):
        shape = (self.dim_x, self.dim_y)
        ub = np.ascontiguousarray(dL_du, dtype=np.float64).reshape(shape)
        df, dg = np.zeros(shape), np.zeros(shape)
        check(self.lib, self.h, self.lib.vlp2d_mg_adjoint_exact(self.h, ptr(ub), float(rtol), int(max_cycles),
                                                                ptr(df), ptr(dg)))
        return df, dg

    def forward_cycles(self, K):
        check(self.lib, self.h, self.lib.vlp2d_mg_forward_cycles(self.h, int(K)))
        self._un0_changed = True

    def adjoint_cycles(self, dL_du):
        shape = (self.dim_x, self.dim_y)
        ub = np.ascontiguousarray(dL_du, dtype=np.float64).reshape(shape)
        df, dg, du0 = np.empty(shape), np.empty(shape), np.empty(shape)  # every node is written
        check(self.lib, self.h, self.lib.vlp2d_mg_adjoint_cycles(self.h, ptr(ub), ptr(df), ptr(dg), ptr(du0)))
        return df, dg, du0

    def info(self):
        inf = _lib.Info()
        check(self.lib, self.h, self.lib.vlp2d_mg_get_info(self.h, ctypes.byref(inf)))
        return {"levels": inf.levels, "nx": inf.nx, "ny": inf.ny, "device_bytes": inf.device_bytes,
                "device_peak_bytes": inf.device_peak_bytes, "checkpoints": inf.checkpoints}

    def kernel_times(self):
        """[(name, level, launches, total_ms, algorithmic_bytes)] per kernel
        class and level (times only under the PROFILE flag)."""
        cap = 256
        names = (ctypes.c_char_p * cap)()
        levels = (ctypes.c_int * cap)()
        launches = (ctypes.c_longlong * cap)()
        ms = (ctypes.c_double * cap)()
        nbytes = (ctypes.c_double * cap)()
        n = self.lib.vlp2d_mg_kernel_times(self.h, names, levels, launches, ms, nbytes, cap)
        return [(names[k].decode(), levels[k], launches[k], ms[k], nbytes[k]) for k in range(min(n, cap))]

    def save_iterate(self):
        check(self.lib, self.h, self.lib.vlp2d_mg_save_iterate(self.h))

    def restore_iterate(self):
        check(self.lib, self.h, self.lib.vlp2d_mg_restore_iterate(self.h))
        self._un0_changed = True

    def reset_kernel_times(self):
        check(self.lib, self.h, self.lib.vlp2d_mg_reset_kernel_times(self.h))
