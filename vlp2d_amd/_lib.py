"""ctypes binding of libvlp2d_mg.so (declarations in include/vlp2d_mg.h).

The library must have been built (``python -m vlp2d_amd.build``); there is no
fallback: if it is missing, loading raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libvlp2d_mg.so")

OK = 0
E_ARG, E_SIZE, E_DEVICE, E_OOM, E_NONFINITE, E_NOTCONV, E_CANCELLED, E_STATE = -1, -2, -3, -4, -5, -6, -7, -8
ERROR_NAMES = {
    E_ARG: "VLP2D_E_ARG", E_SIZE: "VLP2D_E_SIZE", E_DEVICE: "VLP2D_E_DEVICE", E_OOM: "VLP2D_E_OOM",
    E_NONFINITE: "VLP2D_E_NONFINITE", E_NOTCONV: "VLP2D_E_NOTCONV", E_CANCELLED: "VLP2D_E_CANCELLED",
    E_STATE: "VLP2D_E_STATE",
}
PREC_FP64, PREC_MIXED = 0, 1
FLAG_NO_GRAPH, FLAG_PROFILE, FLAG_SIMPLE, FLAG_SEPARATE_REFINE, FLAG_COARSE_L2 = 1, 2, 4, 8, 16
FLAG_STORE_PRESMOOTH = 32
FLAG_STREAM_SMALL = 64
FLAG_NO_CYCLE_FUSION = 128
FLAG_GRAPH_SOLVE = 256
FLAG_NO_STEP_FOLD = 512
FLAG_NO_LT_PAIR = 1024
SMOOTH_RBGS, SMOOTH_WJACOBI = 0, 1
RESTRICT_HALF, RESTRICT_FULL = 0, 1
CYCLE_V, CYCLE_FMG = 0, 1

# Every entry point declared in include/vlp2d_mg.h.
EXPORTS = [
    "vlp2d_mg_default_opts", "vlp2d_mg_version", "vlp2d_mg_create", "vlp2d_mg_upload", "vlp2d_mg_cycle",
    "vlp2d_mg_download", "vlp2d_mg_residual_norm", "vlp2d_mg_solve", "vlp2d_mg_adjoint_exact",
    "vlp2d_mg_forward_cycles", "vlp2d_mg_adjoint_cycles", "vlp2d_mg_get_info", "vlp2d_mg_kernel_times",
    "vlp2d_mg_reset_kernel_times", "vlp2d_mg_save_iterate", "vlp2d_mg_restore_iterate", "vlp2d_mg_cancel",
    "vlp2d_mg_destroy", "vlp2d_mg_last_error", "vlp2d_mg_create_expr", "vlp2d_mg_set_boundary_expr",
    "vlp2d_mg_init_guess", "vlp2d_mg_points_min_max", "vlp2d_mg_difference", "vlp2d_expr_eval",
    "vlp2d_mg_solve_noalloc",
]
NOALLOC_SCRATCH_BYTES = 64 << 20
INIT_MEAN, INIT_ARITHMETIC_MEAN, INIT_LINEAR, INIT_WEIGHT_LINEAR = 0, 1, 2, 3


class VLP2DError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {message}")
        self.code = code


class Opts(ctypes.Structure):
    _fields_ = [("nu1", ctypes.c_int), ("nu2", ctypes.c_int), ("nu_coarse", ctypes.c_int),
                ("precision", ctypes.c_int), ("flags", ctypes.c_uint), ("coarse_nodes", ctypes.c_int),
                ("max_inner", ctypes.c_int), ("smoother", ctypes.c_int), ("omega", ctypes.c_double),
                ("restriction", ctypes.c_int), ("cycle", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("cycles", ctypes.c_int), ("levels", ctypes.c_int), ("r0", ctypes.c_double),
                ("rel_residual", ctypes.c_double), ("seconds", ctypes.c_double),
                ("munknowns_per_s", ctypes.c_double), ("vcycles", ctypes.c_int)]


class Info(ctypes.Structure):
    _fields_ = [("levels", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int),
                ("device_bytes", ctypes.c_longlong), ("device_peak_bytes", ctypes.c_longlong),
                ("checkpoints", ctypes.c_int)]


_dp = ctypes.POINTER(ctypes.c_double)
_h = ctypes.c_void_p
_lib = None


def load_library(path=LIB_PATH):
    """Load libvlp2d_mg.so and declare its signatures.  Raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run `python -m vlp2d_amd.build` (no CPU fallback exists)")
    lib = ctypes.CDLL(path)
    sig = {
        "vlp2d_mg_default_opts": (None, [ctypes.POINTER(Opts)]),
        "vlp2d_mg_version": (ctypes.c_int, []),
        "vlp2d_mg_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp,
                                           ctypes.c_double, ctypes.c_int, ctypes.POINTER(Opts),
                                           ctypes.POINTER(_h)]),
        "vlp2d_mg_upload": (ctypes.c_int, [_h, _dp]),
        "vlp2d_mg_cycle": (ctypes.c_int, [_h, _dp]),
        "vlp2d_mg_download": (ctypes.c_int, [_h, _dp]),
        "vlp2d_mg_residual_norm": (ctypes.c_int, [_h, _dp]),
        "vlp2d_mg_solve": (ctypes.c_int, [_h, ctypes.c_double, ctypes.c_int, ctypes.POINTER(Stats)]),
        "vlp2d_mg_adjoint_exact": (ctypes.c_int, [_h, _dp, ctypes.c_double, ctypes.c_int, _dp, _dp]),
        "vlp2d_mg_forward_cycles": (ctypes.c_int, [_h, ctypes.c_int]),
        "vlp2d_mg_adjoint_cycles": (ctypes.c_int, [_h, _dp, _dp, _dp, _dp]),
        "vlp2d_mg_get_info": (ctypes.c_int, [_h, ctypes.POINTER(Info)]),
        "vlp2d_mg_kernel_times": (ctypes.c_int, [_h, ctypes.POINTER(ctypes.c_char_p),
                                                 ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_longlong),
                                                 _dp, _dp, ctypes.c_int]),
        "vlp2d_mg_reset_kernel_times": (ctypes.c_int, [_h]),
        "vlp2d_mg_save_iterate": (ctypes.c_int, [_h]),
        "vlp2d_mg_restore_iterate": (ctypes.c_int, [_h]),
        "vlp2d_mg_cancel": (None, [_h]),
        "vlp2d_mg_destroy": (None, [_h]),
        "vlp2d_mg_last_error": (ctypes.c_char_p, [_h]),
        "vlp2d_mg_create_expr": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                ctypes.c_char_p, ctypes.c_double, ctypes.c_int, ctypes.POINTER(Opts),
                                                ctypes.POINTER(_h)]),
        "vlp2d_mg_set_boundary_expr": (ctypes.c_int, [_h, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                                      ctypes.c_char_p, ctypes.c_char_p, _dp, _dp]),
        "vlp2d_mg_init_guess": (ctypes.c_int, [_h, ctypes.c_int, ctypes.c_double]),
        "vlp2d_mg_points_min_max": (ctypes.c_int, [_h, _dp, _dp]),
        "vlp2d_mg_difference": (ctypes.c_int, [_h, ctypes.c_char_p, _dp, _dp, _dp]),
        "vlp2d_expr_eval": (ctypes.c_int, [ctypes.c_char_p, _dp, _dp, ctypes.c_longlong, _dp]),
        "vlp2d_mg_solve_noalloc": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                                  ctypes.c_int, ctypes.POINTER(Opts), ctypes.c_int, ctypes.c_void_p,
                                                  ctypes.POINTER(Stats)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def ptr(a):
    """double* of a C-contiguous float64 numpy array (None -> NULL)."""
    if a is None:
        return None
    if a.dtype.name != "float64" or not a.flags.c_contiguous:
        raise TypeError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(_dp)


def check(lib, handle, rc):
    if rc != OK:
        msg = lib.vlp2d_mg_last_error(handle)
        raise VLP2DError(rc, msg.decode() if msg else "")
    return rc
