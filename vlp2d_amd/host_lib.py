"""ctypes binding of the C++ host layer (lib/libvlp2d_host.so, host/vlp2d_host.h):
the reference's input files and its MultiGrid model run on the GPU."""
import ctypes
import os

from . import _lib

HOST_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libvlp2d_host.so")
SOLVE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "vlp2d_solve")
EXPR_LEN = 1024

_host = None


class Input(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 256), ("x_max", ctypes.c_char * EXPR_LEN),
                ("y_max", ctypes.c_char * EXPR_LEN), ("seg_x", ctypes.c_int), ("seg_y", ctypes.c_int),
                ("f_rhs", ctypes.c_char * EXPR_LEN), ("f_left", ctypes.c_char * EXPR_LEN),
                ("f_right", ctypes.c_char * EXPR_LEN), ("f_top", ctypes.c_char * EXPR_LEN),
                ("f_bottom", ctypes.c_char * EXPR_LEN), ("f_boundary", ctypes.c_char * EXPR_LEN),
                ("f_analytic", ctypes.c_char * EXPR_LEN), ("eps_double", ctypes.c_double),
                ("mg_iters", ctypes.c_int * 4)]


class RunResult(ctypes.Structure):
    _fields_ = [("seg_x", ctypes.c_int), ("seg_y", ctypes.c_int), ("hx", ctypes.c_double), ("hy", ctypes.c_double),
                ("eps", ctypes.c_double), ("cycles", ctypes.c_int), ("elapsed_iters", ctypes.c_int),
                ("converged", ctypes.c_int), ("last_metric", ctypes.c_double), ("init_seconds", ctypes.c_double),
                ("solve_seconds", ctypes.c_double), ("vmin", ctypes.c_double), ("vmax", ctypes.c_double),
                ("dmin", ctypes.c_double), ("dmax", ctypes.c_double)]


def load_host(path=HOST_PATH):
    """Load libvlp2d_host.so (and the solver library it links).  Raises if absent."""
    global _host
    if _host is not None:
        return _host
    _lib.load_library()
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run `python -m vlp2d_amd.build`")
    h = ctypes.CDLL(path)
    h.vlp2d_input_parse.restype = ctypes.c_int
    h.vlp2d_input_parse.argtypes = [ctypes.c_char_p, ctypes.POINTER(Input), ctypes.c_char_p, ctypes.c_int]
    h.vlp2d_run_input.restype = ctypes.c_int
    h.vlp2d_run_input.argtypes = [ctypes.POINTER(Input), ctypes.c_int, ctypes.c_int, ctypes.POINTER(_lib.Opts),
                                  ctypes.c_int, ctypes.POINTER(RunResult), ctypes.c_char_p, ctypes.c_int]
    _host = h
    return h


def parse_input(path):
    """Parsed input file (Input); raises VLP2DError on a malformed file."""
    h = load_host()
    out = Input()
    err = ctypes.create_string_buffer(4096)
    rc = h.vlp2d_input_parse(path.encode(), ctypes.byref(out), err, len(err))
    if rc != 0:
        raise _lib.VLP2DError(rc, err.value.decode())
    return out


def run_input(inp, init_kind=_lib.INIT_MEAN, device=0, options=None, max_cycles=0):
    """One MultiGrid run of a parsed input on the GPU (RunResult)."""
    h = load_host()
    out = RunResult()
    err = ctypes.create_string_buffer(4096)
    opts = options.to_c() if options is not None else None
    rc = h.vlp2d_run_input(ctypes.byref(inp), int(init_kind), int(device), ctypes.byref(opts) if opts else None,
                           int(max_cycles), ctypes.byref(out), err, len(err))
    if rc != 0:
        raise _lib.VLP2DError(rc, err.value.decode())
    return out
