"""The C-ABI library loads and exports every entry point include/vlp2d_mg.h
declares; argument validation that happens before any device call."""
import ctypes
import os
import re

import pytest

from vlp2d_amd import _lib
from vlp2d_amd._lib import load_library

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "vlp2d_mg.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(vlp2d_(?:mg|expr)_\w+)\s*\(", txt)))


def test_header_matches_binding_table():
    assert _declared() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = load_library()
    for name in _declared():
        assert hasattr(lib, name), name


def test_default_opts_and_version():
    lib = load_library()
    o = _lib.Opts()
    lib.vlp2d_mg_default_opts(ctypes.byref(o))
    assert (o.nu1, o.nu2, o.nu_coarse, o.precision, o.flags, o.coarse_nodes, o.max_inner, o.smoother, o.omega,
            o.restriction, o.cycle) == (3, 3, 3, 0, 0, 0, 1, 0, 0.8, 0, 0)
    assert lib.vlp2d_mg_version() >> 16 == 1


@pytest.mark.parametrize("sx,sy", [(6, 8), (8, 2), (0, 8), (65536, 4), (-4, 4)])
def test_create_rejects_bad_sizes_before_touching_a_device(sx, sy):
    lib = load_library()
    h = ctypes.c_void_p()
    rc = lib.vlp2d_mg_create(sx, sy, 0.1, 0.1, None, 1e-13, 0, None, ctypes.byref(h))
    assert rc == _lib.E_SIZE and not h.value
    assert b"power" in lib.vlp2d_mg_last_error(None)


def test_create_rejects_bad_steps_and_options():
    lib = load_library()
    h = ctypes.c_void_p()
    assert lib.vlp2d_mg_create(8, 8, -1.0, 0.1, None, 1e-13, 0, None, ctypes.byref(h)) == _lib.E_ARG
    o = _lib.Opts(2, 2, 0, 0, 0, 0, 1, 0, 0.8)
    assert lib.vlp2d_mg_create(8, 8, 0.1, 0.1, None, 1e-13, 0, ctypes.byref(o), ctypes.byref(h)) == _lib.E_ARG
    o = _lib.Opts(2, 2, 1, 7, 0, 0, 1, 0, 0.8)
    assert lib.vlp2d_mg_create(8, 8, 0.1, 0.1, None, 1e-13, 0, ctypes.byref(o), ctypes.byref(h)) == _lib.E_ARG
    for inner in (0, 9, -1):  # FP32 cycles per refinement step: 1..8
        o = _lib.Opts(2, 2, 3, 1, 0, 0, inner, 0, 0.8)
        assert lib.vlp2d_mg_create(8, 8, 0.1, 0.1, None, 1e-13, 0, ctypes.byref(o), ctypes.byref(h)) == _lib.E_ARG
        assert b"max_inner" in lib.vlp2d_mg_last_error(None)
    for sm, w in ((2, 0.8), (1, 0.0), (1, 1.5), (1, float("nan"))):  # smoother 0/1, omega in (0, 1]
        o = _lib.Opts(2, 2, 3, 0, 0, 0, 1, sm, w)
        assert lib.vlp2d_mg_create(8, 8, 0.1, 0.1, None, 1e-13, 0, ctypes.byref(o), ctypes.byref(h)) == _lib.E_ARG
        assert b"omega" in lib.vlp2d_mg_last_error(None)
    for rs in (2, -1):  # restriction 0/1
        o = _lib.Opts(2, 2, 3, 0, 0, 0, 1, 0, 0.8, rs)
        assert lib.vlp2d_mg_create(8, 8, 0.1, 0.1, None, 1e-13, 0, ctypes.byref(o), ctypes.byref(h)) == _lib.E_ARG
        assert b"restriction" in lib.vlp2d_mg_last_error(None)
    for prec, cyc in ((0, 2), (0, -1), (1, 1)):  # cycle 0/1; FMG in FP64 only
        o = _lib.Opts(2, 2, 3, prec, 0, 0, 1, 0, 0.8, 0, cyc)
        assert lib.vlp2d_mg_create(8, 8, 0.1, 0.1, None, 1e-13, 0, ctypes.byref(o), ctypes.byref(h)) == _lib.E_ARG
        assert b"cycle" in lib.vlp2d_mg_last_error(None)
    assert lib.vlp2d_mg_create(8, 8, 0.1, 0.1, None, 1e-13, 0, None, None) == _lib.E_ARG


def test_create_rejects_nonfinite_rhs():
    import numpy as np
    lib = load_library()
    f = np.zeros((9, 9))
    f[3, 4] = np.nan
    h = ctypes.c_void_p()
    rc = lib.vlp2d_mg_create(8, 8, 0.1, 0.1, _lib.ptr(f), 1e-13, 0, None, ctypes.byref(h))
    assert rc == _lib.E_NONFINITE


def test_null_handle_calls_are_rejected():
    lib = load_library()
    assert lib.vlp2d_mg_cycle(None, None) == _lib.E_ARG
    assert lib.vlp2d_mg_solve(None, 1e-10, 10, None) == _lib.E_ARG
    lib.vlp2d_mg_destroy(None)
    lib.vlp2d_mg_cancel(None)


def test_noalloc_solve_checks_arguments_before_touching_a_device():
    lib = load_library()
    st = _lib.Stats()
    fake = 1 << 20  # never dereferenced: every case fails before a device call
    call = lambda sx, sy, h, u, pitch, rtol=1e-10, o=None: lib.vlp2d_mg_solve_noalloc(
        sx, sy, h, h, None, u, pitch, rtol, 10, o, 0, None, ctypes.byref(st))
    assert call(6, 8, 0.1, fake, 32) == _lib.E_SIZE
    assert call(8, 8, -0.1, fake, 32) == _lib.E_ARG
    assert call(8, 8, 0.1, None, 32) == _lib.E_ARG
    assert call(8, 8, 0.1, fake, 9) == _lib.E_ARG and b"pitch" in lib.vlp2d_mg_last_error(None)
    assert call(8, 8, 0.1, fake + 8, 32) == _lib.E_ARG and b"aligned" in lib.vlp2d_mg_last_error(None)
    assert call(8, 8, 0.1, fake, 32, rtol=-1.0) == _lib.E_ARG
    o = _lib.Opts()
    lib.vlp2d_mg_default_opts(ctypes.byref(o))
    o.flags = _lib.FLAG_PROFILE
    assert call(8, 8, 0.1, fake, 32, o=ctypes.byref(o)) == _lib.E_ARG
    assert b"PROFILE" in lib.vlp2d_mg_last_error(None)
