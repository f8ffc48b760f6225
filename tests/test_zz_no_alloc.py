"""GPU: the allocation-free solve vlp2d_mg_solve_noalloc, on device arrays the
caller owns, returns the iterate of vlp2d_mg_create + upload + solve +
download on the same inputs bit for bit, with the same cycle count, on two
calls in a row (the second on a smaller grid in the other precision).  The
second call changes the device's free memory by nothing, and a grid whose
hierarchy does not fit the scratch array is rejected before anything runs."""
import ctypes

import numpy as np
import pytest
import torch

from vlp2d_amd import MultiGridSchemeHIP, Options, _lib, problems
from vlp2d_amd._lib import load_library

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _pitch(ny):
    return (ny + 31) // 32 * 32


def _handle_solve(p, u0, opts, rtol, max_cycles):
    s = MultiGridSchemeHIP(p.seg_x, p.seg_y, p.hx, p.hy, p.f, options=opts)
    s.un0[:] = u0
    s.init_after_boundaries_and_initial_iteration_inited()
    st = s.solve(rtol, max_cycles)
    u = s.get_array().copy()
    s.cleanup()
    return u, st


def _device_arrays(p, u0):
    nx, ny = p.seg_x + 1, p.seg_y + 1
    dev = torch.device("cuda:0")
    u = torch.zeros((nx, _pitch(ny)), dtype=torch.float64, device=dev)
    u[:, :ny] = torch.from_numpy(u0).to(dev)
    f = None
    if p.f is not None:
        f = torch.zeros((nx, _pitch(ny)), dtype=torch.float64, device=dev)
        f[:, :ny] = torch.from_numpy(np.ascontiguousarray(p.f)).to(dev)
    torch.cuda.synchronize()
    return u, f


def _noalloc_call(lib, p, u, f, opts, rtol, max_cycles):
    st = _lib.Stats()
    o = opts.to_c()
    rc = lib.vlp2d_mg_solve_noalloc(p.seg_x, p.seg_y, p.hx, p.hy, None if f is None else f.data_ptr(), u.data_ptr(),
                                    _pitch(p.seg_y + 1), rtol, max_cycles, ctypes.byref(o), 0, None, ctypes.byref(st))
    return rc, st


def test_noalloc_solve_matches_handle_solve_twice():
    lib = load_library()
    rtol = 1e-10
    cases = [
        ("1025^2 Gaussian bumps, mixed V(1,3)", problems.config(3, 1024),
         Options.performance(True, nu1=1, nu2=3, max_inner=2)),
        ("513 x 257 sine modes, FP64 V(2,2)", problems.config(2, 512), Options.performance(False)),
    ]
    for k, (name, p, opts) in enumerate(cases):
        u0 = p.initial("zero")
        _say(f"noalloc call {k + 1}: {name}: handle solve")
        ref, st_ref = _handle_solve(p, u0, opts, rtol, 60)
        u, f = _device_arrays(p, u0)
        free0 = torch.cuda.mem_get_info(0)[0]
        _say(f"noalloc call {k + 1}: {name}: allocation-free solve")
        rc, st = _noalloc_call(lib, p, u, f, opts, rtol, 60)
        free1 = torch.cuda.mem_get_info(0)[0]
        assert rc == _lib.OK, lib.vlp2d_mg_last_error(None)
        got = u[:, :p.seg_y + 1].cpu().numpy()
        _say(f"noalloc call {k + 1}: {st.cycles} cycles ({st_ref['cycles']} with a handle), rel residual "
             f"{st.rel_residual:.3e}, {st.seconds * 1e3:.3f} ms, free device memory change {free1 - free0} B")
        assert st.cycles == st_ref["cycles"] and st.vcycles == st_ref["vcycles"]
        assert st.rel_residual == st_ref["rel_residual"]
        assert np.array_equal(got, ref)
        if k > 0:
            # the scratch array was set up by the first call: nothing is allocated now
            assert free1 == free0


def test_noalloc_rejects_what_it_cannot_hold():
    lib = load_library()
    p = problems.config(3, 2048)  # 2049^2: level 0 alone needs more than the 64 MiB scratch
    u, f = _device_arrays(p, p.initial("zero"))
    keep = u.clone()
    rc, _ = _noalloc_call(lib, p, u, f, Options.performance(False), 1e-10, 20)
    assert rc == _lib.E_SIZE
    assert b"MiB of device scratch" in lib.vlp2d_mg_last_error(None)
    assert torch.equal(u, keep)
    st = _lib.Stats()
    o = Options.performance(False).to_c()
    rc = lib.vlp2d_mg_solve_noalloc(p.seg_x, p.seg_y, p.hx, p.hy, f.data_ptr(), u.data_ptr(), p.seg_y + 1, 1e-10,
                                    20, ctypes.byref(o), 0, None, ctypes.byref(st))
    assert rc == _lib.E_ARG and b"pitch" in lib.vlp2d_mg_last_error(None)
