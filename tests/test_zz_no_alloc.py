"""nsp_diagnostics_noalloc (the allocation-free variant of nsp_diagnostics)
through the C ABI on the GPU: two calls in a row, the second on a smaller
grid, each against nsp_diagnostics on the same state.  Tolerances: the
maxima (wall velocity, wall divergence) are identical; KE and the divergence
norm are sums combined in another order (relative 1e-13); the centre-line
velocity, friction velocity and flux use the same sums on the device instead
of the host (relative 1e-14)."""
import ctypes

import numpy as np
import pytest

import nsp

pytestmark = pytest.mark.gpu

KEYS = ["u_cl", "u_tau", "ke", "div_l2", "wall_max", "wall_div_max", "flux"]
TOL = dict(u_cl=1e-14, u_tau=1e-14, ke=1e-13, div_l2=1e-13, wall_max=0.0, wall_div_max=0.0, flux=1e-14)


def _hip():
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    return hip


def _noalloc(hip, pipe, dev):
    st = nsp.lib().nsp_diagnostics_noalloc(pipe.h, dev)
    assert st == 0, nsp.lib().nsp_last_error()
    assert hip.hipStreamSynchronize(ctypes.c_void_p(pipe.stream())) == 0
    out = np.zeros(7)
    assert hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), dev, 7 * 8, 2) == 0  # device to host
    return dict(zip(KEYS, out))


def test_noalloc_diagnostics_match_existing_entry_point():
    hip = _hip()
    dev = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dev), 7 * 8) == 0
    shapes = [dict(nr=32, nt=64, nz=65, lz=4 * np.pi, re=2000.0, dt=0.005),
              dict(nr=16, nt=8, nz=9, lz=2 * np.pi, re=500.0, dt=0.01)]
    try:
        for j, shape in enumerate(shapes):
            print(f"[gpu] noalloc diagnostics call {j} {shape}", flush=True)
            p = nsp.Pipe(**shape)
            p.seeded_ic(20261202 + j, 4, 4, 0, 0.05)
            p.step(2)
            want = p.diagnostics()
            got = _noalloc(hip, p, dev)
            for k in KEYS:
                assert abs(got[k] - want[k]) <= TOL[k] * abs(want[k]), (j, k, got[k], want[k])
            p.close()
    finally:
        hip.hipFree(dev)
