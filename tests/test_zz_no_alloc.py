"""The allocation-free 1D run (gks1d_run_noalloc, include/gks1d.h) against
the context-based gks1d_run on the same inputs, through the C ABI.

Both run the same kernels, so the bar is bit-for-bit equality of the final
state, the DF sensor, the time, the step count and every step's dt.  Two
calls in a row share the GPU's scratch array: the first a 400-cell Sod tube
with the DF sensor (fast kernels), the second a smaller 200-cell Shu-Osher
tube (IEEE kernels), so the second call runs over what the first one left
in the scratch array.  A third pair continues a run across two calls with
the caller's own DF-sensor array.

Device buffers for the caller's side are allocated here through the HIP
runtime the library itself links (the entry point allocates nothing).
"""
import ctypes
import sys

import numpy as np
import pytest

from gks_amd.params import BC_FREE, BC_REFLECT, shu_osher_params, sod_params, blastwave_params
from gks_amd.problems import blastwave_prim, shu_osher_prim, sod_prim
from gks_amd.solver1d import Solver1D, lib1, prim_to_convar_1d

pytestmark = pytest.mark.gpu

_D = ctypes.POINTER(ctypes.c_double)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libgks1d.so links."""

    def __init__(self):
        lib1()  # loads libamdhip64.so.7 as a dependency first
        self.h = ctypes.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.h.hipDeviceSynchronize.argtypes = []

    def alloc_copy(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, shape):
        out = np.empty(shape)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self, p):
        assert self.h.hipFree(p) == 0


def run_ctx(p, u0, bc, tstop, max_steps):
    s = Solver1D(p)
    s.set_convar(u0)
    n, dts = s.run(bc, tstop, max_steps, record_dt=True)
    out = dict(u=s.get_convar(), alpha=s.get_alpha(), t=s.t, steps=s.step_count, dt=dts)
    s.close()
    return out


def run_noalloc(hip, p, dU, dA, bc, tstop, max_steps, t=0.0, step=0):
    bca = (ctypes.c_int32 * 2)(*bc)
    tt = ctypes.c_double(t)
    ss = ctypes.c_int32(step)
    hist = np.zeros(max(1, max_steps))
    rc = lib1().gks1d_run_noalloc(ctypes.byref(p), dU, dA, bca, tstop, max_steps, ctypes.byref(tt),
                                  ctypes.byref(ss), hist.ctypes.data_as(_D), None)
    assert rc == 0, lib1().gks1d_last_error()
    return tt.value, ss.value, hist[:ss.value - step]


def assert_same(a, b_u, b_t, b_steps, b_dt, b_alpha=None):
    assert b_steps == a["steps"], (b_steps, a["steps"])
    assert b_t == a["t"]
    np.testing.assert_array_equal(b_dt, a["dt"])
    np.testing.assert_array_equal(b_u, a["u"])
    if b_alpha is not None:
        np.testing.assert_array_equal(b_alpha, a["alpha"])


CASES = [
    ("sod400_df_fast", sod_params(400, use_df=1, fast_math=1), sod_prim, [BC_FREE, BC_FREE], 0.2),
    ("shu_osher200_df_ieee", shu_osher_params(200, use_df=1, fast_math=0), shu_osher_prim, [BC_FREE, BC_FREE], 0.3),
]


def test_noalloc_matches_gks1d_run_two_calls_in_a_row():
    hip = Hip()
    for name, p, icf, bc, tstop in CASES:
        say(f"no-alloc: {name} (nodex {p.nodex}, fast_math {p.fast_math})")
        with np.errstate(invalid="ignore", divide="ignore"):
            u0 = prim_to_convar_1d(icf(p), p.gamma)
        ref = run_ctx(p, u0, bc, tstop, 100000)
        dU = hip.alloc_copy(u0)
        try:
            t, steps, dts = run_noalloc(hip, p, dU, None, bc, tstop, 100000)
            u = hip.get(dU, (3, p.nx))
        finally:
            hip.free(dU)
        assert_same(ref, u, t, steps, dts)
        say(f"no-alloc: {name} {steps} steps, bit-identical to gks1d_run")


def test_noalloc_continues_a_run_with_the_callers_alpha():
    """Two calls that continue one run (k1 then k2 steps) with the caller's
    DF-sensor array equal one context running the same two stretches."""
    hip = Hip()
    p = blastwave_params(300, use_df=1, fast_math=1)
    bc = [BC_REFLECT, BC_REFLECT]
    u0 = prim_to_convar_1d(blastwave_prim(p), p.gamma)
    s = Solver1D(p)
    s.set_convar(u0)
    s.run(bc, 0.038, 40)
    n2, dts2 = s.run(bc, 0.038, 60, record_dt=True)
    ref = dict(u=s.get_convar(), alpha=s.get_alpha(), t=s.t, steps=s.step_count, dt=dts2)
    s.close()
    dU = hip.alloc_copy(u0)
    dA = hip.alloc_copy(np.zeros(p.nx))
    try:
        t, step, _ = run_noalloc(hip, p, dU, dA, bc, 0.038, 40)
        t, step, dts = run_noalloc(hip, p, dU, dA, bc, 0.038, 60, t=t, step=step)
        u = hip.get(dU, (3, p.nx))
        a = hip.get(dA, (p.nx,))
    finally:
        hip.free(dU)
        hip.free(dA)
    assert_same(ref, u, t, step, dts, a)
    say(f"no-alloc: blast wave 300 cells, 40 + 60 steps with the caller's alpha, bit-identical")
