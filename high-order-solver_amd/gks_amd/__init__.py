"""gks_amd — Python host over the MI355X GKS hot path (lib/libgks2d.so).

The library is the product: every method below calls the C ABI declared in
include/gks2d.h.  There is no CPU fallback; if the HIP library is missing or
no GPU is visible, the calls raise.
"""
import ctypes
import os

import numpy as np

from .params import (BC_FREE, BC_PERIODIC, BC_REFLECT, GKS_OK, STATUS_NAMES, Params,  # noqa: F401
                     planar_shock_params)

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libgks2d.so")

_D = ctypes.POINTER(ctypes.c_double)
_P = ctypes.c_void_p
_I64 = ctypes.c_int64

# exported symbol -> (restype, argtypes)
_SIGS = {
    "gks2d_create": (ctypes.c_int, [ctypes.POINTER(Params), ctypes.POINTER(_P)]),
    "gks2d_destroy": (ctypes.c_int, [_P]),
    "gks2d_set_convar": (ctypes.c_int, [_P, _D, _I64, _I64]),
    "gks2d_get_convar": (ctypes.c_int, [_P, _D, _I64, _I64]),
    "gks2d_set_alpha": (ctypes.c_int, [_P, _D, _D, _I64]),
    "gks2d_get_alpha": (ctypes.c_int, [_P, _D, _D, _I64]),
    "gks2d_set_cfl_primvar": (ctypes.c_int, [_P, _D, _I64, _I64]),
    "gks2d_copy_old": (ctypes.c_int, [_P]),
    "gks2d_cfl": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_double, _D]),
    "gks2d_boundary": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int]),
    "gks2d_primvar": (ctypes.c_int, [_P]),
    "gks2d_reconstruct": (ctypes.c_int, [_P]),
    "gks2d_flux": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_double]),
    "gks2d_update": (ctypes.c_int, [_P, ctypes.c_int]),
    "gks2d_update_alpha": (ctypes.c_int, [_P]),
    "gks2d_step": (ctypes.c_int, [_P, ctypes.c_double, ctypes.POINTER(ctypes.c_int32)]),
    "gks2d_run": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int32), ctypes.c_double, ctypes.c_int32, _D,
                                 ctypes.POINTER(ctypes.c_int32)]),
    "gks2d_sync": (ctypes.c_int, [_P]),
    "gks2d_flux_points": (ctypes.c_int, [ctypes.POINTER(Params), _I64, _D, _D, ctypes.c_double, _D, _D, _D]),
    "gks2d_get_face_fluxes": (ctypes.c_int, [_P, _D, _D]),
    "gks2d_keep_face_fluxes": (ctypes.c_int, [_P, ctypes.c_int]),
    "gks2d_timing": (ctypes.c_int, [_P, ctypes.c_int]),
    "gks2d_timing_phases": (ctypes.c_int, [_P, ctypes.c_uint32]),
    "gks2d_get_timing": (ctypes.c_int, [_P, _D, ctypes.POINTER(ctypes.c_int64)]),
    "gks2d_last_error": (ctypes.c_char_p, []),
}

PHASES = ["bc", "primvar", "normal_x", "normal_y", "tangent_x", "tangent_y", "flux_x", "flux_y", "update",
          "alpha", "copy_old", "normal_list_x", "normal_list_y", "flux_list_x", "flux_list_y"]

_lib = None


class GksError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


def lib():
    """Loads lib/libgks2d.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != GKS_OK:
        raise GksError(rc, lib().gks2d_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(_D)


def flux_points(p, left, right, dt):
    """Centre state + GKS2D flux for n points on the GPU (n x 12 inputs)."""
    left = np.ascontiguousarray(left, dtype=np.float64)
    right = np.ascontiguousarray(right, dtype=np.float64)
    n = left.shape[0]
    c = np.zeros((n, 12))
    f = np.zeros((n, 4))
    d = np.zeros((n, 4))
    _check(lib().gks2d_flux_points(ctypes.byref(p), n, _dp(left), _dp(right), dt, _dp(c), _dp(f), _dp(d)))
    return c, f, d


class Solver2D:
    """One 2D block on the GPU: the reference's per-step sweeps on device data.

    State arrays on the host side are SoA float64 [4, nx, ny] (conservative
    variables incl. ghost cells), alpha [nx, ny].
    """

    def __init__(self, p):
        self.p = p.copy()
        self.nx, self.ny = p.nx, p.ny
        self.gp = p.gausspoint
        h = _P()
        _check(lib().gks2d_create(ctypes.byref(self.p), ctypes.byref(h)))
        self.h = h
        self.t = 0.0
        self.step_count = 0

    def close(self):
        if getattr(self, "h", None):
            lib().gks2d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- state
    @property
    def nc(self):
        return self.nx * self.ny

    def set_convar(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        assert u.shape == (4, self.nx, self.ny)
        _check(lib().gks2d_set_convar(self.h, _dp(u), 8, 8 * self.nc))

    def get_convar(self):
        u = np.empty((4, self.nx, self.ny))
        _check(lib().gks2d_get_convar(self.h, _dp(u), 8, 8 * self.nc))
        return u

    def set_alpha(self, ax, ay):
        ax = np.ascontiguousarray(ax, dtype=np.float64)
        ay = np.ascontiguousarray(ay, dtype=np.float64)
        _check(lib().gks2d_set_alpha(self.h, _dp(ax), _dp(ay), 8))

    def get_alpha(self):
        ax = np.empty((self.nx, self.ny))
        ay = np.empty((self.nx, self.ny))
        _check(lib().gks2d_get_alpha(self.h, _dp(ax), _dp(ay), 8))
        return ax, ay

    def set_cfl_primvar(self, prim):
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        _check(lib().gks2d_set_cfl_primvar(self.h, _dp(prim), 8, 8 * self.nc))

    def init_prim(self, prim):
        """Initial condition from primitive variables, as IC_for_riemann_2d
        does: convar = Primvar_to_convar_2D(primvar) for every cell, and the
        step-0 CFL reads the primitive field itself."""
        from .problems import prim_to_convar
        self.set_convar(prim_to_convar(prim, self.p.gamma))
        self.set_cfl_primvar(prim)
        self.t = 0.0
        self.step_count = 0

    def init_convar(self, u):
        """Start from given conservative cell averages: the step-0 CFL reads
        their primitive variables (Convar_to_primvar_2D)."""
        from .problems import convar_to_prim
        u = np.asarray(u, dtype=np.float64)
        self.set_convar(u)
        with np.errstate(invalid="ignore", divide="ignore"):
            self.set_cfl_primvar(convar_to_prim(u, self.p.gamma))
        self.t = 0.0
        self.step_count = 0

    # ---------------------------------------------------------------- sweeps
    def copy_old(self):
        _check(lib().gks2d_copy_old(self.h))

    def cfl(self, tstop):
        d = ctypes.c_double()
        _check(lib().gks2d_cfl(self.h, self.t, tstop, ctypes.byref(d)))
        return d.value

    def boundary(self, side, kind):
        _check(lib().gks2d_boundary(self.h, side, kind))

    def primvar(self):
        _check(lib().gks2d_primvar(self.h))

    def reconstruct(self):
        _check(lib().gks2d_reconstruct(self.h))

    def flux(self, stage, dt):
        _check(lib().gks2d_flux(self.h, stage, dt))

    def update(self, stage):
        _check(lib().gks2d_update(self.h, stage))

    def update_alpha(self):
        _check(lib().gks2d_update_alpha(self.h))

    def sync(self):
        _check(lib().gks2d_sync(self.h))

    def step(self, dt, bc):
        bca = (ctypes.c_int32 * 4)(*bc)
        _check(lib().gks2d_step(self.h, dt, bca))

    def run(self, bc, tstop, max_steps=2**31 - 1):
        """The driver's time loop on the device; returns steps taken."""
        bca = (ctypes.c_int32 * 4)(*bc)
        t = ctypes.c_double(self.t)
        s = ctypes.c_int32(self.step_count)
        rc = lib().gks2d_run(self.h, bca, tstop, int(max_steps), ctypes.byref(t), ctypes.byref(s))
        n = s.value - self.step_count
        self.t, self.step_count = t.value, s.value
        _check(rc)
        return n

    def timing(self, enable=True, phases=None):
        """Resets and (with enable) starts per-phase HIP-event timing, of
        every phase or only of the named ones (each timed launch costs the
        step two events)."""
        if enable and phases is not None:
            mask = 0
            for ph in phases:
                mask |= 1 << PHASES.index(ph)
            _check(lib().gks2d_timing_phases(self.h, mask))
        else:
            _check(lib().gks2d_timing(self.h, 1 if enable else 0))

    def get_timing(self):
        """{phase: (total ms, launches)} accumulated since timing(True)."""
        ms = np.zeros(len(PHASES))
        n = (ctypes.c_int64 * len(PHASES))()
        _check(lib().gks2d_get_timing(self.h, _dp(ms), n))
        return {ph: (float(ms[k]), int(n[k])) for k, ph in enumerate(PHASES)}

    def keep_face_fluxes(self, on=True):
        _check(lib().gks2d_keep_face_fluxes(self.h, 1 if on else 0))

    def face_fluxes(self):
        xf = np.zeros((self.nx + 1, self.ny, self.gp, 8))
        yf = np.zeros((self.nx, self.ny + 1, self.gp, 8))
        _check(lib().gks2d_get_face_fluxes(self.h, _dp(xf), _dp(yf)))
        return xf, yf
