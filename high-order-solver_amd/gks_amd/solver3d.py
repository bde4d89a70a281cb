"""3D GKS path on the GPU (lib/libgks3d.so, include/gks3d.h).

``Solver3D`` runs the reference's CubicTube time loop (reference
``src/3Dproblem.cpp:109-166``: WENO5-AO face / line / point reconstruction,
GKS3D at 2 x 2 Gauss points per face, S2O4) with lambda from the
conservative state and free or periodic boundaries; every call goes through
the C ABI and raises if the library is missing.
"""
import ctypes
import os

import numpy as np

from .params import GKS_OK, STATUS_NAMES, Params3D
from .problems import prim_to_convar_3d

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB3_PATH = os.path.join(PKG_DIR, "lib", "libgks3d.so")

_D = ctypes.POINTER(ctypes.c_double)
_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32P = ctypes.POINTER(ctypes.c_int32)

SIGS3 = {
    "gks3d_create": (ctypes.c_int, [ctypes.POINTER(Params3D), ctypes.POINTER(_P)]),
    "gks3d_destroy": (ctypes.c_int, [_P]),
    "gks3d_set_convar": (ctypes.c_int, [_P, _D, _I64, _I64]),
    "gks3d_get_convar": (ctypes.c_int, [_P, _D, _I64, _I64]),
    "gks3d_set_cfl_primvar": (ctypes.c_int, [_P, _D, _I64, _I64]),
    "gks3d_run": (ctypes.c_int, [_P, _I32P, ctypes.c_double, ctypes.c_int32, _D, _I32P]),
    "gks3d_sync": (ctypes.c_int, [_P]),
    "gks3d_boundary": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32]),
    "gks3d_primvar": (ctypes.c_int, [_P]),
    "gks3d_copy_old": (ctypes.c_int, [_P]),
    "gks3d_cfl": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_double, _D]),
    "gks3d_flux": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_double]),
    "gks3d_update": (ctypes.c_int, [_P, ctypes.c_int32]),
    "gks3d_timing": (ctypes.c_int, [_P, ctypes.c_int]),
    "gks3d_timing_phases": (ctypes.c_int, [_P, ctypes.c_uint32]),
    "gks3d_get_timing": (ctypes.c_int, [_P, _D, ctypes.POINTER(ctypes.c_int64)]),
    "gks3d_flux_points": (ctypes.c_int, [ctypes.POINTER(Params3D), _I64, _D, _D, ctypes.c_double, _D, _D, _D]),
    "gks3d_last_error": (ctypes.c_char_p, []),
    "gks3d_set_classifier": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_float), ctypes.c_float]),
    "gks3d_classifier_counts": (ctypes.c_int, [_P, ctypes.POINTER(_I64)]),
}

PHASES3 = ["bc", "primvar", "face", "line", "flux", "update", "copy_old", "classify"]

_lib3 = None


class Gks3dError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


def lib3():
    """Loads lib/libgks3d.so (raises if it has not been built)."""
    global _lib3
    if _lib3 is None:
        if not os.path.exists(LIB3_PATH):
            raise RuntimeError(f"HIP library not built: {LIB3_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB3_PATH)
        for name, (res, args) in SIGS3.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib3 = L
    return _lib3


def _check(rc):
    if rc != GKS_OK:
        raise Gks3dError(rc, lib3().gks3d_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(_D)


def flux_points_3d(p, left, right, dt):
    """GKS3D (gks2nd, S2O4 pair) at n Gauss points on the GPU: left / right
    n x 20 (convar, der1x, der1y, der1z in the face frame) -> (center n x
    20, f n x 5, derf n x 5)."""
    left = np.ascontiguousarray(left, dtype=np.float64)
    right = np.ascontiguousarray(right, dtype=np.float64)
    n = left.shape[0]
    c, f, d = np.zeros((n, 20)), np.zeros((n, 5)), np.zeros((n, 5))
    _check(lib3().gks3d_flux_points(ctypes.byref(p), n, _dp(left), _dp(right), dt, _dp(c), _dp(f), _dp(d)))
    return c, f, d


class Solver3D:
    """One cube on the GPU.  Host arrays: [5, nx, ny, nz] incl. ghosts
    (z fastest, as the reference's cell index)."""

    def __init__(self, p):
        self.p = p.copy()
        self.shape = (5, p.nx, p.ny, p.nz)
        h = _P()
        _check(lib3().gks3d_create(ctypes.byref(self.p), ctypes.byref(h)))
        self.h = h
        self.t = 0.0
        self.step_count = 0

    def close(self):
        if getattr(self, "h", None):
            lib3().gks3d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _strides(self):
        nc = self.shape[1] * self.shape[2] * self.shape[3]
        return 8, 8 * nc

    def set_convar(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        assert u.shape == self.shape
        _check(lib3().gks3d_set_convar(self.h, _dp(u), *self._strides()))

    def get_convar(self):
        u = np.empty(self.shape)
        _check(lib3().gks3d_get_convar(self.h, _dp(u), *self._strides()))
        return u

    def set_cfl_primvar(self, prim):
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        assert prim.shape == self.shape
        _check(lib3().gks3d_set_cfl_primvar(self.h, _dp(prim), *self._strides()))

    def init_prim(self, prim):
        """ICfor_cubic_sod-style start: convar = Primvar_to_convar_3D(primvar),
        the first CFL from primvar."""
        prim = np.asarray(prim, dtype=np.float64)
        with np.errstate(invalid="ignore", divide="ignore"):
            self.set_convar(prim_to_convar_3d(prim, self.p.gamma))
        self.set_cfl_primvar(prim)
        self.t = 0.0
        self.step_count = 0

    def init_convar(self, u, prim=None):
        """Start from conservative variables; the first CFL from ``prim`` or
        from the primitive variables of ``u``."""
        from .problems import convar_to_prim_3d
        u = np.asarray(u, dtype=np.float64)
        self.set_convar(u)
        if prim is None:
            with np.errstate(invalid="ignore", divide="ignore"):
                prim = convar_to_prim_3d(u, self.p.gamma)
        self.set_cfl_primvar(prim)
        self.t = 0.0
        self.step_count = 0

    def run(self, bc, tstop, max_steps=2**31 - 1):
        """The driver's loop; bc[6] per side (x-, x+, y-, y+, z-, z+).
        Returns the number of steps taken."""
        bca = (ctypes.c_int32 * 6)(*bc)
        t = ctypes.c_double(self.t)
        s = ctypes.c_int32(self.step_count)
        rc = lib3().gks3d_run(self.h, bca, tstop, int(max_steps), ctypes.byref(t), ctypes.byref(s))
        n = s.value - self.step_count
        self.t, self.step_count = t.value, s.value
        _check(rc)
        return n

    def set_classifier(self, params, threshold):
        """The learned reconstruction (gks3d_set_classifier): the classifier
        of gks_learn.h (params: its float vector, None disables it)."""
        if params is None:
            _check(lib3().gks3d_set_classifier(self.h, None, 0.0))
            return
        P = np.ascontiguousarray(params, dtype=np.float32)
        _check(lib3().gks3d_set_classifier(self.h, P.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), float(threshold)))

    def classifier_counts(self):
        """(linear, total) reconstructions since the classifier was set, per
        kind: face values, line values, point values."""
        c = (_I64 * 6)()
        _check(lib3().gks3d_classifier_counts(self.h, c))
        return {"face": (int(c[0]), int(c[1])), "line": (int(c[2]), int(c[3])), "point": (int(c[4]), int(c[5]))}

    def sync(self):
        _check(lib3().gks3d_sync(self.h))

    # one sweep at a time (the reference's loop on the host)
    def boundary(self, side, kind):
        _check(lib3().gks3d_boundary(self.h, side, kind))

    def primvar(self):
        _check(lib3().gks3d_primvar(self.h))

    def copy_old(self):
        _check(lib3().gks3d_copy_old(self.h))

    def cfl(self, tstop):
        dt = ctypes.c_double()
        _check(lib3().gks3d_cfl(self.h, self.t, tstop, ctypes.byref(dt)))
        return dt.value

    def flux(self, stage, dt):
        _check(lib3().gks3d_flux(self.h, stage, dt))

    def update(self, stage):
        _check(lib3().gks3d_update(self.h, stage))

    def timing(self, on=True, phases=None):
        """Resets and (with on) starts per-phase HIP-event timing, of every
        phase or only of the named ones."""
        if on and phases is not None:
            mask = 0
            for ph in phases:
                mask |= 1 << PHASES3.index(ph)
            _check(lib3().gks3d_timing_phases(self.h, mask))
        else:
            _check(lib3().gks3d_timing(self.h, 1 if on else 0))

    def get_timing(self):
        """{phase: (summed ms, launches)} since timing(True)."""
        ms = np.zeros(len(PHASES3))
        n = (ctypes.c_int64 * len(PHASES3))()
        _check(lib3().gks3d_get_timing(self.h, _dp(ms), n))
        return {ph: (float(ms[k]), int(n[k])) for k, ph in enumerate(PHASES3)}
