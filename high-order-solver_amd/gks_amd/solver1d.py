"""1D GKS path on the GPU (lib/libgks1d.so, include/gks1d.h).

``Solver1D`` runs the reference's SodTubeProblem time loop
(reference ``src/1Dproblem.cpp:79-148``) inside one persistent workgroup;
every call goes through the C ABI and raises if the library is missing.
"""
import ctypes
import os

import numpy as np

from .params import GKS_OK, STATUS_NAMES, Params1D

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB1_PATH = os.path.join(PKG_DIR, "lib", "libgks1d.so")

_D = ctypes.POINTER(ctypes.c_double)
_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32P = ctypes.POINTER(ctypes.c_int32)

SIGS1 = {
    "gks1d_create": (ctypes.c_int, [ctypes.POINTER(Params1D), ctypes.POINTER(_P)]),
    "gks1d_destroy": (ctypes.c_int, [_P]),
    "gks1d_set_convar": (ctypes.c_int, [_P, _D, _I64, _I64]),
    "gks1d_get_convar": (ctypes.c_int, [_P, _D, _I64, _I64]),
    "gks1d_get_alpha": (ctypes.c_int, [_P, _D, _I64]),
    "gks1d_get_face_fluxes": (ctypes.c_int, [_P, ctypes.c_int32, _D]),
    "gks1d_run": (ctypes.c_int, [_P, _I32P, ctypes.c_double, ctypes.c_int32, _D, _I32P, _D]),
    "gks1d_flux_points": (ctypes.c_int, [ctypes.POINTER(Params1D), _I64, _D, _D, ctypes.c_double, _D, _D, _D]),
    "gks1d_last_error": (ctypes.c_char_p, []),
    "gks1d_boundary": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32]),
    "gks1d_copy_old": (ctypes.c_int, [_P]),
    "gks1d_cfl": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_double, _D]),
    "gks1d_reconstruct": (ctypes.c_int, [_P, ctypes.c_int32]),
    "gks1d_flux": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_double]),
    "gks1d_update": (ctypes.c_int, [_P, ctypes.c_int32]),
    "gks1d_update_alpha": (ctypes.c_int, [_P]),
    "gks1d_run_noalloc": (ctypes.c_int, [ctypes.POINTER(Params1D), _P, _P, _I32P, ctypes.c_double, ctypes.c_int32,
                                         _D, _I32P, _D, _P]),
}

_lib1 = None


class Gks1dError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


def lib1():
    """Loads lib/libgks1d.so (raises if it has not been built)."""
    global _lib1
    if _lib1 is None:
        if not os.path.exists(LIB1_PATH):
            raise RuntimeError(f"HIP library not built: {LIB1_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB1_PATH)
        for name, (res, args) in SIGS1.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib1 = L
    return _lib1


def _check(rc):
    if rc != GKS_OK:
        raise Gks1dError(rc, lib1().gks1d_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(_D)


def flux_points_1d(p, left, right, dt):
    """Center_collision + GKS for n points on the GPU (n x 6 inputs)."""
    left = np.ascontiguousarray(left, dtype=np.float64)
    right = np.ascontiguousarray(right, dtype=np.float64)
    n = left.shape[0]
    c = np.zeros((n, 6))
    f = np.zeros((n, 3))
    d = np.zeros((n, 3))
    _check(lib1().gks1d_flux_points(ctypes.byref(p), n, _dp(left), _dp(right), dt, _dp(c), _dp(f), _dp(d)))
    return c, f, d


def prim_to_convar_1d(prim, gamma):
    """Primvar_to_convar_1D (reference basic_function.cpp:853-858) over [3, n]."""
    rho, u, p = prim
    return np.stack([rho, rho * u, rho * (p / rho / (gamma - 1) + 0.5 * (u * u))])


class Solver1D:
    """One 1D tube on the GPU.  Host arrays: convar [3, nx] incl. ghosts."""

    def __init__(self, p):
        self.p = p.copy()
        self.nx = p.nx
        h = _P()
        _check(lib1().gks1d_create(ctypes.byref(self.p), ctypes.byref(h)))
        self.h = h
        self.t = 0.0
        self.step_count = 0

    def close(self):
        if getattr(self, "h", None):
            lib1().gks1d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_convar(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        assert u.shape == (3, self.nx)
        _check(lib1().gks1d_set_convar(self.h, _dp(u), 8, 8 * self.nx))

    def get_convar(self):
        u = np.empty((3, self.nx))
        _check(lib1().gks1d_get_convar(self.h, _dp(u), 8, 8 * self.nx))
        return u

    def get_alpha(self):
        a = np.empty(self.nx)
        _check(lib1().gks1d_get_alpha(self.h, _dp(a), 8))
        return a

    def init_prim(self, prim):
        """ICforSodTube-style start: convar = Primvar_to_convar_1D(primvar)."""
        with np.errstate(invalid="ignore", divide="ignore"):
            self.set_convar(prim_to_convar_1d(np.asarray(prim, dtype=np.float64), self.p.gamma))
        self.t = 0.0
        self.step_count = 0

    def face_fluxes(self, stage):
        out = np.zeros((self.nx + 1, 6))
        _check(lib1().gks1d_get_face_fluxes(self.h, stage, _dp(out)))
        return out

    def run(self, bc, tstop, max_steps=10_000_000, record_dt=False):
        """The driver's loop on the device; returns steps taken (and the dt
        of each step when record_dt)."""
        bca = (ctypes.c_int32 * 2)(*bc)
        t = ctypes.c_double(self.t)
        s = ctypes.c_int32(self.step_count)
        hist = np.zeros(max(1, int(max_steps))) if record_dt else None
        rc = lib1().gks1d_run(self.h, bca, tstop, int(max_steps), ctypes.byref(t), ctypes.byref(s),
                              _dp(hist) if record_dt else None)
        n = s.value - self.step_count
        self.t, self.step_count = t.value, s.value
        _check(rc)
        return (n, hist[:n]) if record_dt else n

    # one sweep at a time (the reference's loop on the host)
    def boundary(self, side, kind):
        _check(lib1().gks1d_boundary(self.h, side, kind))

    def copy_old(self):
        _check(lib1().gks1d_copy_old(self.h))

    def cfl(self, tstop):
        dt = ctypes.c_double()
        _check(lib1().gks1d_cfl(self.h, self.t, tstop, ctypes.byref(dt)))
        return dt.value

    def reconstruct(self, with_df):
        _check(lib1().gks1d_reconstruct(self.h, 1 if with_df else 0))

    def flux(self, stage, dt):
        _check(lib1().gks1d_flux(self.h, stage, dt))

    def update(self, stage):
        _check(lib1().gks1d_update(self.h, stage))

    def update_alpha(self):
        _check(lib1().gks1d_update_alpha(self.h))
