"""Batches of independent 2D problems (include/gks2d_batch.h).

    b = Batch([planar_shock_params(n) for n in sizes])
    for p, s in enumerate(b.problems):        # Solver2D views of the contexts
        s.init_prim(riemann_prim(s.p, riemann_states(seed + p)))
    b.autotune([BC_FREE] * 4)
    b.run([BC_FREE] * 4, 0.2)

Every problem keeps its own time loop (dt, tstop, step count); one captured
HIP graph per step runs all of them.
"""
import ctypes

import numpy as np

from . import GksError, Solver2D, lib
from .params import Params

_P = ctypes.c_void_p
_I32P = ctypes.POINTER(ctypes.c_int32)
_DP = ctypes.POINTER(ctypes.c_double)
_SIGS = {
    "gks2d_batch_create": (ctypes.c_int, [ctypes.POINTER(Params), ctypes.c_int32, ctypes.POINTER(_P)]),
    "gks2d_batch_destroy": (ctypes.c_int, [_P]),
    "gks2d_batch_problem": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.POINTER(_P)]),
    "gks2d_batch_set_shape": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32]),
    "gks2d_batch_autotune": (ctypes.c_int, [_P, _I32P, ctypes.c_int32, _I32P, _I32P, _DP]),
    "gks2d_batch_run": (ctypes.c_int, [_P, _I32P, _DP, ctypes.c_int32, ctypes.c_int32, _DP, _I32P]),
}
_bound = False


def L():
    global _bound
    lb = lib()
    if not _bound:
        for name, (res, args) in _SIGS.items():
            f = getattr(lb, name)
            f.restype = res
            f.argtypes = args
        _bound = True
    return lb


def _check(rc):
    if rc != 0:
        raise GksError(rc, lib().gks2d_last_error().decode())


class _View(Solver2D):
    """Solver2D over a context the batch owns (never destroyed here)."""

    def __init__(self, p, handle):  # noqa: super().__init__ would create a context
        self.p = p.copy()
        self.nx, self.ny = p.nx, p.ny
        self.gp = p.gausspoint
        self.h = handle
        self.t = 0.0
        self.step_count = 0

    def close(self):
        self.h = None


class Batch:
    def __init__(self, params):
        params = list(params)
        arr = (Params * len(params))(*params)
        h = _P()
        _check(L().gks2d_batch_create(arr, len(params), ctypes.byref(h)))
        self.h = h
        self.problems = []
        for k, p in enumerate(params):
            ph = _P()
            _check(L().gks2d_batch_problem(self.h, k, ctypes.byref(ph)))
            self.problems.append(_View(p, ph))

    def __len__(self):
        return len(self.problems)

    def close(self):
        if getattr(self, "h", None):
            for s in self.problems:
                s.h = None
            L().gks2d_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _bc(self, bc):
        bc = np.asarray(bc, dtype=np.int32)
        if bc.ndim == 1:
            bc = np.tile(bc, (len(self), 1))
        assert bc.shape == (len(self), 4)
        return np.ascontiguousarray(bc)

    def set_shape(self, tanflux_faces=128, largest_first=True):
        _check(L().gks2d_batch_set_shape(self.h, int(tanflux_faces), 1 if largest_first else 0))

    def autotune(self, bc, steps=3):
        """Times every tile shape on the current state (which it restores)
        and keeps the fastest; returns (faces per flux workgroup,
        largest_first, ms per step)."""
        bc = self._bc(bc)
        tf, lf, ms = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_double()
        _check(L().gks2d_batch_autotune(self.h, bc.ctypes.data_as(_I32P), int(steps), ctypes.byref(tf),
                                        ctypes.byref(lf), ctypes.byref(ms)))
        return tf.value, bool(lf.value), ms.value

    def run(self, bc, tstop, max_steps=2**31 - 1, check_every=16):
        """Advances every problem to its tstop (scalar or one per problem);
        returns the steps each took in this call."""
        bc = self._bc(bc)
        tstop = np.broadcast_to(np.asarray(tstop, dtype=np.float64), (len(self),)).copy()
        t = np.array([s.t for s in self.problems], dtype=np.float64)
        st = np.array([s.step_count for s in self.problems], dtype=np.int32)
        st0 = st.copy()
        rc = L().gks2d_batch_run(self.h, bc.ctypes.data_as(_I32P), tstop.ctypes.data_as(_DP), int(max_steps),
                                 int(check_every), t.ctypes.data_as(_DP), st.ctypes.data_as(_I32P))
        for k, s in enumerate(self.problems):
            s.t, s.step_count = float(t[k]), int(st[k])
        _check(rc)
        return st - st0
