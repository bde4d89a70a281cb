"""ctypes access to the oracle (oracle/liboracle.so, the CPU restatement) and,
when it has been built, to the compiled reference (oracle/_ref/libgksref.so).
Both export the same C API (orc2d_* / ref2d_*).  Test infrastructure only."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libgksref.so")

_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int32)


def dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_D)


def _bind(lib, prefix):
    from gks_amd.params import Params
    pp = ctypes.POINTER(Params)
    sig = {
        "create": (_P, [pp]),
        "destroy": (None, [_P]),
        "set_prim": (ctypes.c_int, [_P, _D]),
        "set_convar": (ctypes.c_int, [_P, _D]),
        "primvar": (ctypes.c_int, [_P]),
        "set_alpha": (ctypes.c_int, [_P, _D, _D]),
        "get": (ctypes.c_int, [_P, _D, _D, _D, _D, _D, _I]),
        "set_time": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_int32]),
        "run": (ctypes.c_int, [_P, _I, ctypes.c_double, ctypes.c_int32, _D]),
        "copy_old": (ctypes.c_int, [_P]),
        "cfl": (ctypes.c_int, [_P, ctypes.c_double, _D]),
        "stage": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_double, _I]),
        "get_face_fluxes": (ctypes.c_int, [_P, ctypes.c_int32, _D, _D]),
        "get_gauss": (ctypes.c_int, [_P, _D, _D]),
        "flux_points": (ctypes.c_int, [pp, ctypes.c_int64, _D, _D, ctypes.c_double, _D, _D, _D]),
        "weno5_normal": (ctypes.c_int, [pp, ctypes.c_int64, _D, _D, ctypes.c_double, _D, _D]),
        "weno5_tangent": (ctypes.c_int, [pp, ctypes.c_int64, _D, _D, ctypes.c_double, _D]),
        "weno_normal": (ctypes.c_int, [pp, ctypes.c_int64, ctypes.c_int32, _D, _D, ctypes.c_double, _D, _D]),
        "weno_tangent": (ctypes.c_int, [pp, ctypes.c_int64, ctypes.c_int32, _D, _D, _D, _D, ctypes.c_double, _D,
                                        _D]),
    }
    fns = {}
    for name, (res, args) in sig.items():
        f = getattr(lib, prefix + name)
        f.restype = res
        f.argtypes = args
        fns[name] = f
    return fns


class CpuSolver:
    """One 2D block in the oracle or in the compiled reference."""

    def __init__(self, fns, p):
        self.f = fns
        self.p = p
        self.h = fns["create"](ctypes.byref(p))
        if not self.h:
            raise ValueError("unsupported parameters")
        self.nx, self.ny = p.nx, p.ny
        self.gp = p.gausspoint

    def close(self):
        if self.h:
            self.f["destroy"](self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_prim(self, prim):
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        assert prim.shape == (4, self.nx, self.ny)
        return self.f["set_prim"](self.h, dptr(prim))

    def set_convar(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        return self.f["set_convar"](self.h, dptr(u))

    def set_alpha(self, ax, ay):
        ax = np.ascontiguousarray(ax, dtype=np.float64)
        ay = np.ascontiguousarray(ay, dtype=np.float64)
        return self.f["set_alpha"](self.h, dptr(ax), dptr(ay))

    def init_convar(self, u):
        """start from conservative cell averages; primvar from them"""
        self.set_convar(u)
        return self.f["primvar"](self.h)

    def get(self):
        u = np.zeros((4, self.nx, self.ny))
        pr = np.zeros((4, self.nx, self.ny))
        ax = np.zeros((self.nx, self.ny))
        ay = np.zeros((self.nx, self.ny))
        t = ctypes.c_double()
        s = ctypes.c_int32()
        self.f["get"](self.h, dptr(u), dptr(pr), dptr(ax), dptr(ay), ctypes.byref(t), ctypes.byref(s))
        return dict(convar=u, primvar=pr, alpha_x=ax, alpha_y=ay, t=t.value, step=s.value)

    def set_time(self, t, step):
        self.f["set_time"](self.h, t, step)

    def run(self, bc, tstop, max_steps):
        bca = (ctypes.c_int32 * 4)(*bc)
        hist = np.zeros(max(1, min(max_steps, 10_000_000)))
        rc = self.f["run"](self.h, bca, tstop, max_steps, dptr(hist))
        return rc, hist

    def copy_old(self):
        self.f["copy_old"](self.h)

    def cfl(self, tstop):
        d = ctypes.c_double()
        self.f["cfl"](self.h, tstop, ctypes.byref(d))
        return d.value

    def stage(self, stage, dt, bc):
        bca = (ctypes.c_int32 * 4)(*bc)
        return self.f["stage"](self.h, stage, dt, bca)

    def face_fluxes(self, stage):
        xf = np.zeros((self.nx + 1, self.ny, self.gp, 8))
        yf = np.zeros((self.nx, self.ny + 1, self.gp, 8))
        self.f["get_face_fluxes"](self.h, stage, dptr(xf), dptr(yf))
        return xf, yf

    def gauss(self):
        xg = np.zeros((self.nx + 1, self.ny, self.gp, 2, 12))
        yg = np.zeros((self.nx, self.ny + 1, self.gp, 2, 12))
        self.f["get_gauss"](self.h, dptr(xg), dptr(yg))
        return xg, yg


class CpuLib:
    def __init__(self, path, prefix):
        self.lib = ctypes.CDLL(path)
        self.fns = _bind(self.lib, prefix)

    def solver(self, p):
        return CpuSolver(self.fns, p)

    def flux_points(self, p, left, right, dt):
        left = np.ascontiguousarray(left, dtype=np.float64)
        right = np.ascontiguousarray(right, dtype=np.float64)
        n = left.shape[0]
        c = np.zeros((n, 12))
        f = np.zeros((n, 4))
        d = np.zeros((n, 4))
        self.fns["flux_points"](ctypes.byref(p), n, dptr(left), dptr(right), dt, dptr(c), dptr(f), dptr(d))
        return c, f, d

    def weno5_normal(self, p, w5, alpha5, h):
        w5 = np.ascontiguousarray(w5, dtype=np.float64)
        alpha5 = np.ascontiguousarray(alpha5, dtype=np.float64)
        n = w5.shape[0]
        lo = np.zeros((n, 8))
        ro = np.zeros((n, 8))
        self.fns["weno5_normal"](ctypes.byref(p), n, dptr(w5), dptr(alpha5), h, dptr(lo), dptr(ro))
        return lo, ro

    def weno5_tangent(self, p, line5, alpha5, h):
        line5 = np.ascontiguousarray(line5, dtype=np.float64)
        alpha5 = np.ascontiguousarray(alpha5, dtype=np.float64)
        n = line5.shape[0]
        out = np.zeros((n, 2, 12))
        self.fns["weno5_tangent"](ctypes.byref(p), n, dptr(line5), dptr(alpha5), h, dptr(out))
        return out


    def weno_normal(self, p, order, w, alpha, h):
        """WENO7/9_AO_with_df: w (n, order, 4), alpha (n, order) -> left, right (n, 8)"""
        w = np.ascontiguousarray(w, dtype=np.float64)
        alpha = np.ascontiguousarray(alpha, dtype=np.float64)
        n = w.shape[0]
        lo = np.zeros((n, 8))
        ro = np.zeros((n, 8))
        rc = self.fns["weno_normal"](ctypes.byref(p), n, order, dptr(w), dptr(alpha), h, dptr(lo), dptr(ro))
        assert rc == 0
        return lo, ro

    def weno_tangent(self, p, order, lineL, lineR, a1, a2, h):
        """weno_7th/9th tangential: lines (n, order, 8) -> gpL, gpR (n, NG, 12)"""
        lineL = np.ascontiguousarray(lineL, dtype=np.float64)
        lineR = np.ascontiguousarray(lineR, dtype=np.float64)
        a1 = np.ascontiguousarray(a1, dtype=np.float64)
        a2 = np.ascontiguousarray(a2, dtype=np.float64)
        n = lineL.shape[0]
        ng = 3 if order == 7 else 4
        gl = np.zeros((n, ng, 12))
        gr = np.zeros((n, ng, 12))
        rc = self.fns["weno_tangent"](ctypes.byref(p), n, order, dptr(lineL), dptr(lineR), dptr(a1), dptr(a2), h,
                                      dptr(gl), dptr(gr))
        assert rc == 0
        return gl, gr


_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        _oracle = CpuLib(ORACLE_SO, "orc2d_")
    return _oracle


def have_ref():
    return os.path.exists(REF_SO)


def reference():
    global _ref
    if _ref is None:
        _ref = CpuLib(REF_SO, "ref2d_")
    return _ref


# ------------------------------------------------------------------------ 1D
def _bind1(lib, prefix):
    from gks_amd.params import Params1D
    pp = ctypes.POINTER(Params1D)
    sig = {
        "create": (_P, [pp]),
        "destroy": (None, [_P]),
        "set_prim": (ctypes.c_int, [_P, _D]),
        "get": (ctypes.c_int, [_P, _D, _D, _D, _I]),
        "run": (ctypes.c_int, [_P, _I, ctypes.c_double, ctypes.c_int32, _D]),
        "get_face_fluxes": (ctypes.c_int, [_P, ctypes.c_int32, _D]),
        "flux_points": (ctypes.c_int, [pp, ctypes.c_int64, _D, _D, ctypes.c_double, _D, _D, _D]),
    }
    fns = {}
    for name, (res, args) in sig.items():
        f = getattr(lib, prefix + name)
        f.restype = res
        f.argtypes = args
        fns[name] = f
    return fns


class CpuSolver1D:
    """One 1D block (SodTubeProblem-style loop) in the oracle or the reference."""

    def __init__(self, fns, p):
        self.f = fns
        self.p = p
        self.h = fns["create"](ctypes.byref(p))
        if not self.h:
            raise ValueError("unsupported parameters")
        self.nx = p.nx

    def close(self):
        if self.h:
            self.f["destroy"](self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_prim(self, prim):
        prim = np.ascontiguousarray(prim, dtype=np.float64)
        assert prim.shape == (3, self.nx)
        return self.f["set_prim"](self.h, dptr(prim))

    def get(self):
        u = np.zeros((3, self.nx))
        a = np.zeros(self.nx)
        t = ctypes.c_double()
        s = ctypes.c_int32()
        self.f["get"](self.h, dptr(u), dptr(a), ctypes.byref(t), ctypes.byref(s))
        return dict(convar=u, alpha=a, t=t.value, step=s.value)

    def run(self, bc, tstop, max_steps):
        bca = (ctypes.c_int32 * 2)(*bc)
        hist = np.zeros(max(1, min(max_steps, 10_000_000)))
        rc = self.f["run"](self.h, bca, tstop, max_steps, dptr(hist))
        return rc, hist

    def face_fluxes(self, stage):
        out = np.zeros((self.nx + 1, 6))
        self.f["get_face_fluxes"](self.h, stage, dptr(out))
        return out


class CpuLib1D:
    def __init__(self, path, prefix):
        self.lib = ctypes.CDLL(path)
        self.fns = _bind1(self.lib, prefix)

    def solver(self, p):
        return CpuSolver1D(self.fns, p)

    def flux_points(self, p, left, right, dt):
        left = np.ascontiguousarray(left, dtype=np.float64)
        right = np.ascontiguousarray(right, dtype=np.float64)
        n = left.shape[0]
        c = np.zeros((n, 6))
        f = np.zeros((n, 3))
        d = np.zeros((n, 3))
        self.fns["flux_points"](ctypes.byref(p), n, dptr(left), dptr(right), dt, dptr(c), dptr(f), dptr(d))
        return c, f, d


_oracle1 = None
_ref1 = None


def oracle1d():
    global _oracle1
    if _oracle1 is None:
        _oracle1 = CpuLib1D(ORACLE_SO, "orc1d_")
    return _oracle1


def reference1d():
    global _ref1
    if _ref1 is None:
        _ref1 = CpuLib1D(REF_SO, "ref1d_")
    return _ref1


# ------------------------------------------------------------------------ 3D
class Oracle3:
    """The 3D restatement (oracle/solver3d.hpp) through orc3d_*."""

    def __init__(self):
        from gks_amd.params import Params3D
        self.lib = ctypes.CDLL(ORACLE_SO)
        pp = ctypes.POINTER(Params3D)
        sig = {
            "orc3d_params_layout": (ctypes.c_int64, [ctypes.POINTER(ctypes.c_int64)]),
            "orc3d_create": (_P, [pp]),
            "orc3d_destroy": (None, [_P]),
            "orc3d_set_convar": (ctypes.c_int, [_P, _D]),
            "orc3d_set_cfl_primvar": (ctypes.c_int, [_P, _D]),
            "orc3d_get": (ctypes.c_int, [_P, _D, _D, _I]),
            "orc3d_run": (ctypes.c_int, [_P, _I, ctypes.c_double, ctypes.c_int32, _D]),
            "orc3d_flux_points": (ctypes.c_int, [pp, ctypes.c_int64, _D, _D, ctypes.c_double, _D, _D, _D]),
            "orc3d_weno_face": (ctypes.c_int, [pp, ctypes.c_int64, _D, ctypes.c_double, _D, _D]),
            "orc3d_weno_tangential": (ctypes.c_int, [pp, ctypes.c_int64, ctypes.c_int32, _D, ctypes.c_double, _D]),
        }
        for name, (res, args) in sig.items():
            f = getattr(self.lib, name)
            f.restype = res
            f.argtypes = args
        self.f = self.lib

    def flux_points(self, p, left, right, dt):
        left = np.ascontiguousarray(left, dtype=np.float64)
        right = np.ascontiguousarray(right, dtype=np.float64)
        n = left.shape[0]
        c, f, d = np.zeros((n, 20)), np.zeros((n, 5)), np.zeros((n, 5))
        self.f.orc3d_flux_points(ctypes.byref(p), n, dptr(left), dptr(right), dt, dptr(c), dptr(f), dptr(d))
        return c, f, d

    def weno_face(self, p, w, h):
        w = np.ascontiguousarray(w, dtype=np.float64)
        n = w.shape[0]
        lo, ro = np.zeros((n, 10)), np.zeros((n, 10))
        self.f.orc3d_weno_face(ctypes.byref(p), n, dptr(w), h, dptr(lo), dptr(ro))
        return lo, ro

    def weno_tangential(self, p, st, h):
        st = np.ascontiguousarray(st, dtype=np.float64)
        n, nc = st.shape[0], st.shape[2]
        out = np.zeros((n, 2, nc + 5))
        self.f.orc3d_weno_tangential(ctypes.byref(p), n, nc, dptr(st), h, dptr(out))
        return out

    def run(self, p, convar, prim, bc, tstop, max_steps):
        """(convar, t, steps, rc, dt history) after the driver loop."""
        h = self.f.orc3d_create(ctypes.byref(p))
        assert h
        try:
            u = np.ascontiguousarray(convar, dtype=np.float64)
            pr = np.ascontiguousarray(prim, dtype=np.float64)
            self.f.orc3d_set_convar(h, dptr(u))
            self.f.orc3d_set_cfl_primvar(h, dptr(pr))
            hist = np.zeros(max(1, max_steps))
            rc = self.f.orc3d_run(h, (ctypes.c_int32 * 6)(*bc), tstop, max_steps, dptr(hist))
            out = np.empty_like(u)
            t = ctypes.c_double()
            s = ctypes.c_int32()
            self.f.orc3d_get(h, dptr(out), ctypes.byref(t), ctypes.byref(s))
            return out, t.value, s.value, rc, hist[:s.value]
        finally:
            self.f.orc3d_destroy(h)


_oracle3 = None


def oracle3():
    global _oracle3
    if _oracle3 is None:
        _oracle3 = Oracle3()
    return _oracle3


class Ref3:
    """The compiled reference's 3D leaves and driver loop through ref3d_*
    (oracle/ref_harness.cpp: the reference's code with Interface3d::gauss
    allocated and a GKS3D that takes lambda from the conservative state)."""

    def __init__(self):
        from gks_amd.params import Params3D
        self.lib = ctypes.CDLL(REF_SO)
        pp = ctypes.POINTER(Params3D)
        sig = {
            "ref3d_create": (_P, [pp]),
            "ref3d_destroy": (None, [_P]),
            "ref3d_set_state": (ctypes.c_int, [_P, _D, _D]),
            "ref3d_get": (ctypes.c_int, [_P, _D, _D, _I]),
            "ref3d_run": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_int32, _D]),
            "ref3d_weno_face": (ctypes.c_int, [pp, ctypes.c_int64, _D, ctypes.c_double, _D, _D]),
            "ref3d_weno_tangential": (ctypes.c_int, [pp, ctypes.c_int64, ctypes.c_int32, _D, ctypes.c_double, _D]),
            "ref3d_gks3d": (ctypes.c_int, [pp, ctypes.c_int64, _D, _D, ctypes.c_double, _D, _D]),
        }
        for name, (res, args) in sig.items():
            f = getattr(self.lib, name)
            f.restype = res
            f.argtypes = args
        self.f = self.lib

    def gks3d(self, p, left, right, dt):
        left = np.ascontiguousarray(left, dtype=np.float64)
        right = np.ascontiguousarray(right, dtype=np.float64)
        n = left.shape[0]
        f, d = np.zeros((n, 5)), np.zeros((n, 5))
        self.f.ref3d_gks3d(ctypes.byref(p), n, dptr(left), dptr(right), dt, dptr(f), dptr(d))
        return f, d

    def weno_face(self, p, w, h):
        w = np.ascontiguousarray(w, dtype=np.float64)
        n = w.shape[0]
        lo, ro = np.zeros((n, 10)), np.zeros((n, 10))
        self.f.ref3d_weno_face(ctypes.byref(p), n, dptr(w), h, dptr(lo), dptr(ro))
        return lo, ro

    def weno_tangential(self, p, st, h):
        st = np.ascontiguousarray(st, dtype=np.float64)
        n, nc = st.shape[0], st.shape[2]
        out = np.zeros((n, 2, nc + 5))
        assert self.f.ref3d_weno_tangential(ctypes.byref(p), n, nc, dptr(st), h, dptr(out)) == 0
        return out

    def run(self, p, convar, prim, tstop, max_steps):
        """(convar, t, steps, dt history) after the driver loop (free boundaries)."""
        h = self.f.ref3d_create(ctypes.byref(p))
        assert h
        try:
            u = np.ascontiguousarray(convar, dtype=np.float64)
            pr = np.ascontiguousarray(prim, dtype=np.float64)
            self.f.ref3d_set_state(h, dptr(u), dptr(pr))
            hist = np.zeros(max(1, max_steps))
            self.f.ref3d_run(h, tstop, max_steps, dptr(hist))
            out = np.empty_like(u)
            t = ctypes.c_double()
            s = ctypes.c_int32()
            self.f.ref3d_get(h, dptr(out), ctypes.byref(t), ctypes.byref(s))
            return out, t.value, s.value, hist[:s.value]
        finally:
            self.f.ref3d_destroy(h)


_ref3 = None


def reference3():
    global _ref3
    if _ref3 is None:
        _ref3 = Ref3()
    return _ref3
