"""ctypes mirror of ``gks2d_params`` (include/gks2d.h) and scheme presets.

The presets restate the settings the reference's problem drivers assign to
its globals (e.g. ``PlanarShock``, reference ``src/2Dproblem.cpp:8-53``).
"""
import ctypes
import math

BC_FREE, BC_PERIODIC, BC_REFLECT = 0, 1, 2
SIDE_LEFT, SIDE_RIGHT, SIDE_DOWN, SIDE_UP = 0, 1, 2, 3
VAR_CONSERVATIVE, VAR_CHARACTERISTIC = 0, 1
TAU_EULER, TAU_NS = 0, 1

GKS_OK = 0
STATUS_NAMES = {
    0: "GKS_OK", 1: "GKS_ERR_ARG", 2: "GKS_ERR_UNSUPPORTED", 3: "GKS_ERR_HIP",
    4: "GKS_ERR_BLOWUP", 5: "GKS_ERR_NOMEM", 6: "GKS_ERR_STATE",
}


class Params(ctypes.Structure):
    _fields_ = [
        ("nodex", ctypes.c_int32), ("nodey", ctypes.c_int32), ("ghost", ctypes.c_int32),
        ("gausspoint", ctypes.c_int32), ("weno_order", ctypes.c_int32),
        ("recon_variable", ctypes.c_int32), ("use_df", ctypes.c_int32),
        ("tau_type", ctypes.c_int32), ("K", ctypes.c_int32), ("prandtl_fix", ctypes.c_int32),
        ("gamma", ctypes.c_double), ("mu", ctypes.c_double), ("nu", ctypes.c_double),
        ("pr", ctypes.c_double), ("c1_euler", ctypes.c_double), ("c2_euler", ctypes.c_double),
        ("left", ctypes.c_double), ("right", ctypes.c_double), ("down", ctypes.c_double),
        ("up", ctypes.c_double), ("cfl", ctypes.c_double),
        ("fast_math", ctypes.c_int32), ("df_frozen", ctypes.c_int32),
    ]

    @property
    def nx(self):
        return self.nodex + 2 * self.ghost

    @property
    def ny(self):
        return self.nodey + 2 * self.ghost

    def copy(self, **kw):
        p = Params()
        ctypes.pointer(p)[0] = self
        for k, v in kw.items():
            setattr(p, k, v)
        return p


GAUSS_FOR_ORDER = {5: 2, 7: 3, 9: 4}


def planar_shock_params(n, weno_order=5, cfl=0.5, **kw):
    """Settings of PlanarShock (reference src/2Dproblem.cpp:8-53) on an n x n
    grid of [0,1]^2: K=3, gamma=1.4, characteristic variables, DF factor on,
    GKS2D gks2nd_2d with Euler tau c1=0.05 c2=1, S2O4, 5 ghost layers."""
    p = Params(
        nodex=n, nodey=n, ghost=5, gausspoint=GAUSS_FOR_ORDER[weno_order],
        weno_order=weno_order, recon_variable=VAR_CHARACTERISTIC, use_df=1,
        tau_type=TAU_EULER, K=3, prandtl_fix=0, gamma=1.4, mu=-1.0, nu=-1.0, pr=1.0,
        c1_euler=0.05, c2_euler=1.0, left=0.0, right=1.0, down=0.0, up=1.0, cfl=cfl,
        fast_math=0, df_frozen=0,
    )
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Params1D(ctypes.Structure):
    """ctypes mirror of ``gks1d_params`` (include/gks1d.h)."""
    _fields_ = [
        ("nodex", ctypes.c_int32), ("ghost", ctypes.c_int32), ("recon_variable", ctypes.c_int32),
        ("use_df", ctypes.c_int32), ("tau_type", ctypes.c_int32), ("K", ctypes.c_int32),
        ("fast_math", ctypes.c_int32), ("df_from_start", ctypes.c_int32),
        ("gamma", ctypes.c_double), ("mu", ctypes.c_double), ("nu", ctypes.c_double),
        ("pr", ctypes.c_double), ("c1_euler", ctypes.c_double), ("c2_euler", ctypes.c_double),
        ("left", ctypes.c_double), ("right", ctypes.c_double), ("cfl", ctypes.c_double),
    ]

    @property
    def nx(self):
        return self.nodex + 2 * self.ghost

    @property
    def dx(self):
        return (self.right - self.left) / self.nodex

    def copy(self, **kw):
        p = Params1D()
        ctypes.pointer(p)[0] = self
        for k, v in kw.items():
            setattr(p, k, v)
        return p


def sod_params(n=400, **kw):
    """Settings of SodTubeProblem (reference src/1Dproblem.cpp:3-43): K=4,
    gamma=1.4, Euler tau c1=0.05 c2=1, WENO5-AO (wenoz) in characteristic
    variables, Center_collision, GKS gks2nd, S2O4, CFL 0.5 on [0, 1].
    The reference leaves is_using_df_factor at its default (false,
    reconstruction.cpp:10); ``use_df=1`` gives the DF variant of the loop."""
    p = Params1D(nodex=n, ghost=5, recon_variable=VAR_CHARACTERISTIC, use_df=0, tau_type=TAU_EULER, K=4,
                 fast_math=0, df_from_start=0, gamma=1.4, mu=-1.0, nu=-1.0, pr=1.0, c1_euler=0.05,
                 c2_euler=1.0, left=0.0, right=1.0, cfl=0.5)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def blastwave_params(n=400, **kw):
    """Blastwave (reference src/1Dproblem.cpp:176-321): as Sod with c1=0.01,
    c2=5 and reflective walls; run to t=0.038."""
    kw = dict(dict(c1_euler=0.01, c2_euler=5.0), **kw)
    return sod_params(n, **kw)


def shu_osher_params(n=200, **kw):
    """ShuOsher (reference src/1Dproblem.cpp:347-491): as Sod on [0, 10],
    free boundaries; run to t=1.8."""
    kw = dict(dict(left=0.0, right=10.0), **kw)
    return sod_params(n, **kw)


def sinwave1d_params(n, cfl=0.1, **kw):
    """accuracy_sinwave_1d (reference src/1Dproblem.cpp:550-690): periodic
    density wave on [0, 2], WENO5_AO_with_DF in conservative variables from
    the first step (no alpha update), c1=c2=0, run to t=2."""
    kw = dict(dict(recon_variable=VAR_CONSERVATIVE, df_from_start=1, c1_euler=0.0, c2_euler=0.0,
                   left=0.0, right=2.0, cfl=cfl), **kw)
    return sod_params(n, **kw)


def vortex_params(n, weno_order=7, cfl=0.5, **kw):
    """Isentropic vortex advection on the periodic square [0, 10]^2 (a
    workload of this build; the reference has no vortex driver): K=3,
    gamma=1.4, the smooth-flow collision time of the reference's accuracy
    drivers (Euler tau with c1=c2=0, sinwave_2d, src/2Dproblem.cpp:1358-1361),
    WENO7-AO (3 Gauss points) with the DF factor in characteristic
    variables, S2O4, CFL 0.5."""
    p = planar_shock_params(n, weno_order=weno_order, cfl=cfl, c1_euler=0.0, c2_euler=0.0,
                            left=0.0, right=10.0, down=0.0, up=10.0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def sinwave2d_params(n, cfl=0.05, **kw):
    """sinwave_2d (reference src/2Dproblem.cpp:1348-1527): the density wave
    rho = 1 + 0.2 sin(pi (x + y)), u = v = p = 1, on the periodic [0, 2]^2
    at n x n, K=3, gamma=1.4, Euler tau with c1=c2=0 (tau = 0), Mu=0,
    Pr=0.73, WENO9-AO with the DF factor (WENO9_AO_with_df_normal /
    _tangent) in conservative variables, 4 Gauss points (Lobatto), GKS2D
    gks2nd_2d, S2O4, CFL 0.05; Update_alpha is commented out there, so the
    DF sensor stays zero (df_frozen)."""
    p = planar_shock_params(n, weno_order=9, cfl=cfl, recon_variable=VAR_CONSERVATIVE, c1_euler=0.0,
                            c2_euler=0.0, mu=0.0, pr=0.73, left=0.0, right=2.0, down=0.0, up=2.0, df_frozen=1)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Params3D(ctypes.Structure):
    """ctypes mirror of ``gks3d_params`` (include/gks3d.h)."""
    _fields_ = [
        ("nodex", ctypes.c_int32), ("nodey", ctypes.c_int32), ("nodez", ctypes.c_int32),
        ("ghost", ctypes.c_int32), ("recon_variable", ctypes.c_int32), ("tau_type", ctypes.c_int32),
        ("K", ctypes.c_int32), ("fast_math", ctypes.c_int32),
        ("gamma", ctypes.c_double), ("mu", ctypes.c_double), ("nu", ctypes.c_double),
        ("c1_euler", ctypes.c_double), ("c2_euler", ctypes.c_double),
        ("left", ctypes.c_double), ("right", ctypes.c_double), ("down", ctypes.c_double),
        ("up", ctypes.c_double), ("back", ctypes.c_double), ("front", ctypes.c_double),
        ("cfl", ctypes.c_double),
    ]

    @property
    def nx(self):
        return self.nodex + 2 * self.ghost

    @property
    def ny(self):
        return self.nodey + 2 * self.ghost

    @property
    def nz(self):
        return self.nodez + 2 * self.ghost

    @property
    def dx(self):
        return (self.right - self.left) / self.nodex

    @property
    def dy(self):
        return (self.up - self.down) / self.nodey

    @property
    def dz(self):
        return (self.front - self.back) / self.nodez

    def copy(self, **kw):
        p = Params3D()
        ctypes.pointer(p)[0] = self
        for k, v in kw.items():
            setattr(p, k, v)
        return p


def cubic_params(n, **kw):
    """Settings of CubicTube (reference src/3Dproblem.cpp:3-108): K=2,
    gamma=1.4, Euler tau c1=0.05 c2=1, WENO5-AO (wenoz) in characteristic
    variables, 3 ghost layers, 2x2 Gauss points, GKS3D gks2nd, S2O4, CFL 0.5
    on [0,1]^3."""
    p = Params3D(nodex=n, nodey=n, nodez=n, ghost=3, recon_variable=VAR_CHARACTERISTIC, tau_type=TAU_EULER,
                 K=2, fast_math=0, gamma=1.4, mu=-1.0, nu=-1.0, c1_euler=0.05, c2_euler=1.0,
                 left=0.0, right=1.0, down=0.0, up=1.0, back=0.0, front=1.0, cfl=0.5)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def tgv_params(n, re=1600.0, **kw):
    """Viscous Taylor-Green vortex on the periodic cube [0, 2 pi]^3 (a
    workload of this build; the reference has no 3D viscous driver): rho0=1,
    U0=1, L=1, gamma=1.4 (K=2), Navier-Stokes collision time with
    mu = rho0 U0 L / Re (Get_Tau_NS, basic_function.cpp:638-661; the
    numerical part c2 |jump| dt vanishes on smooth data, c1=0), WENO5-AO in
    characteristic variables, 3 ghost layers, CFL 0.5."""
    two_pi = 2.0 * math.pi
    p = cubic_params(n, tau_type=TAU_NS, mu=1.0 / re, c1_euler=0.0, c2_euler=1.0, left=0.0, right=two_pi,
                     down=0.0, up=two_pi, back=0.0, front=two_pi)
    for k, v in kw.items():
        setattr(p, k, v)
    return p
