"""ctypes mirror of ``gks2d_params`` (include/gks2d.h) and scheme presets.

The presets restate the settings the reference's problem drivers assign to
its globals (e.g. ``PlanarShock``, reference ``src/2Dproblem.cpp:8-53``).
"""
import ctypes

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
        ("fast_math", ctypes.c_int32), ("reserved", ctypes.c_int32),
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
        fast_math=0, reserved=0,
    )
    for k, v in kw.items():
        setattr(p, k, v)
    return p
