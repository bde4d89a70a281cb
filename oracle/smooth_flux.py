"""Test infrastructure: numpy restatement of the smooth-flow GKS flux that
the learned path uses at Gauss points whose two sides took the linear
weights (include/gks_learn.h, gks2d_flux_points_smooth).  Not part of the
product; only tests import it.

The formula has no counterpart in the reference.  It is pinned to the
reference's GKS2D (flux_function.cpp:345-698, through the oracle's
flux_points) by the identity it is built on: with equal left and right
states and slopes, GKS2D's (dt, dt/2) combination equals this flux (the
numerical collision time drops out), see tests/test_smooth_flux.py.

Layout of a Gauss-point state (12 values): conservative variables, their
normal derivatives, their tangential derivatives, in the face frame.
"""
import numpy as np


def moments(u, v, lam, K):
    """Whole-space moments <u^i>, <v^j> (i, j <= 6) of a Maxwellian and the
    internal-energy moments <xi^2>, <xi^4> (basic_function.cpp:110-187,
    the uwhole / vwhole recursions)."""
    n = u.shape[0]
    uw = np.zeros((n, 7))
    vw = np.zeros((n, 7))
    uw[:, 0] = 1.0
    uw[:, 1] = u
    vw[:, 0] = 1.0
    vw[:, 1] = v
    ol = 1.0 / lam
    for i in range(2, 7):
        uw[:, i] = u * uw[:, i - 1] + 0.5 * (i - 1) * ol * uw[:, i - 2]
        vw[:, i] = v * vw[:, i - 1] + 0.5 * (i - 1) * ol * vw[:, i - 2]
    xi = np.stack([np.ones(n), 0.5 * K * ol, 0.25 * (K * K + 2 * K) * ol * ol], 1)
    return uw, vw, xi


def contract(uw, vw, xi, a, u0, v0):
    """<u^u0 v^v0 psi (a0 + a1 u + a2 v + a3 (u^2+v^2+xi^2)/2)> for
    psi = (1, u, v, (u^2+v^2+xi^2)/2) (basic_function.cpp:603-636)."""
    def m(i, j, k):
        return uw[:, i] * vw[:, j] * xi[:, k]
    u, v = u0, v0
    p0 = a[:, 0] * m(u, v, 0) + a[:, 1] * m(u + 1, v, 0) + a[:, 2] * m(u, v + 1, 0) + \
        a[:, 3] * 0.5 * (m(u + 2, v, 0) + m(u, v + 2, 0) + m(u, v, 1))
    p1 = a[:, 0] * m(u + 1, v, 0) + a[:, 1] * m(u + 2, v, 0) + a[:, 2] * m(u + 1, v + 1, 0) + \
        a[:, 3] * 0.5 * (m(u + 3, v, 0) + m(u + 1, v + 2, 0) + m(u + 1, v, 1))
    p2 = a[:, 0] * m(u, v + 1, 0) + a[:, 1] * m(u + 1, v + 1, 0) + a[:, 2] * m(u, v + 2, 0) + \
        a[:, 3] * 0.5 * (m(u + 2, v + 1, 0) + m(u, v + 3, 0) + m(u, v + 1, 1))
    p3 = 0.5 * (a[:, 0] * (m(u + 2, v, 0) + m(u, v + 2, 0) + m(u, v, 1)) +
                a[:, 1] * (m(u + 3, v, 0) + m(u + 1, v + 2, 0) + m(u + 1, v, 1)) +
                a[:, 2] * (m(u + 2, v + 1, 0) + m(u, v + 3, 0) + m(u, v + 1, 1)) +
                a[:, 3] * 0.5 * (m(u + 4, v, 0) + m(u, v + 4, 0) + m(u, v, 2) + 2 * m(u + 2, v + 2, 0) +
                                 2 * m(u + 2, v, 1) + 2 * m(u, v + 2, 1)))
    return np.stack([p0, p1, p2, p3], 1)


def micro_slope(der, rho, u, v, lam, K):
    """Microscopic slope a from a macroscopic derivative (basic_function.cpp:571-583)."""
    r4 = der[:, 3] / rho - 0.5 * (u * u + v * v + 0.5 * (K + 2) / lam) * der[:, 0] / rho
    r3 = (der[:, 2] - v * der[:, 0]) / rho
    r2 = (der[:, 1] - u * der[:, 0]) / rho
    a3 = (4.0 / (K + 2)) * lam * lam * (2 * r4 - 2 * u * r2 - 2 * v * r3)
    a2 = 2 * lam * r3 - v * a3
    a1 = 2 * lam * r2 - u * a3
    a0 = der[:, 0] / rho - u * a1 - v * a2 - 0.5 * a3 * (u * u + v * v + 0.5 * (K + 2) / lam)
    return np.stack([a0, a1, a2, a3], 1)


def smooth_flux(left, right, dt, K=3, tau_type=0, mu=-1.0, nu=-1.0, pr=1.0, prandtl_fix=0):
    """(f, derf) of n Gauss points: f = rho0 dt [<u psi g0> - tau (<u psi g0
    (a u + b v)> + <u psi g0 A>)], derf = rho0 dt^2 <u psi g0 A>, with the
    interface state and slopes averaged over the two sides."""
    left = np.asarray(left, dtype=np.float64)
    right = np.asarray(right, dtype=np.float64)
    c = 0.5 * (left + right)
    w, wx, wy = c[:, 0:4], c[:, 4:8], c[:, 8:12]
    rho = w[:, 0]
    u = w[:, 1] / rho
    v = w[:, 2] / rho
    lam = (K + 2.0) * 0.25 * (rho / (w[:, 3] - 0.5 * rho * (u * u + v * v)))
    uw, vw, xi = moments(u, v, lam, K)
    unit = np.tile([1.0, 0.0, 0.0, 0.0], (len(rho), 1))
    g0u = contract(uw, vw, xi, unit, 1, 0)
    ax = micro_slope(wx, rho, u, v, lam, K)
    ay = micro_slope(wy, rho, u, v, lam, K)
    axu = contract(uw, vw, xi, ax, 1, 0)
    axuu = contract(uw, vw, xi, ax, 2, 0)
    ayv = contract(uw, vw, xi, ay, 0, 1)
    ayuv = contract(uw, vw, xi, ay, 1, 1)
    A = micro_slope(-rho[:, None] * (axu + ayv), rho, u, v, lam, K)
    Au = contract(uw, vw, xi, A, 1, 0)
    if tau_type == 0:
        tau = np.zeros_like(rho)
    elif mu > 0:
        tau = 2.0 * mu * lam / rho
    elif nu > 0:
        tau = 2 * nu * lam
    else:
        tau = np.zeros_like(rho)
    f = rho[:, None] * dt * (g0u - tau[:, None] * ((axuu + ayuv) + Au))
    d = rho[:, None] * dt * dt * Au
    if prandtl_fix:
        qs = axuu[:, 3] + Au[:, 3] - u * (axuu[:, 1] + Au[:, 1]) - v * (axuu[:, 2] + Au[:, 2])
        f[:, 3] += (1.0 / pr - 1.0) * (-rho * tau * dt * qs)
    return f, d
