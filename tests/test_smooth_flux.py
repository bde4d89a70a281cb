"""The smooth-flow GKS flux of the learned path (oracle/smooth_flux.py)
against the reference's GKS2D (the oracle's flux_points, pinned to the
compiled reference elsewhere): with equal left and right states and slopes
the two are the same flux, so they must agree to rounding; with unequal
sides they must not (the test would pass trivially if the smooth flux
ignored one side)."""
import os
import sys

import numpy as np
import pytest

from gks_amd.params import planar_shock_params
from oracle_lib import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import smooth_flux as sf  # noqa: E402


def states(n, seed):
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.5, 2.0, n)
    u = rng.uniform(-1, 1, n)
    v = rng.uniform(-1, 1, n)
    p = rng.uniform(0.5, 2.0, n)
    w = np.stack([rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v)], 1)
    s = np.concatenate([w, 0.3 * w * rng.uniform(-1, 1, (n, 4)), 0.3 * w * rng.uniform(-1, 1, (n, 4))], 1)
    return s


GASES = {
    "euler": dict(),
    "ns": dict(tau_type=1, mu=2e-3, prandtl_fix=1, pr=0.72),
}


@pytest.mark.parametrize("gas", sorted(GASES))
def test_smooth_flux_is_gks2d_when_the_sides_agree(gas):
    p = planar_shock_params(16).copy(**GASES[gas])
    L = states(400, 1)
    dt = 1e-3
    _, f, d = oracle().flux_points(p, L, L, dt)
    fs, ds = sf.smooth_flux(L, L, dt, K=p.K, tau_type=p.tau_type, mu=p.mu, nu=p.nu, pr=p.pr,
                            prandtl_fix=p.prandtl_fix)
    scale_f = np.abs(f).max(axis=1, keepdims=True)
    scale_d = np.abs(d).max(axis=1, keepdims=True)
    assert np.max(np.abs(fs - f) / scale_f) < 1e-13
    assert np.max(np.abs(ds - d) / scale_d) < 1e-9


def test_smooth_flux_differs_from_gks2d_across_a_jump():
    p = planar_shock_params(16)
    L = states(200, 2)
    R = states(200, 3)
    _, f, _ = oracle().flux_points(p, L, R, 1e-3)
    fs, _ = sf.smooth_flux(L, R, 1e-3, K=p.K)
    assert np.median(np.abs(fs - f) / np.abs(f).max(axis=1, keepdims=True)) > 1e-3
