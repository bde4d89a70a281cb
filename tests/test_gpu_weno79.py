"""Parity of the HIP WENO7-AO / WENO9-AO paths (3 and 4 Gauss points per
face, lib/libgks2d.so through the C ABI) with the oracle and the reference's
golden runs.  Same tolerances as tests/test_gpu_parity.py: per interface and
Gauss point the (f, derf) 8-vector to 1e-10 relative (with the Euler-flux
floor), states at final time to 1e-9 in L1."""
import os

import numpy as np
import pytest

import gks_amd
from gks_amd.params import BC_FREE, BC_PERIODIC, BC_REFLECT, planar_shock_params
from gks_amd.problems import riemann_prim, riemann_states
from oracle_lib import oracle
from test_gpu_parity import FLUX_RTOL, L1_TOL, face_floor, l1, one_stage_both, say, vec_rel
from test_oracle_79 import RUNS79

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
G = 5
VARIANTS = [0, 1]


@pytest.mark.parametrize("fast", VARIANTS)
@pytest.mark.parametrize("order,recon", [(7, 1), (7, 0), (9, 1), (9, 0)])
def test_stage_face_fluxes_match_oracle_79(order, recon, fast):
    n = 24
    bc = [BC_REFLECT, BC_FREE, BC_PERIODIC, BC_PERIODIC]
    p = planar_shock_params(n, weno_order=order, recon_variable=recon, fast_math=fast)
    o = oracle().solver(p)
    o.set_prim(riemann_prim(p, riemann_states(4)))
    assert o.run(bc, 0.25, 8)[0] == 0
    st = o.get()
    say(f"[weno{order}] stage fluxes recon={recon} fast={fast}")
    for stage in (0, 1):
        s, o2 = one_stage_both(p, st["convar"], st["alpha_x"], st["alpha_y"], 1.1e-3, bc, stage)
        xf, yf = s.face_fluxes()
        oxf, oyf = o2.face_fluxes(stage)
        xs = (slice(G, n + G + 1), slice(G, n + G))
        ys = (slice(G, n + G), slice(G, n + G + 1))
        xg, yg = o2.gauss()
        fx = face_floor(xg[xs][..., 0, :], xg[xs][..., 1, :], 1.1e-3).reshape(-1)
        fy = face_floor(yg[ys][..., 0, :], yg[ys][..., 1, :], 1.1e-3).reshape(-1)
        ex = vec_rel(xf[xs].reshape(-1, 8), oxf[xs].reshape(-1, 8), fx)
        ey = vec_rel(yf[ys].reshape(-1, 8), oyf[ys].reshape(-1, 8), fy)
        say(f"   stage {stage} max rel (f, derf) x {ex.max():.2e} y {ey.max():.2e}")
        assert max(ex.max(), ey.max()) <= FLUX_RTOL
        u = s.get_convar()[:, G:-G, G:-G]
        uo = o2.get()["convar"][:, G:-G, G:-G]
        assert np.all(l1(u, uo, p) <= 1e-13)
        s.close()


@pytest.mark.parametrize("fast", VARIANTS)
@pytest.mark.parametrize("name", list(RUNS79))
def test_runs79_match_reference_golden(name, fast):
    g = np.load(os.path.join(GOLD, "runs79.npz"))
    order, recon, n, seed, bc, steps = RUNS79[name]
    p = planar_shock_params(n, weno_order=order, recon_variable=recon, fast_math=fast)
    say(f"[weno{order}] golden run {name} fast={fast}")
    s = gks_amd.Solver2D(p)
    s.init_prim(riemann_prim(p, riemann_states(seed)))
    k = s.run(bc, 0.25, steps)
    assert k == steps
    e = l1(s.get_convar()[:, G:-G, G:-G], g[name + "_convar"], p)
    say("   L1", e)
    assert np.all(e <= L1_TOL)
    assert abs(s.t - float(g[name + "_t"])) <= 1e-14
    s.close()


@pytest.mark.parametrize("order", [7, 9])
def test_weno79_full_run_matches_oracle(order):
    n = 40
    p = planar_shock_params(n, weno_order=order)
    prim = riemann_prim(p, riemann_states(6))
    say(f"[weno{order}] full run n={n} to t=0.25")
    o = oracle().solver(p)
    o.set_prim(prim)
    rc, hist = o.run([BC_FREE] * 4, 0.25, 100000)
    assert rc == 0
    st = o.get()
    s = gks_amd.Solver2D(p)
    s.init_prim(prim)
    k = s.run([BC_FREE] * 4, 0.25)
    assert k == st["step"]
    e = l1(s.get_convar()[:, G:-G, G:-G], st["convar"][:, G:-G, G:-G], p)
    say("   steps", k, "L1", e)
    assert np.all(e <= L1_TOL)
    s.close()


@pytest.mark.parametrize("order", [7, 9])
def test_fused_and_separate_tangential_launches_agree(order, monkeypatch):
    """The default WENO7/9 sweep reconstructs the tangential Gauss-point
    states inside the flux launch (LDS); GKS_TANFLUX=0 keeps them in HBM
    between two launches.  Both evaluate the same per-point code, so after
    20 steps of a four-quadrant Riemann problem the states agree to rounding."""
    n = 48
    p = planar_shock_params(n, weno_order=order)
    prim = riemann_prim(p, riemann_states(3))
    out = []
    for flag in ("1", "0"):
        monkeypatch.setenv("GKS_TANFLUX", flag)
        say(f"[weno{order}] GKS_TANFLUX={flag} n={n}, 20 steps")
        s = gks_amd.Solver2D(p)
        s.init_prim(prim)
        assert s.run([BC_FREE] * 4, 1.0, 20) == 20
        out.append((s.get_convar(), s.t))
        s.close()
    assert out[0][1] == out[1][1]
    e = l1(out[0][0][:, G:-G, G:-G], out[1][0][:, G:-G, G:-G], p)
    say("   L1 fused vs separate", e)
    assert np.all(e <= 1e-13)
