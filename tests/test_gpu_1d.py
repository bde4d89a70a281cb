"""Parity of the 1D HIP path (lib/libgks1d.so, one persistent workgroup,
called through the C ABI) with the compiled reference's golden vectors and
with the oracle.

Tolerances, as for the 2D path: per interface, the 6-vector (f, derf)
relative to max(its max-norm, 1e-4 * dt * Euler-flux scale) <= 1e-10; the
state at final time L1 = sum |dU| dx per component <= 1e-9 for O(1) states
(times the component's own L1 norm where that exceeds 1: the blast wave
carries energies of 2500); the same number
of steps and dt history to 1e-9 of the step size (the CFL scan is the
reference's own sequential one, so dt differs only through the state it
reads, and the last step's tstop - t through the accumulated t).
"""
import os
import sys

import numpy as np
import pytest

from gks_amd.params import BC_FREE, BC_PERIODIC, sod_params
from gks_amd.problems import sod_prim
from gks_amd.solver1d import Solver1D, flux_points_1d
from oracle_lib import oracle1d
from test_oracle_1d import RUNS, case_ic

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
G = 5
FLUX_RTOL = 1e-10
L1_TOL = 1e-9
DT_RTOL = 1e-9
VARIANTS = [0, 1]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def euler_scale_1d(S, gamma=1.4):
    r, m, E = S[..., 0], S[..., 1], S[..., 2]
    u = m / r
    p = (gamma - 1) * (E - 0.5 * m * m / r)
    return np.max(np.abs(np.stack([m, m * u + p, u * (E + p)], -1)), axis=-1)


def joint_rel(f, d, fo, do, floor):
    a = np.concatenate([f, d], axis=-1)
    b = np.concatenate([fo, do], axis=-1)
    den = np.maximum(np.max(np.abs(b), axis=-1), floor)
    return np.max(np.abs(a - b), axis=-1) / np.maximum(den, 1e-300)


def gpu_run(p, ic, bc, tstop, max_steps=100000):
    s = Solver1D(p)
    s.init_prim(ic)
    n, dts = s.run(bc, tstop, max_steps, record_dt=True)
    return s, n, dts


@pytest.mark.parametrize("fast", VARIANTS)
@pytest.mark.parametrize("case", ["euler", "c0", "ns"])
def test_flux_points_1d_match_golden(case, fast):
    g = np.load(os.path.join(GOLD, "gks1d_flux_points.npz"))
    kw = {"euler": {}, "c0": dict(c1_euler=0.0, c2_euler=0.0),
          "ns": dict(tau_type=1, mu=1e-3, c1_euler=0.0, c2_euler=1.0)}[case]
    p = sod_params(16, fast_math=fast, **kw)
    dt = float(g[f"{case}_dt"])
    say(f"[1d flux points] case={case} fast={fast} n={g[case + '_left'].shape[0]}")
    c, f, d = flux_points_1d(p, g[f"{case}_left"], g[f"{case}_right"], dt)
    co = g[f"{case}_center"]
    crel = np.max(np.abs(c - co), axis=-1) / np.maximum(np.max(np.abs(co), axis=-1), 1e-300)
    assert crel.max() < 1e-12, crel.max()
    floor = 1e-4 * dt * np.maximum(euler_scale_1d(g[f"{case}_left"]), euler_scale_1d(g[f"{case}_right"]))
    rel = joint_rel(f, d, g[f"{case}_f"], g[f"{case}_derf"], floor)
    say(f"   max rel {rel.max():.3e}")
    assert rel.max() <= FLUX_RTOL


@pytest.mark.parametrize("fast", VARIANTS)
@pytest.mark.parametrize("name", list(RUNS))
def test_runs_1d_match_golden(name, fast):
    g = np.load(os.path.join(GOLD, "runs1d.npz"))
    c = RUNS[name]
    p = sod_params(c["n"], fast_math=fast, **c["kw"])
    say(f"[1d run] {name} fast={fast} n={c['n']} tstop={c['tstop']}")
    s, n, dts = gpu_run(p, case_ic(name, p), c["bc"], c["tstop"])
    steps = int(g[f"{name}_step"])
    assert n == steps and s.step_count == steps
    assert abs(s.t - float(g[f"{name}_t"])) <= 1e-13 * max(1.0, abs(s.t))
    dref = g[f"{name}_dt"]
    # the last dt is tstop - t + 1e-15: relative to the step size, not to itself
    ddt = float(np.max(np.abs(dts - dref)) / np.max(dref))
    u = s.get_convar()[:, G:-G]
    l1 = np.sum(np.abs(u - g[f"{name}_convar"]), axis=1) * p.dx
    scale = np.maximum(1.0, np.sum(np.abs(g[f"{name}_convar"]), axis=1) * p.dx)
    say(f"   steps {n}  L1 {l1}  (state L1 {scale})  dt {ddt:.2e}")
    assert np.all(l1 <= L1_TOL * scale), l1
    assert ddt <= DT_RTOL, ddt
    a = s.get_alpha()[G:-G]
    assert np.sum(np.abs(a - g[f"{name}_alpha"])) * p.dx <= 1e-6 * max(1.0, np.sum(np.abs(g[f"{name}_alpha"])) * p.dx)
    if name == "sod400_df":
        dt_last = float(dref[-1])
        for k in range(2):
            fo = g[f"{name}_faces{k}"]
            fg = s.face_fluxes(k)[G:p.nodex + G + 1]
            # the face states are not stored by the golden run: floor from the
            # cell states on either side
            uc = g[f"{name}_convar"].T
            sc = euler_scale_1d(np.concatenate([uc[:1], uc, uc[-1:]]))
            floor = 1e-4 * dt_last * np.maximum(sc[:-1], sc[1:])
            rel = joint_rel(fg[:, :3], fg[:, 3:], fo[:, :3], fo[:, 3:], floor)
            say(f"   stage {k} faces max rel {rel.max():.3e}")
            assert rel.max() <= FLUX_RTOL
    s.close()


@pytest.mark.parametrize("fast", VARIANTS)
def test_1d_matches_oracle_random_riemann(fast):
    """Random two-state tubes with mixed boundaries, against the oracle run
    on the same inputs (beyond the golden cases)."""
    for seed in range(3):
        rng = np.random.default_rng(100 + seed)
        p = sod_params(96 + 32 * seed, recon_variable=seed % 2, use_df=1, fast_math=fast)
        prim = np.zeros((3, p.nx))
        cut = rng.integers(G + 10, p.nx - G - 10)
        prim[:, :cut] = np.array([rng.uniform(0.1, 2), rng.uniform(-1, 1), rng.uniform(0.05, 2)])[:, None]
        prim[:, cut:] = np.array([rng.uniform(0.1, 2), rng.uniform(-1, 1), rng.uniform(0.05, 2)])[:, None]
        bc = [[BC_FREE, 2], [2, 2], [BC_FREE, BC_FREE]][seed]
        say(f"[1d random] seed={seed} n={p.nodex} recon={p.recon_variable} fast={fast} bc={bc}")
        o = oracle1d().solver(p)
        o.set_prim(prim)
        rc, hist = o.run(bc, 0.15, 200)
        assert rc == 0
        st = o.get()
        s, n, dts = gpu_run(p, prim, bc, 0.15, 200)
        assert n == st["step"]
        ddt = float(np.max(np.abs(dts - hist[:n])) / np.max(hist[:n]))
        l1 = np.sum(np.abs(s.get_convar()[:, G:-G] - st["convar"][:, G:-G]), axis=1) * p.dx
        say(f"   steps {n} L1 {l1}  dt {ddt:.2e}")
        assert np.all(l1 <= L1_TOL)
        assert ddt <= DT_RTOL, ddt
        s.close()


def test_1d_run_in_pieces_equals_one_run_and_is_deterministic():
    p = sod_params(400, use_df=1)
    ic = sod_prim(p)
    say("[1d pieces] sod400_df, 350 steps in one call / in pieces of 37 / again")
    a, n, _ = gpu_run(p, ic, [BC_FREE, BC_FREE], 0.2)
    b = Solver1D(p)
    b.init_prim(ic)
    total = 0
    while b.t < 0.2:
        total += b.run([BC_FREE, BC_FREE], 0.2, 37)
    assert total == n
    assert b.t == a.t
    assert np.array_equal(a.get_convar(), b.get_convar())
    c, _, _ = gpu_run(p, ic, [BC_FREE, BC_FREE], 0.2)
    assert np.array_equal(a.get_convar(), c.get_convar())
    assert np.array_equal(a.get_alpha(), c.get_alpha())


def test_1d_sinwave_periodic_conserves():
    from gks_amd.params import sinwave1d_params
    from gks_amd.problems import sinwave_prim_1d
    p = sinwave1d_params(80)
    ic = sinwave_prim_1d(p)
    say("[1d sinwave] n=80 periodic to t=2")
    s, n, _ = gpu_run(p, ic, [BC_PERIODIC, BC_PERIODIC], 2.0)
    u = s.get_convar()[:, G:-G]
    from gks_amd.solver1d import prim_to_convar_1d
    u0 = prim_to_convar_1d(ic, p.gamma)[:, G:-G]
    assert np.allclose(u.sum(axis=1), u0.sum(axis=1), rtol=1e-13)


@pytest.mark.parametrize("use_df,bc", [(1, (BC_FREE, BC_FREE)), (0, (2, 2)), (1, (BC_PERIODIC, BC_PERIODIC))])
def test_sweep_by_sweep_loop_equals_the_persistent_run(use_df, bc):
    """SodTubeProblem's loop on the host, one ABI call per sweep (copy old,
    Get_CFL, per stage the two side fills, reconstruction with WENO5_AO or,
    from step 1, WENO5_AO_with_DF, flux, update, DF sensor), gives the
    persistent kernel's state, alpha and dt sequence bit for bit."""
    p = sod_params(200, use_df=use_df)
    steps = 25
    a = Solver1D(p)
    a.init_prim(sod_prim(p))
    _, hist = a.run(list(bc), 0.2, steps, record_dt=True)
    b = Solver1D(p)
    b.init_prim(sod_prim(p))
    dts = []
    for _ in range(steps):
        b.copy_old()
        dt = b.cfl(0.2)
        with_df = bool(p.df_from_start) or (bool(use_df) and b.step_count > 0)
        for stage in range(2):
            b.boundary(0, bc[0])
            b.boundary(1, bc[1])
            b.reconstruct(with_df)
            b.flux(stage, dt)
            b.update(stage)
            if use_df:
                b.update_alpha()
        b.t += dt
        b.step_count += 1
        dts.append(dt)
    print(f"[1d sweeps] use_df {use_df} bc {bc}: t {a.t} vs {b.t}", flush=True)
    assert np.array_equal(np.array(dts), hist)
    assert a.t == b.t
    assert np.array_equal(a.get_convar(), b.get_convar())
    assert np.array_equal(a.get_alpha(), b.get_alpha())
    a.close()
    b.close()
