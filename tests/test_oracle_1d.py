"""Pins the 1D oracle (oracle/solver1d.hpp) before it checks the HIP 1D path:

* against golden vectors the compiled reference produced (tests/golden/
  runs1d.npz, gks1d_flux_points.npz);
* against the compiled reference directly, when oracle/_ref is built;
* against what the mathematics fixes: the half-space Maxwellian moments by
  quadrature, free-stream exactness and mirror symmetry of the GKS flux, the
  exact Riemann solution of the Sod tube, conservation on periodic and
  walled domains, and the convergence order on the smooth sine wave.
"""
import os

import numpy as np
import pytest
from scipy import integrate, optimize

from gks_amd.params import (BC_FREE, BC_PERIODIC, BC_REFLECT, blastwave_params, sinwave1d_params,
                            sod_params)
from gks_amd.problems import blastwave_prim, sinwave_exact_1d, sinwave_prim_1d, sod_prim
from oracle_lib import have_ref, oracle1d, reference1d

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
G = 5

# the golden cases (tests/golden/make_golden.py, RUNS1D)
RUNS = {
    "sod400": dict(n=400, kw={}, bc=[BC_FREE, BC_FREE], tstop=0.2),
    "sod400_df": dict(n=400, kw=dict(use_df=1), bc=[BC_FREE, BC_FREE], tstop=0.2),
    "sod100_cons_df": dict(n=100, kw=dict(use_df=1, recon_variable=0), bc=[BC_FREE, BC_FREE], tstop=0.2),
    "blastwave400": dict(n=400, kw=dict(c1_euler=0.01, c2_euler=5.0), bc=[BC_REFLECT] * 2, tstop=0.038),
    "shu_osher200_df": dict(n=200, kw=dict(use_df=1, left=0.0, right=10.0), bc=[BC_FREE, BC_FREE], tstop=1.8),
    "sinwave40": dict(n=40, kw=dict(recon_variable=0, df_from_start=1, c1_euler=0.0, c2_euler=0.0, right=2.0,
                                    cfl=0.1), bc=[BC_PERIODIC] * 2, tstop=2.0),
}


def case_ic(name, p):
    from gks_amd.problems import shu_osher_prim
    for key, fn in (("sod", sod_prim), ("blastwave", blastwave_prim), ("shu_osher", shu_osher_prim),
                    ("sinwave", sinwave_prim_1d)):
        if name.startswith(key):
            return fn(p)
    raise KeyError(name)


def run_oracle(p, ic, bc, tstop, max_steps=100000, lib=None):
    s = (lib or oracle1d()).solver(p)
    s.set_prim(ic)
    rc, hist = s.run(bc, tstop, max_steps)
    assert rc == 0
    st = s.get()
    st["dt"] = hist[:st["step"]]
    st["solver"] = s
    return st


# ------------------------------------------------------------ golden vectors
@pytest.mark.parametrize("case", ["euler", "c0", "ns"])
def test_flux_points_1d_match_reference_golden(case):
    g = np.load(os.path.join(GOLD, "gks1d_flux_points.npz"))
    kw = {"euler": {}, "c0": dict(c1_euler=0.0, c2_euler=0.0),
          "ns": dict(tau_type=1, mu=1e-3, c1_euler=0.0, c2_euler=1.0)}[case]
    p = sod_params(16, **kw)
    c, f, d = oracle1d().flux_points(p, g[f"{case}_left"], g[f"{case}_right"], float(g[f"{case}_dt"]))
    assert np.array_equal(c, g[f"{case}_center"])
    assert np.array_equal(f, g[f"{case}_f"])
    assert np.array_equal(d, g[f"{case}_derf"])


@pytest.mark.parametrize("name", list(RUNS))
def test_runs_1d_match_reference_golden(name):
    g = np.load(os.path.join(GOLD, "runs1d.npz"))
    c = RUNS[name]
    p = sod_params(c["n"], **c["kw"])
    st = run_oracle(p, case_ic(name, p), c["bc"], c["tstop"])
    assert st["step"] == int(g[f"{name}_step"])
    assert st["t"] == float(g[f"{name}_t"])
    assert np.array_equal(st["dt"], g[f"{name}_dt"])
    assert np.array_equal(st["convar"][:, G:-G], g[f"{name}_convar"])
    assert np.array_equal(st["alpha"][G:-G], g[f"{name}_alpha"])
    if name == "sod400_df":
        for k in range(2):
            assert np.array_equal(st["solver"].face_fluxes(k)[G:p.nodex + G + 1], g[f"{name}_faces{k}"])


@pytest.mark.skipif(not have_ref(), reason="compiled reference (oracle/_ref) not built")
@pytest.mark.parametrize("seed,recon,df", [(0, 1, 1), (1, 0, 0), (2, 1, 0), (3, 0, 1)])
def test_oracle_1d_equals_compiled_reference(seed, recon, df):
    """Random two-state Riemann problems, both reconstruction variables."""
    rng = np.random.default_rng(seed)
    p = sod_params(64 + 8 * seed, recon_variable=recon, use_df=df)
    prim = np.zeros((3, p.nx))
    cut = rng.integers(G + 10, p.nx - G - 10)
    for side, sl in ((0, slice(0, cut)), (1, slice(cut, p.nx))):
        prim[:, sl] = np.array([rng.uniform(0.1, 2), rng.uniform(-1, 1), rng.uniform(0.05, 2)])[:, None]
    a = run_oracle(p, prim, [BC_FREE, BC_REFLECT], 0.15, 60)
    b = run_oracle(p, prim, [BC_FREE, BC_REFLECT], 0.15, 60, lib=reference1d())
    assert a["step"] == b["step"]
    assert np.array_equal(a["dt"], b["dt"])
    assert np.array_equal(a["convar"][:, G:-G], b["convar"][:, G:-G])
    assert np.array_equal(a["alpha"], b["alpha"])


# ------------------------------------------------------- mathematics: flux
def conv1(rho, u, p, gamma=1.4):
    return np.array([rho, rho * u, p / (gamma - 1) + 0.5 * rho * u * u])


def test_collision_state_is_half_space_quadrature():
    """With zero slopes the collision state is rho_L times the u>0 moments of
    the left Maxwellian plus rho_R times the u<0 moments of the right one;
    the internal (xi) degrees contribute <xi^2> = K / (2 lambda)."""
    K = 4
    p = sod_params(16)
    L = conv1(1.3, 0.4, 0.9)
    R = conv1(0.5, -0.7, 0.3)
    left = np.concatenate([L, np.zeros(3)])[None, :]
    right = np.concatenate([R, np.zeros(3)])[None, :]
    c, _, _ = oracle1d().flux_points(p, left, right, 1e-3)

    def half(state, sign):
        rho, m, E = state
        u0 = m / rho
        lam = (K + 1) * 0.25 * rho / (E - 0.5 * rho * u0 * u0)
        w = lambda u: rho * np.sqrt(lam / np.pi) * np.exp(-lam * (u - u0) ** 2)
        lim = (0, np.inf) if sign > 0 else (-np.inf, 0)
        mass = integrate.quad(w, *lim, epsabs=1e-14, epsrel=1e-13)[0]
        mom = integrate.quad(lambda u: u * w(u), *lim, epsabs=1e-14, epsrel=1e-13)[0]
        en = integrate.quad(lambda u: 0.5 * (u * u + K / (2 * lam)) * w(u), *lim, epsabs=1e-14, epsrel=1e-13)[0]
        return np.array([mass, mom, en])

    expect = half(L, +1) + half(R, -1)
    assert np.allclose(c[0, :3], expect, rtol=1e-11, atol=1e-13)
    assert np.array_equal(c[0, 3:], np.zeros(3))


def test_uniform_state_flux_is_dt_times_euler_flux():
    """Equilibrium free stream: F(t) is the Euler flux, so 4 F(dt/2) - F(dt)
    = dt * Euler flux and derf = 0."""
    p = sod_params(16)
    rng = np.random.default_rng(5)
    for _ in range(8):
        rho, u, pr = rng.uniform(0.2, 2), rng.uniform(-2, 2), rng.uniform(0.1, 2)
        U = conv1(rho, u, pr)
        s = np.concatenate([U, np.zeros(3)])[None, :]
        dt = 10 ** rng.uniform(-4, -2)
        _, f, d = oracle1d().flux_points(p, s, s, dt)
        euler = np.array([rho * u, rho * u * u + pr, u * (U[2] + pr)])
        assert np.allclose(f[0], dt * euler, rtol=1e-12, atol=1e-15 * dt)
        assert np.max(np.abs(d[0])) < 1e-12 * dt * np.max(np.abs(euler))


def test_flux_mirror_symmetry():
    """x -> -x: swap sides, negate momentum and the slopes of rho and rhoE;
    the mass and energy fluxes change sign, the momentum flux does not."""
    g = np.load(os.path.join(GOLD, "gks1d_flux_points.npz"))
    p = sod_params(16)
    left, right = g["euler_left"], g["euler_right"]
    dt = float(g["euler_dt"])
    flip = np.array([1.0, -1.0, 1.0, -1.0, 1.0, -1.0])
    _, f, d = oracle1d().flux_points(p, left, right, dt)
    _, fm, dm = oracle1d().flux_points(p, right * flip, left * flip, dt)
    sgn = np.array([-1.0, 1.0, -1.0])
    scale = np.max(np.abs(f), axis=1, keepdims=True)
    assert np.max(np.abs(fm * sgn - f) / scale) < 1e-10
    dscale = np.maximum(np.max(np.abs(d), axis=1, keepdims=True), scale)
    assert np.max(np.abs(dm * sgn - d) / dscale) < 1e-10


# -------------------------------------------------------- mathematics: runs
def sod_exact(x, t, gamma=1.4, L=(1.0, 0.0, 1.0), R=(0.125, 0.0, 0.1), x0=0.5):
    """Exact Riemann solution (two-rarefaction / shock pressure function,
    Newton-free: brentq on the pressure function)."""
    rl, ul, pl = L
    rr, ur, pr = R
    cl, cr = np.sqrt(gamma * pl / rl), np.sqrt(gamma * pr / rr)

    def fk(p, rk, pk, ck):
        if p > pk:
            A, B = 2 / ((gamma + 1) * rk), (gamma - 1) / (gamma + 1) * pk
            return (p - pk) * np.sqrt(A / (p + B))
        return 2 * ck / (gamma - 1) * ((p / pk) ** ((gamma - 1) / (2 * gamma)) - 1)

    ps = optimize.brentq(lambda p: fk(p, rl, pl, cl) + fk(p, rr, pr, cr) + ur - ul, 1e-8, 10, xtol=1e-15)
    us = 0.5 * (ul + ur) + 0.5 * (fk(ps, rr, pr, cr) - fk(ps, rl, pl, cl))
    # left rarefaction (ps < pl), right shock (ps > pr) for Sod
    rsl = rl * (ps / pl) ** (1 / gamma)
    csl = cl * (ps / pl) ** ((gamma - 1) / (2 * gamma))
    rsr = rr * ((ps / pr + (gamma - 1) / (gamma + 1)) / ((gamma - 1) / (gamma + 1) * ps / pr + 1))
    sr = ur + cr * np.sqrt((gamma + 1) / (2 * gamma) * ps / pr + (gamma - 1) / (2 * gamma))
    s = (x - x0) / t
    rho = np.empty_like(s)
    head, tail = ul - cl, us - csl
    rho[s <= head] = rl
    fan = (s > head) & (s <= tail)
    rho[fan] = rl * (2 / (gamma + 1) + (gamma - 1) / ((gamma + 1) * cl) * (ul - s[fan])) ** (2 / (gamma - 1))
    rho[(s > tail) & (s <= us)] = rsl
    rho[(s > us) & (s <= sr)] = rsr
    rho[s > sr] = rr
    return rho


@pytest.mark.parametrize("df", [0, 1])
def test_sod_matches_exact_riemann_solution(df):
    p = sod_params(400, use_df=df)
    st = run_oracle(p, sod_prim(p), [BC_FREE, BC_FREE], 0.2)
    assert abs(st["t"] - 0.2) < 1e-12
    dx = p.dx
    sub = 32
    xs = (np.arange(p.nodex)[:, None] + (np.arange(sub)[None, :] + 0.5) / sub) * dx
    exact = sod_exact(xs.ravel(), st["t"]).reshape(p.nodex, sub).mean(axis=1)
    l1 = np.mean(np.abs(st["convar"][0, G:-G] - exact))
    # WENO5-class schemes give ~1e-3 at 400 cells; a wrong term gives O(1e-2)
    assert l1 < 2.5e-3, l1


def test_sinwave_fifth_order_and_periodic_conservation():
    errs = []
    for n in (20, 40, 80):
        p = sinwave1d_params(n)
        ic = sinwave_prim_1d(p)
        st = run_oracle(p, ic, [BC_PERIODIC, BC_PERIODIC], 2.0)
        exact = sinwave_exact_1d(p)
        errs.append(np.mean(np.abs(st["convar"][0, G:-G] - exact)))
        # conservation: the totals of all three variables
        s = oracle1d().solver(p)
        s.set_prim(ic)
        u0 = s.get()["convar"][:, G:-G].sum(axis=1)
        assert np.allclose(st["convar"][:, G:-G].sum(axis=1), u0, rtol=1e-13)
    orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert orders[-1] > 4.0, (errs, orders)


def test_reflective_walls_conserve_mass_and_energy():
    p = blastwave_params(200)
    ic = blastwave_prim(p)
    st = run_oracle(p, ic, [BC_REFLECT, BC_REFLECT], 0.01)
    s = oracle1d().solver(p)
    s.set_prim(ic)
    u0 = s.get()["convar"][:, G:-G].sum(axis=1)
    u1 = st["convar"][:, G:-G].sum(axis=1)
    assert abs(u1[0] - u0[0]) < 1e-12 * u0[0]
    assert abs(u1[2] - u0[2]) < 1e-12 * u0[2]
