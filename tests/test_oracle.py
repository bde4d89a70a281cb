"""Pins the oracle (CPU restatement, oracle/) before it is trusted as the
checker of the HIP path:

* against golden vectors the compiled reference produced (tests/golden/);
* against the compiled reference directly, when oracle/_ref is built;
* against what the mathematics fixes: half-space Maxwellian moments by
  quadrature, free-stream exactness of the GKS flux, polynomial exactness
  and convergence order of the WENO-AO reconstructions, conservation on a
  periodic domain, and the mirror symmetry of the four-quadrant problem.
"""
import os

import numpy as np
import pytest
from scipy import integrate
from scipy.special import erfc

from gks_amd.params import BC_FREE, BC_PERIODIC, planar_shock_params
from gks_amd.problems import riemann_prim, riemann_states
from oracle_lib import have_ref, oracle, reference

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
G = 5


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300 + np.max(np.abs(b)) * 1e-14)))


# ------------------------------------------------------------ golden vectors
@pytest.mark.parametrize("case", ["euler", "c0", "ns", "prandtl"])
def test_flux_points_match_reference_golden(case):
    g = np.load(os.path.join(GOLD, "gks2d_flux_points.npz"))
    kw = {"euler": {}, "c0": dict(c1_euler=0.0, c2_euler=0.0),
          "ns": dict(tau_type=1, mu=1e-3, c1_euler=0.0, c2_euler=1.0),
          "prandtl": dict(tau_type=1, mu=1e-3, pr=0.72, prandtl_fix=1)}[case]
    p = planar_shock_params(16, **kw)
    c, f, d = oracle().flux_points(p, g[f"{case}_left"], g[f"{case}_right"], float(g[f"{case}_dt"]))
    # same compiler family, same libm, no FMA: the restatement is bit-exact
    np.testing.assert_array_equal(c, g[f"{case}_center"])
    np.testing.assert_array_equal(f, g[f"{case}_f"])
    np.testing.assert_array_equal(d, g[f"{case}_derf"])


@pytest.mark.parametrize("name,recon", [("char", 1), ("cons", 0)])
def test_weno5_df_match_reference_golden(name, recon):
    g = np.load(os.path.join(GOLD, "weno5_df.npz"))
    p = planar_shock_params(16, recon_variable=recon)
    n = g[f"{name}_w5"].shape[0]
    lo, ro = oracle().weno5_normal(p, g[f"{name}_w5"].reshape(n, 20), g[f"{name}_alpha"], float(g[f"{name}_h"]))
    np.testing.assert_array_equal(lo, g[f"{name}_left"])
    np.testing.assert_array_equal(ro, g[f"{name}_right"])
    gp2 = oracle().weno5_tangent(p, g[f"{name}_line5"].reshape(n, 40), g[f"{name}_alpha"], float(g[f"{name}_h"]))
    np.testing.assert_array_equal(gp2, g[f"{name}_gp2"])


RUNS = ["riemann_seed0", "planar_rm2", "periodic", "reflect", "mixed"]


def run_case_from_golden(lib, g, name):
    n = int(g[f"{name}_n"])
    p = planar_shock_params(n)
    prim = riemann_prim(p, g[f"{name}_zones"], tuple(g[f"{name}_discon"]))
    s = lib.solver(p)
    s.set_prim(prim)
    steps = int(g[f"{name}_steps"])
    rc, hist = s.run(list(g[f"{name}_bc"]), float(g[f"{name}_tstop"]), steps)
    assert rc == 0
    return s, hist[:steps]


@pytest.mark.parametrize("name", RUNS)
def test_runs_match_reference_golden(name):
    g = np.load(os.path.join(GOLD, "runs2d.npz"))
    s, hist = run_case_from_golden(oracle(), g, name)
    st = s.get()
    np.testing.assert_array_equal(hist, g[f"{name}_dt"])
    np.testing.assert_array_equal(st["convar"][:, G:-G, G:-G], g[f"{name}_convar"])
    np.testing.assert_array_equal(st["alpha_x"][G:-G, G:-G], g[f"{name}_alpha_x"])
    np.testing.assert_array_equal(st["alpha_y"][G:-G, G:-G], g[f"{name}_alpha_y"])
    if f"{name}_xflux" in g:
        n = int(g[f"{name}_n"])
        xf, yf = s.face_fluxes(1)
        np.testing.assert_array_equal(xf[G:n + G + 1, G:n + G], g[f"{name}_xflux"])
        np.testing.assert_array_equal(yf[G:n + G, G:n + G + 1], g[f"{name}_yflux"])


@pytest.mark.skipif(not have_ref(), reason="compiled reference (oracle/_ref) not built here")
@pytest.mark.parametrize("seed,n,recon", [(21, 20, 1), (22, 26, 0), (23, 18, 1)])
def test_oracle_equals_compiled_reference(seed, n, recon):
    p = planar_shock_params(n, recon_variable=recon)
    prim = riemann_prim(p, riemann_states(seed))
    outs = []
    for lib in (oracle(), reference()):
        s = lib.solver(p)
        s.set_prim(prim)
        rc, hist = s.run([BC_FREE] * 4, 0.25, 12)
        assert rc == 0
        outs.append((s.get(), hist[:12], s.gauss()))
    (a, ha, ga), (b, hb, gb) = outs
    np.testing.assert_array_equal(ha, hb)
    np.testing.assert_array_equal(a["convar"][:, G:-G, G:-G], b["convar"][:, G:-G, G:-G])
    np.testing.assert_array_equal(a["alpha_x"], b["alpha_x"])
    np.testing.assert_array_equal(ga[0][G:n + G + 1, G:n + G], gb[0][G:n + G + 1, G:n + G])
    np.testing.assert_array_equal(ga[1][G:n + G, G:n + G + 1], gb[1][G:n + G, G:n + G + 1])


# ------------------------------------------------------- closed-form checks
def conv(rho, u, v, p, gamma=1.4):
    return np.array([rho, rho * u, rho * v, p / (gamma - 1) + 0.5 * rho * (u * u + v * v)])


def test_collision_state_is_half_space_quadrature():
    """Centre state = rho_L <psi>_{u>0}(g_L) + rho_R <psi>_{u<0}(g_R), each
    half-moment integrated numerically over the 1D Maxwellian in u."""
    K = 3
    p = planar_shock_params(8)
    rng = np.random.default_rng(5)
    for _ in range(6):
        sl = (rng.uniform(0.2, 2), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.1, 2))
        sr = (rng.uniform(0.2, 2), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.1, 2))
        L = np.zeros((1, 12))
        R = np.zeros((1, 12))
        L[0, :4] = conv(*sl)
        R[0, :4] = conv(*sr)
        c, _, _ = oracle().flux_points(p, L, R, 1e-3)
        expect = np.zeros(4)
        for (rho, u, v, pr), lo, hi in ((sl, 0, np.inf), (sr, -np.inf, 0)):
            lam = rho / (2 * pr)  # lambda = rho / (2p)
            g = lambda x, n: x ** n * np.sqrt(lam / np.pi) * np.exp(-lam * (x - u) ** 2)  # noqa: E731
            m0 = integrate.quad(g, lo, hi, args=(0,), epsabs=1e-14, epsrel=1e-13)[0]
            m1 = integrate.quad(g, lo, hi, args=(1,), epsabs=1e-14, epsrel=1e-13)[0]
            m2 = integrate.quad(g, lo, hi, args=(2,), epsabs=1e-14, epsrel=1e-13)[0]
            # <v^2> = v^2 + 1/(2 lam); <xi^2> = K/(2 lam)
            expect += rho * np.array([m0, m1, v * m0, 0.5 * (m2 + m0 * (v * v + 0.5 / lam) + m0 * K * 0.5 / lam)])
        assert rel_err(c[0, :4], expect) < 1e-10


def test_uniform_state_flux_is_dt_times_euler_flux():
    """No gradients and L = R: F(t) = t F_Euler(U) for any collision time, so
    f = 4F(dt/2) - F(dt) = dt F_Euler and derf = 0."""
    rng = np.random.default_rng(6)
    for kw in ({}, dict(c1_euler=0.0, c2_euler=0.0), dict(tau_type=1, mu=1e-2)):
        p = planar_shock_params(8, **kw)
        n = 16
        st = np.zeros((n, 12))
        prim = np.stack([rng.uniform(0.2, 2, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n),
                         rng.uniform(0.1, 2, n)], 1)
        for k in range(n):
            st[k, :4] = conv(*prim[k])
        dt = 3e-3
        c, f, d = oracle().flux_points(p, st, st, dt)
        r, u, v, pr = prim.T
        E = st[:, 3]
        euler = np.stack([r * u, r * u * u + pr, r * u * v, u * (E + pr)], 1)
        np.testing.assert_allclose(c[:, :4], st[:, :4], rtol=1e-13, atol=1e-15)
        np.testing.assert_allclose(f, dt * euler, rtol=1e-12, atol=1e-15)
        assert np.max(np.abs(d)) < 1e-13 * np.max(np.abs(dt * euler))


def test_uniform_gradient_free_flux_respects_mirror_symmetry():
    """Reflecting the normal direction (u -> -u, left <-> right) negates the
    mass, energy and tangential-momentum fluxes and keeps the normal one."""
    p = planar_shock_params(8)
    rng = np.random.default_rng(8)
    n = 12
    L = np.zeros((n, 12))
    R = np.zeros((n, 12))
    for k in range(n):
        L[k, :4] = conv(rng.uniform(0.2, 2), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.1, 2))
        R[k, :4] = conv(rng.uniform(0.2, 2), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.1, 2))
    flip = np.array([1, -1, 1, 1])
    _, f1, _ = oracle().flux_points(p, L, R, 1e-3)
    Lm = R.copy()
    Rm = L.copy()
    Lm[:, :4] *= flip
    Rm[:, :4] *= flip
    _, f2, _ = oracle().flux_points(p, Lm, Rm, 1e-3)
    np.testing.assert_allclose(f2, -f1 * flip, rtol=1e-11, atol=1e-14)


def cell_avg_poly(coef, x0, h):
    """cell averages over [x0 - h/2, x0 + h/2] of sum coef[k] x^k"""
    P = np.polynomial.Polynomial(coef).integ()
    return (P(x0 + h / 2) - P(x0 - h / 2)) / h


def test_weno5_df_exact_for_quadratics():
    """With alpha = 0 (df = 1) and cell averages of a quadratic, every
    candidate polynomial is exact and the AO weights sum to one: the face
    values and first derivatives are exact (conservative variables)."""
    p = planar_shock_params(8, recon_variable=0)
    h = 0.1
    rng = np.random.default_rng(3)
    xs = np.arange(-2, 3) * h
    for _ in range(5):
        coefs = [np.array([rng.uniform(1, 2), rng.normal(), rng.normal(), 0.0]) for _ in range(4)]
        coefs[3][0] += 10.0
        w5 = np.stack([cell_avg_poly(c, xs, h) for c in coefs], 1)  # [5, 4]
        lo, ro = oracle().weno5_normal(p, w5.reshape(1, 20), np.zeros((1, 5)), h)
        for k, c in enumerate(coefs):
            P = np.polynomial.Polynomial(c)
            assert abs(lo[0, k] - P(-h / 2)) < 1e-12 * (1 + abs(P(-h / 2)))
            assert abs(ro[0, k] - P(h / 2)) < 1e-12 * (1 + abs(P(h / 2)))
            assert abs(lo[0, 4 + k] - P.deriv()(-h / 2)) < 1e-9 * (1 + abs(P.deriv()(-h / 2)))
            assert abs(ro[0, 4 + k] - P.deriv()(h / 2)) < 1e-9 * (1 + abs(P.deriv()(h / 2)))
        # tangential: Gauss points -/+ h/(2 sqrt 3), value and derivative
        line5 = np.zeros((1, 5, 8))
        line5[0, :, :4] = w5
        line5[0, :, 4:] = w5 * 0.5
        gp = oracle().weno5_tangent(p, line5.reshape(1, 40), np.zeros((1, 5)), h)
        for q, xq in enumerate((-h / (2 * np.sqrt(3)), h / (2 * np.sqrt(3)))):
            for k, c in enumerate(coefs):
                P = np.polynomial.Polynomial(c)
                assert abs(gp[0, q, k] - P(xq)) < 1e-12 * (1 + abs(P(xq)))
                assert abs(gp[0, q, 8 + k] - P.deriv()(xq)) < 1e-9 * (1 + abs(P.deriv()(xq)))
                assert abs(gp[0, q, 4 + k] - 0.5 * P(xq)) < 1e-12 * (1 + abs(P(xq)))


def test_weno5_df_fifth_order_on_smooth_data():
    """Face value error on cell averages of a smooth function converges at
    (close to) fifth order: pins the large-stencil polynomial."""
    p = planar_shock_params(8, recon_variable=0)
    errs = []
    for h in (0.04, 0.02, 0.01):
        xs = 0.3 + np.arange(-2, 3) * h
        w = (-np.cos(xs + h / 2) + np.cos(xs - h / 2)) / h  # averages of sin
        w5 = np.stack([w + 2.0, w * 0.1, w * 0.2, w + 5.0], 1)
        lo, ro = oracle().weno5_normal(p, w5.reshape(1, 20), np.zeros((1, 5)), h)
        errs.append(abs(ro[0, 0] - 2.0 - np.sin(0.3 + h / 2)))
    orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(orders > 4.5), orders


def smooth_periodic_prim(p):
    nx, ny = p.nx, p.ny
    x = (np.arange(nx) - G + 0.5) / p.nodex
    y = (np.arange(ny) - G + 0.5) / p.nodey
    X, Y = np.meshgrid(x, y, indexing="ij")
    prim = np.stack([1 + 0.2 * np.sin(2 * np.pi * X) * np.sin(2 * np.pi * Y), 0.5 + 0 * X, 0.3 + 0 * X,
                     1.0 + 0 * X])
    prim[:, :G] = 0
    prim[:, -G:] = 0
    prim[:, :, :G] = 0
    prim[:, :, -G:] = 0
    return prim


def test_periodic_run_conserves_mass_momentum_energy():
    """Smooth periodic data (the DF sensor stays below its threshold, so the
    ghost cells' alpha = 0 matches the interior): sums are conserved.  With a
    shock on a periodic seam the reference is not conservative, because ghost
    alpha is never written (reconstruction.cpp:1523-1534)."""
    p = planar_shock_params(24)
    prim = smooth_periodic_prim(p)
    s = oracle().solver(p)
    s.set_prim(prim)
    u0 = s.get()["convar"][:, G:-G, G:-G].sum(axis=(1, 2))
    rc, _ = s.run([BC_PERIODIC] * 4, 0.25, 15)
    assert rc == 0
    u1 = s.get()["convar"][:, G:-G, G:-G].sum(axis=(1, 2))
    scale = np.abs(s.get()["convar"][:, G:-G, G:-G]).sum(axis=(1, 2))
    assert np.all(np.abs(u1 - u0) < 1e-12 * scale)


def test_four_quadrant_mirror_symmetry():
    """Zones symmetric under x <-> y (with u <-> v): the solution stays
    symmetric.  Exercises the y-face rotation and the reversed tangential
    stencil against the x path."""
    p = planar_shock_params(24)
    z = riemann_states(5)
    z[2] = [z[0][0], z[0][2], z[0][1], z[0][3]]  # zone 3 = mirror of zone 1 ... not needed on diagonal
    z = np.array([[1.0, 0.3, 0.3, 1.0], [0.5, -0.2, 0.4, 0.6], [0.8, -0.1, -0.1, 0.9], [0.5, 0.4, -0.2, 0.6]])
    prim = riemann_prim(p, z)
    s = oracle().solver(p)
    s.set_prim(prim)
    rc, _ = s.run([BC_FREE] * 4, 0.25, 10)
    assert rc == 0
    u = s.get()["convar"][:, G:-G, G:-G]
    np.testing.assert_allclose(u[0], u[0].T, rtol=1e-11, atol=1e-13)
    np.testing.assert_allclose(u[1], u[2].T, rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(u[3], u[3].T, rtol=1e-11, atol=1e-13)
