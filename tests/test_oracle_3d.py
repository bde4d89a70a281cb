"""The 3D restatement (oracle/solver3d.hpp) on the CPU.  The reference's 3D
path cannot run (see include/gks3d.h), so it is pinned by what the
reference's working 2D path and the mathematics fix:

* the reference's own 3D leaves, compiled (oracle/ref_harness.cpp): the
  face, line and point reconstructions, and GKS3D where its lambda bug is
  silent (states whose pressure equals lambda);
* the reference's own 3D driver loop, made runnable by the harness (the
  Gauss-point records allocated, the flux with lambda from the state);
* GKS3D on planar data (w = 0, no z slopes, K = 2) is the reference's GKS2D
  (K = 3) through the compiled reference (or the 2D oracle when the
  reference is not built);
* a z-invariant 3D run in conservative variables is the reference's 2D run
  with the same WENO5-AO leaves (the non-DF path, WENO5_AO_normal /
  WENO5_AO_tangent; in characteristic variables the 3D tangential steps
  project the carried derivatives too, the 2D ones do not);
* free-stream exactness of the flux, its v <-> w symmetry, exactness of the
  face / line / point reconstructions on quadratics, the axis-permutation
  symmetry of the cubic Sod problem, conservation on a periodic cube."""
import ctypes

import numpy as np
import pytest

from gks_amd.params import (BC_FREE, BC_PERIODIC, Params3D, cubic_params, planar_shock_params, tgv_params)
from gks_amd.problems import (convar_to_prim_3d, cubic_sod_prim, prim_to_convar_3d, riemann_prim, riemann_states,
                              tgv_convar)
from oracle_lib import have_ref, oracle, oracle3, reference, reference3

needs_ref = pytest.mark.skipif(not have_ref(), reason="compiled reference not built (oracle/Makefile target ref)")


def test_params3_layout_matches_c():
    off = (ctypes.c_int64 * 32)()
    size = oracle3().f.orc3d_params_layout(off)
    assert size == ctypes.sizeof(Params3D)
    for k, (name, _) in enumerate(Params3D._fields_):
        assert off[k] == getattr(Params3D, name).offset, name


def states3(n, seed, slope=0.3, planar=False):
    rng = np.random.default_rng(seed)
    r = rng.uniform(0.5, 2.0, n)
    u, v, w = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    p = rng.uniform(0.5, 2.0, n)
    if planar:
        w = np.zeros(n)
    c = prim_to_convar_3d(np.stack([r, u, v, w, p]), 1.4).T
    out = np.zeros((n, 20))
    out[:, :5] = c
    for d in range(1, 4):
        out[:, 5 * d:5 * d + 5] = slope * c * rng.uniform(-1, 1, (n, 5))
    if planar:
        out[:, 3::5] = 0.0   # no w and no w slopes
        out[:, 15:20] = 0.0  # no z slopes
    return out


def test_free_stream_flux_is_exact():
    p = cubic_params(8)
    s = states3(200, 1, slope=0.0)
    dt = 1e-3
    _, f, d = oracle3().flux_points(p, s, s, dt)
    pr = convar_to_prim_3d(s[:, :5].T, 1.4)
    r, u, v, w, pp = pr
    E = s[:, 4]
    fe = np.stack([r * u, r * u * u + pp, r * u * v, r * u * w, u * (E + pp)], 1) * dt
    assert np.max(np.abs(f - fe) / np.abs(fe).max(axis=1, keepdims=True)) < 1e-13
    assert np.max(np.abs(d)) < 1e-13 * np.abs(fe).max()


@pytest.mark.parametrize("gas", ["euler", "ns"])
def test_planar_flux_is_the_reference_gks2d(gas):
    kw3 = dict(tau_type=1, mu=2e-3) if gas == "ns" else {}
    p3 = cubic_params(8, **kw3)
    p2 = planar_shock_params(8).copy(**kw3)
    L3, R3 = states3(500, 2, planar=True), states3(500, 3, planar=True)
    R3 = L3 + 0.2 * (R3 - L3)
    dt = 1e-3
    _, f3, d3 = oracle3().flux_points(p3, L3, R3, dt)
    keep = [0, 1, 2, 4]
    def to2(s):
        return np.concatenate([s[:, [0, 1, 2, 4]], s[:, [5, 6, 7, 9]], s[:, [10, 11, 12, 14]]], 1)
    lib = reference() if have_ref() else oracle()
    _, f2, d2 = lib.flux_points(p2, to2(L3), to2(R3), dt)
    sf = np.abs(f2).max(axis=1, keepdims=True)
    sd = np.abs(d2).max(axis=1, keepdims=True)
    assert np.max(np.abs(f3[:, keep] - f2) / sf) < 1e-12
    assert np.max(np.abs(d3[:, keep] - d2) / sd) < 1e-9
    assert np.max(np.abs(f3[:, 3]) / sf[:, 0]) < 1e-14


def test_flux_is_symmetric_in_v_and_w():
    p = cubic_params(8, tau_type=1, mu=1e-3)
    L, R = states3(300, 4), states3(300, 5)
    dt = 1e-3
    def swap(s):
        t = s.copy()
        for b in range(4):
            t[:, 5 * b + 2], t[:, 5 * b + 3] = s[:, 5 * b + 3], s[:, 5 * b + 2]
        t[:, 10:15], t[:, 15:20] = t[:, 15:20].copy(), t[:, 10:15].copy()
        return t
    _, f, d = oracle3().flux_points(p, L, R, dt)
    _, fs, ds = oracle3().flux_points(p, swap(L), swap(R), dt)
    sc = np.abs(f).max(axis=1, keepdims=True)
    assert np.max(np.abs(fs[:, [0, 1, 3, 2, 4]] - f) / sc) < 1e-13
    assert np.max(np.abs(ds[:, [0, 1, 3, 2, 4]] - d) / np.abs(d).max(axis=1, keepdims=True)) < 1e-9


def quad_avgs(a, b, c, xc, h):
    return a + b * xc + c * (xc * xc + h * h / 12.0)


def test_face_reconstruction_is_exact_for_quadratics():
    p = cubic_params(8, recon_variable=0)
    rng = np.random.default_rng(6)
    n, h = 100, 0.1
    A = rng.uniform(1, 2, (n, 5))
    B = rng.uniform(-1, 1, (n, 5))
    C = rng.uniform(-1, 1, (n, 5))
    A[:, 4] += 5.0  # keep the energy positive
    xs = (np.arange(5) - 2) * h
    w = np.stack([quad_avgs(A, B, C, x, h) for x in xs], 1)  # (n, 5 cells, 5 vars)
    lo, ro = oracle3().weno_face(p, w, h)
    xl, xr = -0.5 * h, 0.5 * h
    assert np.allclose(lo[:, :5], A + B * xl + C * xl * xl, rtol=0, atol=1e-13)
    assert np.allclose(ro[:, :5], A + B * xr + C * xr * xr, rtol=0, atol=1e-13)
    assert np.allclose(lo[:, 5:], B + 2 * C * xl, rtol=0, atol=1e-11)
    assert np.allclose(ro[:, 5:], B + 2 * C * xr, rtol=0, atol=1e-11)


@pytest.mark.parametrize("nc", [10, 15])
def test_tangential_reconstruction_is_exact_for_quadratics(nc):
    p = cubic_params(8, recon_variable=0)
    rng = np.random.default_rng(7)
    n, h = 100, 0.1
    A = rng.uniform(1, 2, (n, nc))
    B = rng.uniform(-1, 1, (n, nc))
    C = rng.uniform(-1, 1, (n, nc))
    A[:, 4] += 5.0
    xs = (np.arange(5) - 2) * h
    st = np.stack([quad_avgs(A, B, C, x, h) for x in xs], 1)  # (n, 5, nc)
    out = oracle3().weno_tangential(p, st, h)
    for q, xg in enumerate((-h / (2 * np.sqrt(3.0)), h / (2 * np.sqrt(3.0)))):
        assert np.allclose(out[:, q, :nc], A + B * xg + C * xg * xg, rtol=0, atol=1e-13)
        assert np.allclose(out[:, q, nc:], (B + 2 * C * xg)[:, :5], rtol=0, atol=1e-11)


def test_characteristic_reconstruction_of_uniform_data_is_the_data():
    p = cubic_params(8)
    s = states3(50, 8, slope=0.0)[:, :5]
    w = np.repeat(s[:, None, :], 5, axis=1)
    lo, ro = oracle3().weno_face(p, w, 0.1)
    assert np.allclose(lo[:, :5], s, rtol=1e-13, atol=0) and np.allclose(ro[:, :5], s, rtol=1e-13, atol=0)
    assert np.max(np.abs(lo[:, 5:])) < 1e-10


@needs_ref
def test_z_invariant_run_is_the_reference_2d_run():
    """A planar four-quadrant problem, uniform in z (periodic in z), against
    the reference's 2D WENO5-AO path without the DF factor, conservative
    variables.  The two paths round differently (3D and 2D moments), and
    the WENO-Z weights amplify that near flat data: 1e-8 after 12 steps."""
    n, nz, steps = 16, 4, 12
    p2 = planar_shock_params(n).copy(use_df=0, recon_variable=0)
    r2 = reference().solver(p2)
    prim2 = riemann_prim(p2, riemann_states(3))
    r2.set_prim(prim2)
    rc2, h2 = r2.run([BC_FREE] * 4, 1.0, steps)
    u2 = r2.get()["convar"]
    r2.close()
    p3 = cubic_params(n, ghost=5, front=nz / n, recon_variable=0)
    p3.nodez = nz
    pr3 = np.zeros((5, p3.nx, p3.ny, p3.nz))
    for k in range(p3.nz):
        pr3[[0, 1, 2, 4], :, :, k] = prim2
    with np.errstate(invalid="ignore"):  # the ghost layers are unset (zero) until the first boundary fill
        u30 = prim_to_convar_3d(pr3, 1.4)
    u3, t3, s3, rc3, h3 = oracle3().run(p3, u30, pr3,
                                        [BC_FREE, BC_FREE, BC_FREE, BC_FREE, BC_PERIODIC, BC_PERIODIC], 1.0, steps)
    assert rc2 == 0 and rc3 == 0 and s3 == steps
    assert np.max(np.abs(h3 - h2[:steps]) / h2[:steps]) < 1e-10
    G = 5
    a = u3[:, G:-G, G:-G, G + 1][[0, 1, 2, 4]]
    b = u2[:, G:-G, G:-G]
    assert np.max(np.abs(a - b)) < 1e-8 * np.max(np.abs(b))
    assert np.max(np.abs(u3[3, G:-G, G:-G, G:-G])) < 1e-12
    zi = u3[:, G:-G, G:-G, G:-G]
    assert np.max(np.abs(zi - zi[:, :, :, :1])) < 1e-12 * np.max(np.abs(zi))


@pytest.mark.parametrize("rv", [0, 1])
def test_cubic_sod_symmetries(rv):
    """Mirror symmetry about the centre in each axis holds to rounding.  The
    axis permutations hold only to truncation error: every face reconstructs
    along its two tangential directions in a fixed order (x faces y then z,
    y faces x then z, z faces y then x; reconstruction.cpp:3925-4161), and
    WENO's nonlinear weights make the two orders differ."""
    p = cubic_params(10, recon_variable=rv)
    prim = cubic_sod_prim(p)
    u, t, s, rc, _ = oracle3().run(p, prim_to_convar_3d(prim, 1.4), prim, [BC_FREE] * 6, 1.0, 6)
    assert rc == 0 and s == 6
    G = 3
    c = u[:, G:-G, G:-G, G:-G]
    sc = np.abs(c).max()
    for ax, mom in ((0, 1), (1, 2), (2, 3)):
        flip = [slice(None)] * 3
        flip[ax] = slice(None, None, -1)
        flip = tuple(flip)
        for v in range(5):
            sign = -1.0 if v == mom else 1.0
            assert np.max(np.abs(c[v] - sign * c[v][flip])) < 1e-12 * sc
    assert np.max(np.abs(c[0] - c[0].transpose(1, 0, 2))) < 1e-3 * sc
    assert np.max(np.abs(c[0] - c[0].transpose(0, 2, 1))) < 1e-3 * sc
    assert np.max(np.abs(c[1] - c[2].transpose(1, 0, 2))) < 1e-3 * sc
    assert np.max(np.abs(c[1] - c[3].transpose(2, 1, 0))) < 1e-3 * sc


def test_taylor_green_conserves_mass_momentum_energy():
    p = tgv_params(8)
    u0 = tgv_convar(p)
    prim = convar_to_prim_3d(u0, 1.4)
    u, t, s, rc, _ = oracle3().run(p, u0, prim, [BC_PERIODIC] * 6, 10.0, 4)
    assert rc == 0 and s == 4
    G = 3
    a = u0[:, G:-G, G:-G, G:-G].sum(axis=(1, 2, 3))
    b = u[:, G:-G, G:-G, G:-G].sum(axis=(1, 2, 3))
    scale = np.abs(u0[:, G:-G, G:-G, G:-G]).sum(axis=(1, 2, 3))
    scale[3] = scale[1]  # w is zero at t = 0
    assert np.all(np.abs(a - b) <= 1e-13 * scale)


# ------------------------------------------------ the reference's own 3D code
def smooth_stencils(n, seed, nc, h=0.1, amp=0.3):
    """n stencils of 5 cells with nc components: positive states (first 5),
    their derivatives (the rest), smooth in the stencil direction with a
    random kink in some."""
    rng = np.random.default_rng(seed)
    r = rng.uniform(0.5, 2.0, n)
    u, v, w = rng.uniform(-1, 1, (3, n))
    p = rng.uniform(0.5, 2.0, n)
    base = prim_to_convar_3d(np.stack([r, u, v, w, p]), 1.4).T
    out = np.zeros((n, 5, nc))
    x = (np.arange(5) - 2) * h
    for q in range(5):
        out[:, q, :5] = base * (1 + amp * np.sin(3 * x[q] + rng.uniform(0, 6, (n, 1))) * 0.3)
        for b in range(1, nc // 5):
            out[:, q, 5 * b:5 * b + 5] = base * rng.uniform(-1, 1, (n, 5))
    kink = rng.random(n) < 0.3
    out[kink, 3:, :5] *= 1.4
    return out


@needs_ref
@pytest.mark.parametrize("rv", [0, 1])
def test_face_values_are_the_reference_weno5_ao(rv):
    p = cubic_params(8, recon_variable=rv)
    w = smooth_stencils(400, 11 + rv, 5)
    lo, ro = oracle3().weno_face(p, w, 0.1)
    rlo, rro = reference3().weno_face(p, w, 0.1)
    sc = np.abs(rlo).max(axis=1, keepdims=True)
    assert np.max(np.abs(lo - rlo) / sc) < 1e-13
    assert np.max(np.abs(ro - rro) / sc) < 1e-13


@needs_ref
@pytest.mark.parametrize("rv", [0, 1])
@pytest.mark.parametrize("nc", [10, 15])
def test_line_and_point_values_are_the_reference(rv, nc):
    p = cubic_params(8, recon_variable=rv)
    st = smooth_stencils(400, 20 + nc + rv, nc)
    out = oracle3().weno_tangential(p, st, 0.1)
    ref = reference3().weno_tangential(p, st, 0.1)
    sc = np.abs(ref).max(axis=2, keepdims=True)
    assert np.max(np.abs(out - ref) / sc) < 1e-12


def p_equals_lambda_states(n, seed, slope=0.3):
    """Random states whose pressure equals lambda = 0.5 rho / p (K = 2,
    gamma 1.4): p = sqrt(rho / 2); the same state on both sides, slopes
    drawn independently per side."""
    rng = np.random.default_rng(seed)
    r = rng.uniform(0.5, 2.0, n)
    u, v, w = rng.uniform(-1, 1, (3, n))
    pr = np.sqrt(0.5 * r)
    c = prim_to_convar_3d(np.stack([r, u, v, w, pr]), 1.4).T
    L, R = np.zeros((n, 20)), np.zeros((n, 20))
    L[:, :5] = R[:, :5] = c
    for d in range(1, 4):
        L[:, 5 * d:5 * d + 5] = slope * c * rng.uniform(-1, 1, (n, 5))
        R[:, 5 * d:5 * d + 5] = slope * c * rng.uniform(-1, 1, (n, 5))
    return L, R


@needs_ref
@pytest.mark.parametrize("gas", ["euler", "ns"])
def test_flux_is_the_reference_gks3d_where_its_lambda_is_right(gas):
    kw = dict(tau_type=1, mu=2e-3) if gas == "ns" else {}
    p = cubic_params(8, **kw)
    L, R = p_equals_lambda_states(500, 31)
    _, f, d = oracle3().flux_points(p, L, R, 2e-3)
    rf, rd = reference3().gks3d(p, L, R, 2e-3)
    assert np.max(np.abs(f - rf) / np.abs(rf).max(axis=1, keepdims=True)) < 1e-12
    assert np.max(np.abs(d - rd) / np.abs(rd).max(axis=1, keepdims=True)) < 1e-9


@needs_ref
def test_reference_gks3d_departs_where_pressure_is_not_lambda():
    """The departure is real: with p != lambda the reference's GKS3D is not
    the flux of the state (the free-stream flux of a uniform state is off)."""
    p = cubic_params(8)
    s = states3(50, 32, slope=0.0)
    rf, _ = reference3().gks3d(p, s, s, 1e-3)
    _, f, _ = oracle3().flux_points(p, s, s, 1e-3)
    assert np.max(np.abs(rf - f) / np.abs(f).max(axis=1, keepdims=True)) > 1e-3


@needs_ref
@pytest.mark.parametrize("case", ["sod_char", "sod_cons", "smooth_ns"])
def test_run_is_the_reference_3d_driver(case):
    n, steps = 10, 5
    if case == "smooth_ns":
        p = cubic_params(n, tau_type=1, mu=1e-3)
        rng = np.random.default_rng(40)
        prim = np.zeros((5, p.nx, p.ny, p.nz))
        x = np.linspace(0, 1, p.nx)[:, None, None]
        y = np.linspace(0, 1, p.ny)[None, :, None]
        z = np.linspace(0, 1, p.nz)[None, None, :]
        prim[0] = 1 + 0.2 * np.sin(2 * np.pi * (x + 2 * y + z))
        prim[1] = 0.3 * np.cos(2 * np.pi * (y - z))
        prim[2] = 0.2 * np.sin(2 * np.pi * (x + z))
        prim[3] = -0.1 * np.cos(2 * np.pi * x)
        prim[4] = 1 + 0.1 * np.cos(2 * np.pi * (x - y))
    else:
        p = cubic_params(n, recon_variable=1 if case == "sod_char" else 0)
        prim = cubic_sod_prim(p)
    u0 = prim_to_convar_3d(prim, 1.4)
    u, t, s, rc, h = oracle3().run(p, u0, prim, [BC_FREE] * 6, 1.0, steps)
    ur, tr, sr, hr = reference3().run(p, u0, prim, 1.0, steps)
    assert rc == 0 and s == sr == steps
    assert np.max(np.abs(h - hr) / hr) < 1e-13
    G = 3
    a, b = u[:, G:-G, G:-G, G:-G], ur[:, G:-G, G:-G, G:-G]
    assert np.max(np.abs(a - b)) < 1e-11 * np.max(np.abs(b))
