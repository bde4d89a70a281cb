"""Writes the golden vectors under tests/golden/ from the compiled reference
(oracle/_ref/libgksref.so, built by `make -C oracle ref` from the reference's
own sources).  Run in the build container, where the reference exists:

    python tests/golden/make_golden.py

Every fixture is data only: the inputs drawn here and the reference's outputs
for them.  The tests compare the oracle (and through it the HIP path) against
these vectors.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "high-order-solver_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

from gks_amd.params import (BC_FREE, BC_PERIODIC, BC_REFLECT, blastwave_params, planar_shock_params,  # noqa: E402
                            shu_osher_params, sinwave1d_params, sinwave2d_params, sod_params)
from gks_amd.problems import (PLANAR_SHOCK_ZONES, blastwave_prim, riemann_prim, riemann_states,  # noqa: E402
                              shu_osher_prim, sinwave2d_prim, sinwave2d_rho, sinwave_prim_1d, sod_prim)
from oracle_lib import reference, reference1d  # noqa: E402


def random_points(rng, n, gamma=1.4):
    """Physical left/right states (convar, der1x, der1y) in a face frame."""
    def state(m):
        rho = rng.uniform(0.1, 2.0, m)
        u = rng.uniform(-1.5, 1.5, m)
        v = rng.uniform(-1.5, 1.5, m)
        p = rng.uniform(0.05, 2.0, m)
        E = p / (gamma - 1) + 0.5 * rho * (u * u + v * v)
        out = np.zeros((m, 12))
        out[:, 0], out[:, 1], out[:, 2], out[:, 3] = rho, rho * u, rho * v, E
        out[:, 4:12] = rng.normal(0.0, 0.5, (m, 8)) * rng.choice([0.0, 1.0, 10.0], (m, 1))
        return out
    left = state(n)
    right = left.copy()
    k = n // 2  # half the points are small perturbations (smooth faces)
    right[:k] = left[:k] * (1 + rng.normal(0, 1e-3, (k, 12)))
    right[k:] = state(n - k)
    return left, right


def gen_flux_points(ref):
    rng = np.random.default_rng(1234)
    out = {}
    for name, kw in (("euler", {}), ("c0", dict(c1_euler=0.0, c2_euler=0.0)),
                     ("ns", dict(tau_type=1, mu=1e-3, c1_euler=0.0, c2_euler=1.0)),
                     ("prandtl", dict(tau_type=1, mu=1e-3, pr=0.72, prandtl_fix=1))):
        p = planar_shock_params(16, **kw)
        left, right = random_points(rng, 256)
        dt = 10 ** rng.uniform(-4, -2)
        c, f, d = ref.flux_points(p, left, right, dt)
        out.update({f"{name}_left": left, f"{name}_right": right, f"{name}_dt": np.array(dt),
                    f"{name}_center": c, f"{name}_f": f, f"{name}_derf": d})
    np.savez_compressed(os.path.join(HERE, "gks2d_flux_points.npz"), **out)


def gen_weno(ref):
    rng = np.random.default_rng(99)
    out = {}
    n = 256
    for name, recon in (("char", 1), ("cons", 0)):
        p = planar_shock_params(16, recon_variable=recon)
        # 5-cell stencils: smooth ramps, random, and jumps
        base = np.array([1.0, 0.3, -0.2, 2.5])
        w5 = np.zeros((n, 5, 4))
        for k in range(n):
            kind = k % 3
            b = base * rng.uniform(0.5, 1.5, 4)
            b[3] = max(b[3], 1.0 + 0.5 * (b[1] ** 2 + b[2] ** 2) / b[0])
            if kind == 0:
                s = rng.normal(0, 0.02, 4)
                w5[k] = b + np.outer(np.arange(-2, 3), s)
            elif kind == 1:
                w5[k] = b * (1 + rng.normal(0, 0.1, (5, 4)))
            else:
                j = rng.integers(1, 5)
                w5[k] = b
                w5[k, j:] = b * rng.uniform(0.3, 0.9, 4)
            w5[k, :, 3] = np.maximum(w5[k, :, 3], 0.6 + 0.5 * (w5[k, :, 1] ** 2 + w5[k, :, 2] ** 2) / w5[k, :, 0])
        alpha = rng.choice([0.0, 0.0, 1.5, 3.0, 8.0], (n, 5))
        h = 1.0 / 64
        lo, ro = ref.weno5_normal(p, w5.reshape(n, 20), alpha, h)
        line5 = np.zeros((n, 5, 8))
        line5[:, :, :4] = w5
        line5[:, :, 4:] = rng.normal(0, 0.3, (n, 5, 4))
        gp2 = ref.weno5_tangent(p, line5.reshape(n, 40), alpha, h)
        out.update({f"{name}_w5": w5, f"{name}_alpha": alpha, f"{name}_h": np.array(h), f"{name}_left": lo,
                    f"{name}_right": ro, f"{name}_line5": line5, f"{name}_gp2": gp2})
    np.savez_compressed(os.path.join(HERE, "weno5_df.npz"), **out)


def gen_runs(ref):
    out = {}
    cases = [
        # name, n, zones, bc, steps, tstop
        ("riemann_seed0", 32, riemann_states(0), [BC_FREE] * 4, 30, 0.25),
        ("planar_rm2", 40, PLANAR_SHOCK_ZONES, [BC_FREE] * 4, 25, 0.6),
        ("periodic", 24, riemann_states(7), [BC_PERIODIC] * 4, 20, 0.25),
        ("reflect", 24, riemann_states(11), [BC_REFLECT] * 4, 20, 0.25),
        ("mixed", 28, riemann_states(3), [BC_REFLECT, BC_FREE, BC_PERIODIC, BC_PERIODIC], 20, 0.25),
    ]
    for name, n, zones, bc, steps, tstop in cases:
        p = planar_shock_params(n)
        prim = riemann_prim(p, zones, (0.5, 0.5) if name != "planar_rm2" else (0.7, 0.7))
        s = ref.solver(p)
        s.set_prim(prim)
        rc, hist = s.run(bc, tstop, steps)
        assert rc == 0
        st = s.get()
        xf, yf = s.face_fluxes(1)
        G = 5
        out.update({
            f"{name}_n": np.array(n), f"{name}_zones": np.asarray(zones), f"{name}_bc": np.array(bc),
            f"{name}_steps": np.array(steps), f"{name}_tstop": np.array(tstop),
            f"{name}_discon": np.array((0.5, 0.5) if name != "planar_rm2" else (0.7, 0.7)),
            f"{name}_convar": st["convar"][:, G:-G, G:-G], f"{name}_alpha_x": st["alpha_x"][G:-G, G:-G],
            f"{name}_alpha_y": st["alpha_y"][G:-G, G:-G], f"{name}_dt": hist[:steps], f"{name}_t": np.array(st["t"]),
        })
        if name == "riemann_seed0":  # per-face fluxes of the last stage, one case only (size)
            out.update({f"{name}_xflux": xf[G:n + G + 1, G:n + G], f"{name}_yflux": yf[G:n + G, G:n + G + 1]})
        s.close()
    np.savez_compressed(os.path.join(HERE, "runs2d.npz"), **out)


# 1D cases: name -> (params, initial condition, bc, tstop)
RUNS1D = {
    "sod400": (lambda: sod_params(400), sod_prim, [BC_FREE, BC_FREE], 0.2),
    "sod400_df": (lambda: sod_params(400, use_df=1), sod_prim, [BC_FREE, BC_FREE], 0.2),
    "sod100_cons_df": (lambda: sod_params(100, use_df=1, recon_variable=0), sod_prim, [BC_FREE, BC_FREE], 0.2),
    "blastwave400": (lambda: blastwave_params(400), blastwave_prim, [BC_REFLECT, BC_REFLECT], 0.038),
    "shu_osher200_df": (lambda: shu_osher_params(200, use_df=1), shu_osher_prim, [BC_FREE, BC_FREE], 1.8),
    "sinwave40": (lambda: sinwave1d_params(40), sinwave_prim_1d, [BC_PERIODIC, BC_PERIODIC], 2.0),
}


def gen_runs1d(ref1):
    out = {}
    for name, (mk, ic, bc, tstop) in RUNS1D.items():
        p = mk()
        s = ref1.solver(p)
        s.set_prim(ic(p))
        rc, hist = s.run(bc, tstop, 100000)
        assert rc == 0
        st = s.get()
        G = p.ghost
        out.update({f"{name}_convar": st["convar"][:, G:-G], f"{name}_alpha": st["alpha"][G:-G],
                    f"{name}_dt": hist[:st["step"]], f"{name}_t": np.array(st["t"]),
                    f"{name}_step": np.array(st["step"])})
        if name == "sod400_df":  # both stages' face fluxes of the last step
            for k in range(2):
                out[f"{name}_faces{k}"] = s.face_fluxes(k)[G:p.nodex + G + 1]
        s.close()
    np.savez_compressed(os.path.join(HERE, "runs1d.npz"), **out)


def random_points_1d(rng, n, gamma=1.4):
    """Physical left/right 1D states (convar[3], der1[3])."""
    def state(m):
        rho = rng.uniform(0.1, 2.0, m)
        u = rng.uniform(-1.5, 1.5, m)
        p = rng.uniform(0.05, 2.0, m)
        out = np.zeros((m, 6))
        out[:, 0], out[:, 1], out[:, 2] = rho, rho * u, p / (gamma - 1) + 0.5 * rho * u * u
        out[:, 3:6] = rng.normal(0.0, 0.5, (m, 3)) * rng.choice([0.0, 1.0, 10.0], (m, 1))
        return out
    left = state(n)
    right = left.copy()
    k = n // 2
    right[:k] = left[:k] * (1 + rng.normal(0, 1e-3, (k, 6)))
    right[k:] = state(n - k)
    return left, right


FLUX1D_CASES = {"euler": {}, "c0": dict(c1_euler=0.0, c2_euler=0.0),
                "ns": dict(tau_type=1, mu=1e-3, c1_euler=0.0, c2_euler=1.0)}


def gen_flux_points_1d(ref1):
    rng = np.random.default_rng(4321)
    out = {}
    for name, kw in FLUX1D_CASES.items():
        p = sod_params(16, **kw)
        left, right = random_points_1d(rng, 256)
        dt = 10 ** rng.uniform(-4, -2)
        c, f, d = ref1.flux_points(p, left, right, dt)
        out.update({f"{name}_left": left, f"{name}_right": right, f"{name}_dt": np.array(dt),
                    f"{name}_center": c, f"{name}_f": f, f"{name}_derf": d})
    np.savez_compressed(os.path.join(HERE, "gks1d_flux_points.npz"), **out)


def weno_states(rng, n, m):
    """n stencils of m cells (x-like frame conservative states): smooth
    ramps, noisy, and single jumps, with positive pressure"""
    base = np.array([1.0, 0.3, -0.2, 2.5])
    w = np.zeros((n, m, 4))
    for k in range(n):
        b = base * rng.uniform(0.5, 1.5, 4)
        b[3] = max(b[3], 1.0 + 0.5 * (b[1] ** 2 + b[2] ** 2) / b[0])
        kind = k % 3
        if kind == 0:
            w[k] = b + np.outer(np.arange(m) - m // 2, rng.normal(0, 0.02, 4))
        elif kind == 1:
            w[k] = b * (1 + rng.normal(0, 0.1, (m, 4)))
        else:
            j = rng.integers(1, m)
            w[k] = b
            w[k, j:] = b * rng.uniform(0.3, 0.9, 4)
        w[k, :, 3] = np.maximum(w[k, :, 3], 0.6 + 0.5 * (w[k, :, 1] ** 2 + w[k, :, 2] ** 2) / w[k, :, 0])
    return w


def gen_weno79(ref):
    """WENO7/9-AO-DF leaves: normal (both faces of a cell) and tangential
    (both sides of a face, 3 or 4 Gauss points)"""
    rng = np.random.default_rng(79)
    out = {}
    n = 96
    for order in (7, 9):
        for name, recon in (("char", 1), ("cons", 0)):
            p = planar_shock_params(16, weno_order=order, recon_variable=recon)
            key = f"o{order}_{name}"
            h = 1.0 / 64
            w = weno_states(rng, n, order)
            a = rng.choice([0.0, 0.0, 1.5, 3.0, 8.0], (n, order))
            lo, ro = ref.weno_normal(p, order, w, a, h)
            lineL = np.zeros((n, order, 8))
            lineR = np.zeros((n, order, 8))
            lineL[:, :, :4] = weno_states(rng, n, order)
            lineR[:, :, :4] = weno_states(rng, n, order)
            lineL[:, :, 4:] = rng.normal(0, 0.3, (n, order, 4))
            lineR[:, :, 4:] = rng.normal(0, 0.3, (n, order, 4))
            a2 = rng.choice([0.0, 0.0, 1.5, 3.0, 8.0], (n, order))
            gl, gr = ref.weno_tangent(p, order, lineL, lineR, a, a2, h)
            out.update({f"{key}_w": w, f"{key}_alpha": a, f"{key}_alpha2": a2, f"{key}_h": np.array(h),
                        f"{key}_left": lo, f"{key}_right": ro, f"{key}_lineL": lineL, f"{key}_lineR": lineR,
                        f"{key}_gpL": gl, f"{key}_gpR": gr})
    np.savez_compressed(os.path.join(HERE, "weno79_df.npz"), **out)


# WENO7 / WENO9 runs: name -> (order, recon, n, zones seed, bc, steps)
RUNS79 = {
    "w7_char_free": (7, 1, 28, 0, [BC_FREE] * 4, 20),
    "w7_cons_periodic": (7, 0, 24, 7, [BC_PERIODIC] * 4, 15),
    "w9_char_mixed": (9, 1, 24, 3, [BC_REFLECT, BC_FREE, BC_PERIODIC, BC_PERIODIC], 15),
    "w9_cons_free": (9, 0, 24, 11, [BC_FREE] * 4, 15),
}


def gen_runs79(ref):
    out = {}
    G = 5
    for name, (order, recon, n, seed, bc, steps) in RUNS79.items():
        p = planar_shock_params(n, weno_order=order, recon_variable=recon)
        prim = riemann_prim(p, riemann_states(seed))
        s = ref.solver(p)
        s.set_prim(prim)
        rc, hist = s.run(bc, 0.25, steps)
        assert rc == 0
        st = s.get()
        out.update({f"{name}_convar": st["convar"][:, G:-G, G:-G], f"{name}_alpha_x": st["alpha_x"][G:-G, G:-G],
                    f"{name}_alpha_y": st["alpha_y"][G:-G, G:-G], f"{name}_dt": hist[:steps],
                    f"{name}_t": np.array(st["t"])})
        s.close()
    np.savez_compressed(os.path.join(HERE, "runs79.npz"), **out)


SINWAVE2D_MESHES = (10, 20, 40)


def gen_sinwave2d(ref):
    """The reference's 2D accuracy test (sinwave_2d, 2Dproblem.cpp:1348-1527,
    WENO9-AO-DF with 4 Gauss points, DF sensor never updated) on meshes 10,
    20, 40 to t = 2: per mesh the steps, final t and the interior density;
    the errors against the exact cell averages are formed from these by the
    tests (error_for_sinwave_2d's L1 / L2 / Linf)."""
    out = {}
    G = 5
    for n in SINWAVE2D_MESHES:
        p = sinwave2d_params(n)
        s = ref.solver(p)
        s.set_prim(sinwave2d_prim(p))
        rc, hist = s.run([BC_PERIODIC] * 4, 2.0, 10**7)
        assert rc == 0
        st = s.get()
        out.update({f"n{n}_rho": st["convar"][0, G:-G, G:-G], f"n{n}_steps": np.array(st["step"]),
                    f"n{n}_t": np.array(st["t"]), f"n{n}_dt": hist[:st["step"]]})
        s.close()
        print("sinwave2d", n, st["step"], st["t"], flush=True)
    np.savez_compressed(os.path.join(HERE, "sinwave2d.npz"), **out)


def main(which=None):
    ref = reference()
    if which == "sinwave2d":
        gen_sinwave2d(ref)
        return
    gen_flux_points(ref)
    gen_weno(ref)
    gen_runs(ref)
    gen_weno79(ref)
    gen_runs79(ref)
    gen_sinwave2d(ref)
    ref1 = reference1d()
    gen_runs1d(ref1)
    gen_flux_points_1d(ref1)
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else None)
