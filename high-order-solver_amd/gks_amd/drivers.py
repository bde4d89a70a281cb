"""Problem drivers over the GPU solvers, and the reference's output formats.

The drivers restate the set-up and time loop of the reference's problem
functions (reference src/1Dproblem.cpp, src/2Dproblem.cpp) for the
workloads of this build; every step runs on the GPU through the C ABI.

* ``run_sod``      - SodTubeProblem (1Dproblem.cpp:3-149), N cells to t=0.2;
* ``run_riemann``  - PlanarShock-style four-quadrant Riemann problem
                     (2Dproblem.cpp:3-168 with IC_for_riemann_2d, :345-385)
                     from seeded quadrant states;
* ``run_vortex``   - isentropic vortex advection on the periodic [0,10]^2,
                     with the L1 error against the exact solution (the
                     initial field moved by t (1, 1), periodically);
* ``sinwave_2d`` / ``accuracy_sinwave_2d`` - the 2D accuracy test
                     (2Dproblem.cpp:1323-1603): the periodic density wave
                     to t=2 with WENO9-AO-DF (4 Gauss points) on meshes
                     10 * 2^k, L1 / L2 / Linf density errors against the
                     exact cell averages and the observed orders;
* ``write_output1d`` / ``write_output2d`` - the text formats of output1d /
  output2d (output.cpp:3-15, 64-84): density of the interior cells, one
  value per line, six significant digits, x fastest in 2D;
* ``write_error_form`` - output_error_form (output.cpp:17-61), the error /
  order table of the accuracy tests.
"""
import os
import time

import numpy as np

from . import Solver2D
from .params import BC_FREE, BC_PERIODIC, planar_shock_params, sinwave2d_params, sod_params, vortex_params
from .problems import riemann_prim, riemann_states, sinwave2d_prim, sinwave2d_rho, sod_prim, vortex_convar
from .solver1d import Solver1D


def _g(x):
    return "%g" % x


def write_output1d(path, convar, p):
    """output1d: interior densities, one per line (`R.txt`)."""
    G = p.ghost
    with open(path, "w") as f:
        f.write("\n".join(_g(v) for v in convar[0, G:G + p.nodex]) + "\n")


def write_output2d(directory, convar, p, step):
    """output2d: `R-<step>.txt` with the interior densities, x fastest."""
    G = p.ghost
    rho = convar[0, G:G + p.nodex, G:G + p.nodey]
    path = os.path.join(directory, f"R-{step}.txt")
    with open(path, "w") as f:
        f.write("\n".join(_g(v) for v in rho.T.reshape(-1)) + "\n")
    return path


def run_sod(n=400, tstop=0.2, use_df=0, fast_math=1):
    p = sod_params(n, use_df=use_df, fast_math=fast_math)
    s = Solver1D(p)
    s.init_prim(sod_prim(p))
    t0 = time.perf_counter()
    steps = s.run([BC_FREE, BC_FREE], tstop)
    wall = time.perf_counter() - t0
    u = s.get_convar()
    s.close()
    return dict(params=p, convar=u, steps=steps, wall_s=wall, mcells=n * steps / wall / 1e6)


def run_riemann(n, seed=0, tstop=0.25, weno_order=5, fast_math=1, max_steps=2**31 - 1):
    p = planar_shock_params(n, weno_order=weno_order, fast_math=fast_math)
    s = Solver2D(p)
    s.init_prim(riemann_prim(p, riemann_states(seed)))
    s.sync()
    t0 = time.perf_counter()
    steps = s.run([BC_FREE] * 4, tstop, max_steps)
    s.sync()
    wall = time.perf_counter() - t0
    u = s.get_convar()
    s.close()
    return dict(params=p, convar=u, steps=steps, wall_s=wall, mcells=n * n * steps / wall / 1e6)


def vortex_l1(convar, p, t):
    """Mean absolute error per conservative variable over the interior
    against the exact cell averages at time t."""
    G = p.ghost
    exact = vortex_convar(p, shift=(t, t))
    d = np.abs(convar[:, G:G + p.nodex, G:G + p.nodey] - exact[:, G:G + p.nodex, G:G + p.nodey])
    return d.mean(axis=(1, 2))


def run_vortex(n, tstop=10.0, weno_order=7, fast_math=1, max_steps=2**31 - 1, steps_per_call=None):
    """Isentropic vortex to tstop; steps_per_call splits the run into calls
    of that many steps (for progress output); returns the L1 errors."""
    p = vortex_params(n, weno_order=weno_order, fast_math=fast_math)
    s = Solver2D(p)
    s.init_convar(vortex_convar(p))
    s.sync()
    t0 = time.perf_counter()
    steps = 0
    chunk = steps_per_call or max_steps
    while s.t < tstop and steps < max_steps:
        k = s.run([BC_PERIODIC] * 4, tstop, min(chunk, max_steps - steps))
        steps += k
        if steps_per_call:
            print(f"vortex n={n}: step {steps} t={s.t:.4f}", flush=True)
        if k == 0:
            break
    s.sync()
    wall = time.perf_counter() - t0
    u = s.get_convar()
    t = s.t
    s.close()
    return dict(params=p, convar=u, steps=steps, t=t, wall_s=wall, mcells=n * n * steps / wall / 1e6,
                l1=vortex_l1(u, p, t))


def sinwave2d_errors(convar, p, t, tstop):
    """error_for_sinwave_2d (2Dproblem.cpp:1563-1601): L1 (mean), L2 (root
    mean square) and Linf of the interior density against the exact cell
    averages, computed only when t is within 1e-10 of tstop (else None)."""
    if abs(t - tstop) > 1e-10:
        return None
    G = p.ghost
    d = convar[0, G:G + p.nodex, G:G + p.nodey] - sinwave2d_rho(p)[G:G + p.nodex, G:G + p.nodey]
    return [float(np.abs(d).mean()), float(np.sqrt((d * d).mean())), float(np.abs(d).max())]


def sinwave_2d(n, cfl=0.05, tstop=2.0, fast_math=0):
    """One mesh of the 2D accuracy test on the GPU: n x n cells of the
    periodic [0, 2]^2 to tstop; returns the errors, steps and timing."""
    p = sinwave2d_params(n, cfl=cfl, fast_math=fast_math)
    s = Solver2D(p)
    s.init_prim(sinwave2d_prim(p))
    s.sync()
    t0 = time.perf_counter()
    steps = s.run([BC_PERIODIC] * 4, tstop)
    s.sync()
    wall = time.perf_counter() - t0
    u = s.get_convar()
    t = s.t
    s.close()
    return dict(params=p, n=n, steps=steps, t=t, wall_s=wall, errors=sinwave2d_errors(u, p, t, tstop))


def convergence_orders(errors):
    """log2 of the ratio of consecutive meshes' errors (output_error_form)."""
    e = np.asarray(errors, dtype=np.float64)
    return (np.log(e[:-1] / e[1:]) / np.log(2.0)).tolist()


def accuracy_sinwave_2d(mesh_set=4, mesh_number_start=10, cfl=0.05, fast_math=0, path=None):
    """accuracy_sinwave_2d (2Dproblem.cpp:1323-1346): meshes 10 * 2^k,
    k < mesh_set; returns meshes, errors (per mesh [L1, L2, Linf]) and the
    orders between consecutive meshes; with path, writes the table in
    output_error_form's format."""
    meshes = [mesh_number_start * 2 ** k for k in range(mesh_set)]
    runs = [sinwave_2d(n, cfl=cfl, fast_math=fast_math) for n in meshes]
    errors = [r["errors"] for r in runs]
    out = dict(meshes=meshes, errors=errors, orders=convergence_orders(errors),
               steps=[r["steps"] for r in runs], wall_s=[r["wall_s"] for r in runs])
    if path:
        write_error_form(path, cfl, 1.0, meshes, errors)
    return out


def write_error_form(path, cfl, dt_ratio, meshes, errors):
    """output_error_form (output.cpp:17-61): the CFL number and dt ratio,
    then one row per mesh '1/N & L1 & order & L2 & order & Linf & order \\\\'
    (the first mesh has '~' for its orders); errors in scientific notation
    with 6 digits, orders fixed with 2."""
    orders = np.log(np.asarray(errors[:-1]) / np.asarray(errors[1:])) / np.log(2.0)
    lines = [f"the CFL number is {cfl:g}", f"the dt ratio over dx is {dt_ratio:g}"]
    for i, n in enumerate(meshes):
        e = errors[i]
        if i == 0:
            lines.append(f"1/{n} & {e[0]:.6e} & ~ & {e[1]:.6e} & ~ & {e[2]:.6e} & ~ \\\\ ")
        else:
            o = orders[i - 1]
            lines.append(f"1/{n} & {e[0]:.6e} & {o[0]:.2f} & {e[1]:.6e} & {o[1]:.2f} & {e[2]:.6e} & {o[2]:.2f} \\\\ ")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def _run_chunks(s, bc, tstop, chunk, log, tag):
    total = 0
    t0 = time.perf_counter()
    while s.t < tstop:
        n = s.run(bc, tstop, chunk)
        total += n
        if log:
            log(f"{tag}: t={s.t:.4f} steps={total} wall={time.perf_counter() - t0:.1f}s")
        if n == 0:
            break
    s.sync()
    return total, time.perf_counter() - t0


def equal_l1(n, seed=0, tstop=0.25, classifier=None, smooth_flux=True, fast_math=1, chunk=500, log=None):
    """The learned-versus-plain accuracy check of the four-quadrant Riemann
    workload: density L1 of the plain and the learned run at n^2 against a
    plain run at (2n)^2 averaged over 2x2 blocks, all to tstop; plus each
    run's steps and wall time.  classifier: (params, threshold); default the
    packaged one."""
    from . import learn
    P, thr = classifier if classifier is not None else learn.load()[:2]
    bc = [BC_FREE] * 4
    G = 5
    out = {}
    finals = {}
    for tag, m, use in (("plain", n, False), ("learned", n, True), ("reference_2n", 2 * n, False)):
        p = planar_shock_params(m, fast_math=fast_math)
        s = Solver2D(p)
        s.init_prim(riemann_prim(p, riemann_states(seed)))
        if use:
            learn.set_classifier(s, P, thr)
            learn.classifier_options(s, smooth_flux)
        steps, wall = _run_chunks(s, bc, tstop, chunk, log, tag)
        finals[tag] = s.get_convar()[:, G:G + m, G:G + m]
        out[tag] = {"steps": steps, "wall_s": round(wall, 3), "t": s.t,
                    "mcell_updates_per_s": round(m * m * steps / wall / 1e6, 3)}
        if use:
            lin, tot = learn.classifier_stats(s)
            out[tag]["linear_share"] = round(lin / max(tot, 1), 4)
        s.close()
    ref = finals["reference_2n"].reshape(4, n, 2, n, 2).mean(axis=(2, 4))
    for tag in ("plain", "learned"):
        e = np.abs(finals[tag] - ref).mean(axis=(1, 2))
        out[tag]["l1"] = [float(x) for x in e]
    out["l1_ratio_density"] = out["learned"]["l1"][0] / out["plain"]["l1"][0]
    out["speedup"] = out["learned"]["mcell_updates_per_s"] / out["plain"]["mcell_updates_per_s"]
    return out
