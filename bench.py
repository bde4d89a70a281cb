"""Throughput of the GKS hot path on one MI355X.

Workload: the seeded four-quadrant Riemann problem on [0,1]^2 at 2048^2
(5 ghost layers), WENO5-AO with the DF factor in characteristic variables,
GKS2D (gks2nd_2d, Euler collision time c1=0.05, c2=1), S2O4, CFL 0.5,
gamma 1.4, K=3, free boundaries, from the initial condition to t=0.25.  A
step is one full S2O4 step (two stages) of the whole grid;
Mcell-updates/s = N^2 * steps / wall seconds / 1e6, with the state resident
in HBM.

The default measurement (the metric: learned vs plain at equal L1) times
three whole runs to t=0.25 - the plain path, the learned path (trained
troubled-interface classifier, BF16 MFMA inference) and the learned path
through the batched executor - and compares the density of each with a
plain run at (2N)^2 averaged over 2x2 blocks.  `value` is the learned
path's Mcell-updates/s; `steps` is its number of S2O4 steps; --warmup
untimed steps run on a separate solver before each leg.  --window instead
times --steps steps of one --path from t=0.

Prints one JSON line (rank 0).  Also reports:
  roofline          - the plain path's dominant kernel (k_tanflux, y faces):
                      algorithmic FP64 FLOPs per launch (the oracle's
                      operation count) over its average launch duration
                      from HIP events on the solver's stream in the timed
                      run;
  roofline_learned  - the learned path's full-work kernels (listed faces
                      and cells), each priced by the units it processed;
  cpu_baseline      - the compiled reference (oracle/_ref) or, if it is
                      not built, the oracle, timed on this host's cores on
                      a bounded sample of the same workload.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "high-order-solver_amd"))

# Algorithmic FP64 operation counts (+ - * / sqrt between runtime values;
# exp and erfc counted as one each) from the oracle's counting build
# (oracle/flopcount), WENO5-AO-DF, characteristic variables, 2 Gauss points.
FLOPS_PER_CELL_UPDATE = 57829.3 + 40.04 + 24.02  # N=1024 grid, one S2O4 step per interior cell
FLOPS_FLUX_POINT = 3041 + 43 + 5 + 3  # centre + GKS2D + DF alpha_k per (face, Gauss point)
FLOPS_TANGENT_SIDE = 2857  # tangential WENO per (face, side)
FLOPS_NORMAL_CELL = 2253  # normal WENO per (cell, direction)
# linear-weight tangential reconstruction of one face side (learned path):
# the fifth-order candidate at both Gauss points, counted by hand on its
# IEEE expressions (weno5_linear_2gauss): 70 per component with the
# tangential derivative x 4, 32 per component without it x 4 (der1x), plus
# the two order-reduction tests (12 each)
FLOPS_TANGENT_SIDE_LINEAR = 4 * 70 + 4 * 32 + 2 * 12
PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 (vector and matrix)
# HBM bytes per launch of the dominant kernel at 2048^2, fast kernels, from
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (profiles/gks2d_riemann2048_fast_hbm_traffic.md),
# each scaled by the counter's reading for k_copy_old (4 x 2048^2 doubles read
# and written): fetch x1.882, write x0.985
TRAFFIC_2048_FAST = {"flux_x": 0.836e9, "flux_y": 0.844e9}
# FP64 operations the fast flux kernel executes per launch at 2048^2, from
# rocprofv3 --pmc SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 (x64 lanes, FMA = 2)
# (profiles/gks2d_riemann2048_fast_hbm_traffic.md); below the algorithmic
# count: reciprocal multiplication, FMA contraction, whole-space tables
EXECUTED_FLOPS_2048_FAST = {"flux_x": 39.77e9, "flux_y": 39.77e9}
PEAK_HBM_GBS = 8000.0
# 3D units (oracle/flopcount3d, WENO5-AO characteristic, GKS3D gks2nd S2O4):
# FP64 operations per unit of each kernel phase
FLOPS3 = {"face": 2782,   # face values of one cell across one direction's faces
          "line": 3424,   # line values of one face side
          "point": 4796,  # point values of one face side along one line (inside the flux launch)
          "flux": 8115}   # GKS3D at one Gauss point (+ 4 exp, 2 erfc)
FLOPS3_CELL_STEP = 367644  # one interior cell's S2O4 step (2 stages x 3 directions)
# HBM bytes per k3_pflux launch at 256^3, fast kernels, mean of the three
# directions, from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE scaled by the
# k3_copy_old calibration (profiles/gks3d_tgv256_fast_hbm_traffic.md)
TRAFFIC3_256_FAST = {"flux": 13.47e9}
# FP64 operations k3_pflux executes per launch at 256^3, fast kernels
# (SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 x64, FMA = 2;
# profiles/gks3d_tgv256_fast_kernel_trace.md)
EXECUTED_FLOPS3_256_FAST = {"flux": 577.18e9}


def kernel_units(phase, n, gp=2, fused=False):
    """Work units one launch of `phase` processes on an n x n grid (G=5).
    With the tangential reconstruction inside the flux launch (WENO5), a
    flux launch does both: FLOPs per face = 2 sides x tangential + gp x
    flux point, per launch (n+1) n faces."""
    G = 5
    nx = n + 2 * G
    if phase in ("flux_x", "flux_y"):
        if fused:
            return (n + 1) * n, 2 * FLOPS_TANGENT_SIDE + gp * FLOPS_FLUX_POINT
        return gp * (n + 1) * n, FLOPS_FLUX_POINT
    if phase in ("tangent_x", "tangent_y"):
        return 2 * (n + 1) * n, FLOPS_TANGENT_SIDE
    if phase in ("normal_x", "normal_y"):
        return (n + 2) * nx, FLOPS_NORMAL_CELL
    return None, None


def cpu_baseline(n, threads, sample_steps, warm_steps=1):
    """Times the reference (or the oracle) on the host: `warm_steps` untimed
    S2O4 steps (first touch of the ~21 GB of records at 2048^2), then
    `sample_steps` timed steps of the same workload at n x n, in a child
    process."""
    code = f"""
import sys, time, json
sys.path.insert(0, {os.path.join(ROOT, 'high-order-solver_amd')!r}); sys.path.insert(0, {os.path.join(ROOT, 'tests')!r})
from oracle_lib import have_ref, reference, oracle
from gks_amd.params import planar_shock_params
from gks_amd.problems import riemann_prim, riemann_states
kind = 'reference' if have_ref() else 'port'
lib = reference() if kind == 'reference' else oracle()
p = planar_shock_params({n})
prim = riemann_prim(p, riemann_states(0))
s = lib.solver(p)
s.set_prim(prim)
rc, _ = s.run([0, 0, 0, 0], 0.25, {warm_steps}) if {warm_steps} else (0, None)
t0 = time.perf_counter()
rc2, _ = s.run([0, 0, 0, 0], 0.25, {sample_steps})
t1 = time.perf_counter()
print(json.dumps(dict(kind=kind, seconds=t1 - t0, rc=rc or rc2)))
"""
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=900)
    if out.returncode != 0:
        return None
    r = json.loads(out.stdout.strip().splitlines()[-1])
    if r["rc"] != 0:
        return None
    v = n * n * sample_steps / r["seconds"] / 1e6
    return {"value": round(v, 4), "unit": "Mcell-updates/s", "cores": threads, "kind": r["kind"],
            "sample": f"{sample_steps} timed S2O4 steps after {warm_steps} untimed of the same workload at {n}x{n} "
                      f"from t=0, {r['seconds']:.1f} s wall, OMP_NUM_THREADS={threads}"}


BATCH_SIDES = (128, 256, 384, 512, 768, 1024)


def batch_sizes(n, seed):
    """Side lengths of the batched workload: uniform over BATCH_SIDES."""
    import numpy as np
    rng = np.random.default_rng(seed)
    return [int(BATCH_SIDES[k]) for k in rng.integers(0, len(BATCH_SIDES), n)]


def _batch_run(params, prims, bc, tstop, cls=None, warmup=0, shape=None):
    """One whole batched run of every problem to tstop; the tile shape from
    the autotuner (or `shape`), found with the warm-up steps on a separate
    batch.  Returns (steps per problem, wall seconds, final interior
    densities, shape)."""
    from gks_amd import GksError, learn
    from gks_amd.batch import Batch

    def make():
        b = Batch(params)
        for k, s in enumerate(b.problems):
            s.init_prim(prims[k])
            if cls:
                learn.set_classifier(s, cls[0], cls[1])
        return b
    if shape is None:
        w = make()
        tf, lf, _ = w.autotune(bc, steps=2)
        if warmup:
            w.run(bc, tstop, warmup)
        w.close()
        shape = (tf, lf)
    b = make()
    b.set_shape(*shape)
    b.problems[0].sync()
    t0 = time.perf_counter()
    try:
        b.run(bc, tstop)
    except GksError as e:  # a problem whose scheme blew up stops alone; the others finish
        if e.code != 4:
            raise
    b.problems[0].sync()
    wall = time.perf_counter() - t0
    G = 5
    rho = [s.get_convar()[0, G:-G, G:-G] for s in b.problems]
    done = [s.step_count for s in b.problems]
    blown = [k for k, s in enumerate(b.problems) if s.t < tstop]
    b.close()
    return done, wall, rho, shape, blown


def bench_batch(args):
    """The batched workload: args.batch independent seeded four-quadrant
    Riemann problems with side lengths drawn uniformly from BATCH_SIDES,
    WENO5-AO-DF + GKS + S2O4, free boundaries, each to t = 0.2.  Legs, each
    a whole run of every problem: the learned path (packaged classifier,
    BF16 MFMA inference) through the batched executor (one launch per sweep
    over every problem's tiles, one captured HIP graph per step, autotuned
    tile shape) - the measured value; the plain path through the batched
    executor; the plain path one problem after the other (gks2d_run).  Per
    problem: density L1 of the learned batched and the plain single run
    against a plain run at (2n)^2 averaged over 2x2 blocks, their ratio, and
    the L1 difference between the two."""
    import numpy as np

    import gks_amd
    from gks_amd import GksError, learn
    from gks_amd.params import BC_FREE, planar_shock_params
    from gks_amd.problems import riemann_prim, riemann_states

    def log(m):
        print(m, file=sys.stderr, flush=True)

    tstop = 0.2
    sides = batch_sizes(args.batch, args.seed)
    seeds = [1000 + args.seed * 1000 + k for k in range(args.batch)]
    fast = 1 if args.variant == "fast" else 0
    params = [planar_shock_params(n, fast_math=fast) for n in sides]
    prims = [riemann_prim(p, riemann_states(sd)) for p, sd in zip(params, seeds)]
    P, thr, _ = learn.load(args.classifier or learn.DEFAULT_PARAMS)
    bc = [BC_FREE] * 4
    cells = sum(n * n for n in sides)

    def rate(steps, wall):
        return sum(n * n * k for n, k in zip(sides, steps)) / wall / 1e6

    steps_l, wall_l, rho_l, shape_l, blown_l = _batch_run(params, prims, bc, tstop, cls=(P, thr),
                                                           warmup=args.warmup)
    log(f"learned batched: {max(steps_l)} replays, {wall_l:.2f} s, {rate(steps_l, wall_l):.1f} Mcell-updates/s, "
        f"blown up: {blown_l}")
    steps_b, wall_b, _, shape_b, blown_b = _batch_run(params, prims, bc, tstop, warmup=args.warmup)
    log(f"plain batched: {wall_b:.2f} s, {rate(steps_b, wall_b):.1f} Mcell-updates/s, blown up: {blown_b}")
    # the plain path, one problem after the other
    steps_s, wall_s, rho_s, blown_s = [], 0.0, [], []
    for k, p in enumerate(params):
        s = gks_amd.Solver2D(p)
        s.init_prim(prims[k])
        s.sync()
        t0 = time.perf_counter()
        try:
            s.run(bc, tstop)
        except GksError as e:
            if e.code != 4:
                raise
            blown_s.append(k)
        s.sync()
        wall_s += time.perf_counter() - t0
        steps_s.append(s.step_count)
        rho_s.append(s.get_convar()[0, 5:-5, 5:-5])
        s.close()
    log(f"plain one by one: {wall_s:.2f} s, {rate(steps_s, wall_s):.1f} Mcell-updates/s, blown up: {blown_s}")
    value = rate(steps_l, wall_l)
    per = []
    if not args.no_l1:
        for k, p in enumerate(params):
            if k in blown_l or k in blown_s:
                continue
            n = p.nodex
            pr = planar_shock_params(2 * n, fast_math=fast)
            r = gks_amd.Solver2D(pr)
            r.init_prim(riemann_prim(pr, riemann_states(seeds[k])))
            r.run(bc, tstop)
            ref = r.get_convar()[0, 5:-5, 5:-5].reshape(n, 2, n, 2).mean(axis=(1, 3))
            r.close()
            el = float(np.abs(rho_l[k] - ref).mean())
            ep = float(np.abs(rho_s[k] - ref).mean())
            log(f"reference {2 * n}^2 of problem {k}: density L1 learned {el:.4e}, plain {ep:.4e}")
            per.append({"n": n, "seed": seeds[k], "steps": steps_l[k], "l1_learned": el, "l1_plain": ep,
                        "ratio": round(el / ep, 5), "l1_learned_minus_plain": float(np.abs(rho_l[k] - rho_s[k]).mean())})
        log(f"L1 ratios: max {max(e['ratio'] for e in per):.4f}")
    out = {
        "metric": "Mcell-updates/s FP64, batched uneven 2D four-quadrant Riemann problems, WENO5-AO GKS, learned "
                  "BF16 classifier, per-problem L1 vs plain FP64",
        "value": round(value, 3), "unit": "Mcell-updates/s", "n_gpus": 1, "steps": max(steps_l),
        "warmup": args.warmup,
        "window": f"every problem's whole run from the initial condition to t={tstop} ({max(steps_l)} batched "
                  "steps; a problem that reaches t=0.2 sits out the rest)",
        "ms_per_step": round(wall_l / max(steps_l) * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{args.batch} seeded four-quadrant Riemann problems, sides uniform over "
                               f"{list(BATCH_SIDES)} (seed {args.seed}), {cells / 1e6:.2f} M cells, to t={tstop}, "
                               f"WENO5-AO-DF characteristic, GKS2D gks2nd, S2O4, CFL 0.5, free BCs; learned path "
                               f"(packaged classifier, BF16 MFMA inference) through the batched executor "
                               f"({args.variant} kernels)",
                   "problems": args.batch, "sides": sides, "path": "learned", "kernels": args.variant},
        "autotune": {"learned": {"tanflux_faces": shape_l[0], "largest_first": shape_l[1]},
                     "plain": {"tanflux_faces": shape_b[0], "largest_first": shape_b[1]}},
        "blown_up": {"learned_batched": blown_l, "plain_batched": blown_b, "plain_one_by_one": blown_s,
                     "note": "problems whose seeded states drive the scheme to a negative density or pressure "
                             "(as the reference's own CPU solver does for such states) stop alone; the cell "
                             "updates they made count, and they are left out of the L1 comparison"},
        "plain_batched": {"value": round(rate(steps_b, wall_b), 3), "wall_s": round(wall_b, 3)},
        "plain_one_by_one": {"value": round(rate(steps_s, wall_s), 3), "wall_s": round(wall_s, 3)},
        "learned_batched_vs_plain_one_by_one": round(value / rate(steps_s, wall_s), 4),
        "learned_batched_vs_plain_batched": round(value / rate(steps_b, wall_b), 4),
        "roofline": None, "cpu_baseline": None,
    }
    if per:
        rs = [e["ratio"] for e in per]
        out["l1"] = {"reference": f"plain FP64 run of each problem at (2n)^2 to t={tstop}, 2x2 averaged; density",
                     "max_ratio": max(rs), "mean_ratio": round(float(np.mean(rs)), 5),
                     "problems_within_5pct": int(sum(r <= 1.05 for r in rs)), "per_problem": per}
    print(json.dumps(out), flush=True)


def units3(phase, n, G=3):
    """Work units of one launch of a 3D phase on an n^3 cube (the three
    directions' launches have the same count).  The flux launch computes the
    point values too: its unit is a face (4 point sides + 4 Gauss points)."""
    m = n + 2 * G
    if phase == "flux":
        return (n + 1) * n * n
    if phase == "line":
        return 2 * (n + 1) * n * (m - 2 * G + 4)
    if phase == "face":
        return (n + 2) * (m - 2 * G + 4) ** 2
    return None


def cpu_baseline3(n, threads, sample_steps):
    """Times the reference's 3D driver loop (oracle/_ref: its own boundary,
    reconstruction, flux loop and update code, with the harness's
    lambda-from-state GKS3D in flux_function_3d, free boundaries) or, if it
    is not built, the 3D oracle, on the host: `sample_steps` steps of the
    Taylor-Green cube at n^3 in a child process."""
    code = f"""
import sys, time, json
sys.path.insert(0, {os.path.join(ROOT, 'high-order-solver_amd')!r}); sys.path.insert(0, {os.path.join(ROOT, 'tests')!r})
from oracle_lib import have_ref, reference3, oracle3
from gks_amd.params import tgv_params
from gks_amd.problems import tgv_convar, convar_to_prim_3d
p = tgv_params({n})
u0 = tgv_convar(p, quad=2)
prim = convar_to_prim_3d(u0, 1.4)
if have_ref():
    import ctypes
    import numpy as np
    kind = 'reference'
    R = reference3()
    h = R.f.ref3d_create(ctypes.byref(p))
    R.f.ref3d_set_state(h, u0.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), prim.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    hist = np.zeros({sample_steps})
    t0 = time.perf_counter()
    R.f.ref3d_run(h, 100.0, {sample_steps}, hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    t1 = time.perf_counter()
    k = {sample_steps}
else:
    kind = 'port'
    t0 = time.perf_counter()
    u, t, k, rc, _ = oracle3().run(p, u0, prim, [1] * 6, 100.0, {sample_steps})
    t1 = time.perf_counter()
print(json.dumps(dict(kind=kind, seconds=t1 - t0, steps=k)))
"""
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=900)
    if out.returncode != 0:
        return None
    r = json.loads(out.stdout.strip().splitlines()[-1])
    v = n ** 3 * r["steps"] / r["seconds"] / 1e6
    return {"value": round(v, 4), "unit": "Mcell-updates/s", "cores": threads if r["kind"] == "reference" else 1,
            "kind": r["kind"],
            "sample": f"{r['steps']} S2O4 step(s) of the Taylor-Green cube at {n}^3, {r['seconds']:.1f} s wall"
                      + (f", OMP_NUM_THREADS={threads}; the reference's 3D loop with a lambda-from-state GKS3D "
                         "and its free boundaries (it has no periodic 3D boundary; the fills are under 1% of a step)"
                         if r["kind"] == "reference" else ", the 3D oracle, one thread")}


def bench_tgv3d(args):
    """3D workload: the viscous Taylor-Green vortex (Re 1600, Mach 0.1) on
    the periodic cube [0, 2 pi]^3 at n^3 (3 ghost layers), WENO5-AO in
    characteristic variables, 2 x 2 Gauss points, GKS3D (gks2nd,
    Navier-Stokes collision time), S2O4, CFL 0.5."""
    from gks_amd.params import BC_PERIODIC, tgv_params
    from gks_amd.problems import tgv_convar
    from gks_amd.solver3d import Solver3D

    n = args.n3d
    p = tgv_params(n, fast_math=1 if args.variant == "fast" else 0)
    s = Solver3D(p)
    s.init_convar(tgv_convar(p, quad=2))
    bc = [BC_PERIODIC] * 6
    s.run(bc, 100.0, args.warmup)
    s.sync()
    s.timing(True)
    t0 = time.perf_counter()
    k = s.run(bc, 100.0, args.steps)
    s.sync()
    wall = time.perf_counter() - t0
    tim = s.get_timing()
    s.close()
    value = n ** 3 * k / wall / 1e6
    dom = max(tim, key=lambda ph: tim[ph][0])
    avg_ms = tim[dom][0] / max(tim[dom][1], 1)
    u = units3(dom, n)
    roof = None
    fpu = 4 * FLOPS3["point"] + 4 * FLOPS3["flux"] if dom == "flux" else FLOPS3.get(dom)
    if u and fpu:
        achieved = u * fpu / (avg_ms * 1e-3) / 1e12
        roof = {"bound": "fp64-valu", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP64_TFLOPS, 4),
                "traffic": TRAFFIC3_256_FAST.get(dom) if (n == 256 and args.variant == "fast") else None,
                "kernel": "k3_pflux (x, y, z): point values + GKS3D, one launch" if dom == "flux" else f"k3_{dom} (x, y, z)",
                "avg_ms": round(avg_ms, 4), "flops_per_launch": u * fpu}
        ex = EXECUTED_FLOPS3_256_FAST.get(dom) if (n == 256 and args.variant == "fast") else None
        if ex:
            roof["executed_flops_per_launch"] = ex
            roof["executed_frac"] = round(ex / (avg_ms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS, 4)
    out = {
        "metric": "Mcell-updates/s FP64, 3D Taylor-Green WENO5-AO GKS3D S2O4",
        "value": round(value, 3), "unit": "Mcell-updates/s", "n_gpus": 1, "steps": k, "warmup": args.warmup,
        "ms_per_step": round(wall / k * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"Taylor-Green vortex Re 1600 Mach 0.1, periodic [0,2pi]^3 at {n}^3, WENO5-AO "
                               f"characteristic, 2x2 Gauss points, GKS3D gks2nd Navier-Stokes, S2O4, CFL 0.5 "
                               f"({args.variant} kernels)", "n": n, "kernels": args.variant},
        "roofline": roof,
        "step_fp64_tflops": round(FLOPS3_CELL_STEP * n ** 3 * k / wall / 1e12, 3),
        "phase_ms_per_step": {ph: round(v[0] / k, 4) for ph, v in tim.items()},
    }
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline3(args.cpu_n3d, threads, args.cpu_steps)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def bench_tgv_dissipation(args):
    """The Taylor-Green vortex (Re 1600, Mach 0.1, periodic [0, 2 pi]^3) at
    n^3 to t_end by the plain path and by the learned path (the packaged
    troubled-interface classifier on every cell's x / y / z stencils):
    throughput of each (whole run, state in HBM; the kinetic-energy samples
    are taken between timed stretches), the kinetic energy E_k every
    --ek-every steps, the dissipation rate -dE_k/dt (central differences)
    and its largest difference from the plain run's relative to the plain
    peak rate."""
    import numpy as np

    from gks_amd import learn
    from gks_amd.params import BC_PERIODIC, tgv_params
    from gks_amd.problems import kinetic_energy, tgv_convar
    from gks_amd.solver3d import Solver3D

    def log(m):
        print(m, file=sys.stderr, flush=True)

    n, tend = args.n3d, args.t3d
    fast = 1 if args.variant == "fast" else 0
    P, thr, meta = learn.load(args.classifier or learn.DEFAULT_PARAMS)
    # the threshold calibration covers the 2D four-quadrant Riemann problems:
    # the 3D path takes the network's trained (validation) threshold
    thr = learn.trained_threshold(meta, thr)
    bc = [BC_PERIODIC] * 6
    legs = {}
    for tag, cls in (("plain", None), ("learned", (P, thr))):
        p = tgv_params(n, fast_math=fast)
        s = Solver3D(p)
        s.init_convar(tgv_convar(p, quad=2))
        if cls:
            s.set_classifier(*cls)
        ts, es = [0.0], [kinetic_energy(s.get_convar(), p)]
        wall, steps = 0.0, 0
        while s.t < tend:
            s.sync()
            t0 = time.perf_counter()
            k = s.run(bc, tend, args.ek_every)
            s.sync()
            wall += time.perf_counter() - t0
            steps += k
            if k == 0:
                break
            ts.append(s.t)
            es.append(kinetic_energy(s.get_convar(), p))
            if len(ts) % 20 == 0:
                log(f"tgv {tag} {n}^3: t={s.t:.3f} steps={steps} E_k={es[-1]:.6e}")
        legs[tag] = {"t": np.array(ts), "ek": np.array(es), "wall": wall, "steps": steps,
                     "counts": s.classifier_counts() if cls else None}
        s.close()
        log(f"tgv {tag}: {steps} steps in {wall:.1f} s, {n ** 3 * steps / wall / 1e6:.1f} Mcell-updates/s")
    lp, ll = legs["plain"], legs["learned"]
    eps_p = -np.gradient(lp["ek"], lp["t"])
    eps_l = np.interp(lp["t"], ll["t"], -np.gradient(ll["ek"], ll["t"]))
    rel = float(np.max(np.abs(eps_l - eps_p)) / np.max(eps_p))
    vp = n ** 3 * lp["steps"] / lp["wall"] / 1e6
    vl = n ** 3 * ll["steps"] / ll["wall"] / 1e6
    c = ll["counts"]
    out = {
        "metric": "Mcell-updates/s FP64, 3D Taylor-Green WENO5-AO GKS3D S2O4: learned vs plain, kinetic-energy "
                  "dissipation rate",
        "value": round(vl, 3), "unit": "Mcell-updates/s", "n_gpus": 1, "steps": ll["steps"], "warmup": 0,
        "window": f"whole runs from t=0 to t={tend} (kinetic energy sampled every {args.ek_every} steps between "
                  "timed stretches)",
        "ms_per_step": round(ll["wall"] / ll["steps"] * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"Taylor-Green vortex Re 1600 Mach 0.1, periodic [0,2pi]^3 at {n}^3 to t={tend}, "
                               f"WENO5-AO characteristic, 2x2 Gauss points, GKS3D gks2nd Navier-Stokes, S2O4, CFL "
                               f"0.5; learned path (packaged classifier, BF16 MFMA inference) against plain "
                               f"({args.variant} kernels)", "n": n, "kernels": args.variant},
        "plain": {"value": round(vp, 3), "steps": lp["steps"], "wall_s": round(lp["wall"], 3)},
        "speedup_learned_vs_plain": round(vl / vp, 4),
        "linear_share": {k: round(v[0] / max(v[1], 1), 4) for k, v in c.items()},
        "dissipation": {"max_rel_diff": round(rel, 6),
                        "peak_plain": {"t": round(float(lp["t"][np.argmax(eps_p)]), 4), "eps": float(np.max(eps_p))},
                        "peak_learned": {"t": round(float(lp["t"][np.argmax(eps_l)]), 4),
                                         "eps": float(np.max(eps_l))},
                        "t": [round(float(x), 5) for x in lp["t"]],
                        "eps_plain": [float(x) for x in eps_p], "eps_learned": [float(x) for x in eps_l]},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_window(args):
    """The K-step window (--window): --steps timed steps after --warmup
    untimed ones, from t=0, of one path (--path)."""
    import numpy as np  # noqa: F401

    import gks_amd
    from gks_amd.params import BC_FREE, planar_shock_params
    from gks_amd.problems import riemann_prim, riemann_states

    n = args.n
    p = planar_shock_params(n, fast_math=1 if args.variant == "fast" else 0)
    prim = riemann_prim(p, riemann_states(args.seed))
    s = gks_amd.Solver2D(p)
    s.init_prim(prim)
    learned = None
    if args.path == "learned":
        from gks_amd import learn
        P, thr, _ = learn.load(args.classifier or learn.DEFAULT_PARAMS)
        learn.set_classifier(s, P, thr)
        learned = (learn, thr)
    bc = [BC_FREE] * 4
    s.run(bc, 0.25, args.warmup)
    # one CopyFluid_new_to_old outside the run loop (which needs none): the
    # HBM counters' calibration kernel, 4 n^2 doubles read and written
    s.copy_old()
    s.sync()
    # In the timed steps only the flux sweeps (the dominant kernels; on the
    # learned path the launches over listed cells and faces) carry timing
    # events: events around every launch cost the step about 4%.
    s.timing(True, phases=LEARNED_PHASES if learned else ["flux_x", "flux_y"])
    c0 = learned[0].classifier_counts(s) if learned else None
    t0 = time.perf_counter()
    k = s.run(bc, 0.25, 2**31 - 1 if args.full_run else args.steps)
    s.sync()
    t1 = time.perf_counter()
    tim = s.get_timing()
    c1 = learned[0].classifier_counts(s) if learned else None
    assert args.full_run or k == args.steps, (k, args.steps)
    wall = t1 - t0
    value = n * n * k / wall / 1e6
    # per-phase breakdown: a few more steps with every phase timed (not part
    # of the timed window)
    nb = 0 if args.full_run else min(5, args.steps)
    s.timing(True)
    nb = s.run(bc, 0.25, nb) if nb else 0
    s.sync()
    tim_all = s.get_timing()
    s.timing(False)

    roof_learned = None
    if learned:
        roof_learned = learned_rooflines({key: c1[key] - c0[key] for key in c1}, tim)
        tim = {}
    # dominant kernel (of the flux sweeps; the breakdown confirms it)
    dom = max(tim, key=lambda ph: tim[ph][0]) if tim else None
    # the WENO5 tangential reconstruction runs inside the flux launch
    # (unless GKS_TANFLUX=0 asks the library for separate launches)
    fused = p.weno_order == 5 and os.environ.get("GKS_TANFLUX", "1")[:1] != "0"
    units, fpu = kernel_units(dom, n, fused=fused) if dom else (None, None)
    avg_ms = tim[dom][0] / max(tim[dom][1], 1) if dom else None
    roof = None
    if units:
        achieved = units * fpu / (avg_ms * 1e-3) / 1e12
        traffic = (TRAFFIC_2048_FAST.get(dom) if (n == 2048 and args.variant == "fast" and fused and args.path == "plain")
                   else None)
        roof = {"bound": "fp64-valu", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": traffic,
                "kernel": dom + (" (tangential reconstruction + flux, one launch)" if fused and dom.startswith("flux") else ""),
                "avg_ms": round(avg_ms, 4), "flops_per_launch": units * fpu}
        ex = (EXECUTED_FLOPS_2048_FAST.get(dom) if (n == 2048 and args.variant == "fast" and fused and args.path == "plain")
              else None)
        if ex:
            roof["executed_flops_per_launch"] = ex
            roof["executed_frac"] = round(ex / (avg_ms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS, 4)
    total_dev_ms = sum(v[0] for v in tim_all.values())
    learned_info = None
    if learned:
        lin, tot = learned[0].classifier_stats(s)
        learned_info = {"linear_share": round(lin / max(tot, 1), 4), "logit_threshold": learned[1]}
    step_flops = FLOPS_PER_CELL_UPDATE * n * n
    out = {
        "metric": f"Mcell-updates/s FP64, 2D {n}² WENO5-AO GKS, {args.path} path, "
                  + ("whole run" if args.full_run else f"{k}-step window"),
        "value": round(value, 3),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": k,
        "warmup": args.warmup,
        "window": "whole run to t=0.25" if args.full_run else f"{k} steps after {args.warmup} warm-up steps from t=0",
        "ms_per_step": round(wall / k * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"four-quadrant Riemann {n}x{n} seed {args.seed}, WENO5-AO-DF characteristic, "
                               f"GKS2D gks2nd, S2O4, CFL 0.5, free BCs ({args.path} path, {args.variant} kernels)",
                   "n": n, "path": args.path, "kernels": args.variant},
        "roofline": roof,
        "step_fp64_tflops": round(step_flops * k / wall / 1e12, 3),
        "phase_ms_per_step": {ph: round(v[0] / max(nb, 1), 4) for ph, v in tim_all.items()} if nb else None,
        "device_ms_per_step": round(total_dev_ms / nb, 3) if nb else None,
        "phase_breakdown": f"{nb} further steps with every phase timed" if nb else None,
    }
    if learned_info:
        out["learned"] = learned_info
    if roof_learned:
        # the learned path's step has no single dominant kernel of fixed
        # work: its roofline is that of the listed-face flux kernel
        out["roofline"] = dict(roof_learned["k_tanflux_list"], kernel="k_tanflux_list (listed faces, x and y)")
        out["roofline_learned"] = roof_learned
    if args.equal_l1:
        from gks_amd.drivers import equal_l1
        out["equal_l1"] = equal_l1(n, args.seed, log=lambda m: print(m, file=sys.stderr, flush=True))
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, threads, args.cpu_steps)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


LEARNED_PHASES = ["normal_list_x", "normal_list_y", "flux_list_x", "flux_list_y"]


def learned_rooflines(c, lt):
    """The learned path's full-work kernels, each priced by the units it
    processed: k_tanflux_list per listed face 2 flux points + each side's
    tangential reconstruction (WENO5-AO-DF, or the linear weights);
    k_normal_list per listed cell one WENO5-AO-DF normal reconstruction.
    c: classifier counters over the timed steps, lt: HIP-event times of
    LEARNED_PHASES over the same steps."""
    nfaces = c["tangential_total"] // 2  # faces swept, x and y (one count per face side)
    listed = c["listed_faces_x"] + c["listed_faces_y"]
    cheap_faces = nfaces - listed
    lin_sides_listed = c["tangential_linear"] - 2 * cheap_faces
    weno_sides_listed = 2 * listed - lin_sides_listed
    f_list = (listed * 2 * FLOPS_FLUX_POINT + weno_sides_listed * FLOPS_TANGENT_SIDE
              + lin_sides_listed * FLOPS_TANGENT_SIDE_LINEAR)
    ms_list = lt["flux_list_x"][0] + lt["flux_list_y"][0]
    listed_cells = c["normal_total"] - c["normal_linear"]
    ms_nlist = lt["normal_list_x"][0] + lt["normal_list_y"][0]
    a_list = f_list / (ms_list * 1e-3) / 1e12
    a_nlist = listed_cells * FLOPS_NORMAL_CELL / (ms_nlist * 1e-3) / 1e12
    return {
        "k_tanflux_list": {"bound": "fp64-valu", "achieved": round(a_list, 3), "peak": PEAK_FP64_TFLOPS,
                           "unit": "TFLOP/s", "frac": round(a_list / PEAK_FP64_TFLOPS, 4), "traffic": None,
                           "listed_faces": listed, "weno_sides": weno_sides_listed, "linear_sides": lin_sides_listed,
                           "flops": f_list, "ms_total": round(ms_list, 3),
                           "launches": lt["flux_list_x"][1] + lt["flux_list_y"][1]},
        "k_normal_list": {"bound": "fp64-valu", "achieved": round(a_nlist, 3), "peak": PEAK_FP64_TFLOPS,
                          "unit": "TFLOP/s", "frac": round(a_nlist / PEAK_FP64_TFLOPS, 4), "traffic": None,
                          "listed_cells": listed_cells, "flops": listed_cells * FLOPS_NORMAL_CELL,
                          "ms_total": round(ms_nlist, 3),
                          "launches": lt["normal_list_x"][1] + lt["normal_list_y"][1]},
    }


def _leg(p, prim, bc, tstop, cls=None, phases=None, warmup=0, log=None, tag=""):
    """One whole run from the initial condition to tstop on the
    single-problem path (gks2d_run), with `warmup` untimed steps on a
    separate solver first; the timed region is the run itself (state in
    HBM, synchronised on both sides).  Returns steps, wall, per-phase
    event times, classifier counters and the final interior state."""
    import gks_amd
    from gks_amd import learn
    G = p.ghost
    if warmup:
        w = gks_amd.Solver2D(p)
        w.init_prim(prim)
        if cls:
            learn.set_classifier(w, cls[0], cls[1])
        w.run(bc, tstop, warmup)
        w.sync()
        w.close()
    s = gks_amd.Solver2D(p)
    s.init_prim(prim)
    if cls:
        learn.set_classifier(s, cls[0], cls[1])
    if phases:
        s.timing(True, phases=phases)
    s.sync()
    t0 = time.perf_counter()
    k = s.run(bc, tstop)
    s.sync()
    wall = time.perf_counter() - t0
    out = {"steps": k, "wall": wall, "t": s.t, "timing": s.get_timing() if phases else None,
           "counts": learn.classifier_counts(s) if cls else None}
    out["u"] = s.get_convar()[:, G:G + p.nodex, G:G + p.nodey]
    s.close()
    if log:
        log(f"{tag}: {k} steps to t={out['t']:.4f} in {wall:.2f} s = {p.nodex * p.nodey * k / wall / 1e6:.1f} "
            "Mcell-updates/s")
    return out


def _leg_batched(p, prim, bc, tstop, cls, warmup=0, log=None, tag=""):
    """The same run through the batched executor (a batch of one problem:
    the device-side time loop, one captured HIP graph per S2O4 step).  The
    tile shape comes from the autotuner, run with the warm-up steps on a
    separate batch; the timed batch starts from the initial condition."""
    from gks_amd import learn
    from gks_amd.batch import Batch
    G = p.ghost
    w = Batch([p])
    w.problems[0].init_prim(prim)
    learn.set_classifier(w.problems[0], cls[0], cls[1])
    tf, lf, _ = w.autotune(bc, steps=2)
    if warmup:
        w.run(bc, tstop, warmup)
    w.close()
    b = Batch([p])
    s = b.problems[0]
    s.init_prim(prim)
    learn.set_classifier(s, cls[0], cls[1])
    b.set_shape(tf, lf)
    s.sync()
    t0 = time.perf_counter()
    k = int(b.run(bc, tstop)[0])
    s.sync()
    wall = time.perf_counter() - t0
    out = {"steps": k, "wall": wall, "t": s.t, "tanflux_faces": tf, "largest_first": lf,
           "u": s.get_convar()[:, G:G + p.nodex, G:G + p.nodey]}
    b.close()
    if log:
        log(f"{tag}: {k} steps to t={out['t']:.4f} in {wall:.2f} s = {p.nodex * p.nodey * k / wall / 1e6:.1f} "
            "Mcell-updates/s")
    return out


def bench_headline(args):
    """The metric: the four-quadrant Riemann problem at n^2 (seed args.seed)
    run from the initial condition to t = 0.25 by the plain path, the
    learned path and the learned path through the batched executor, each
    timed over its whole run; density L1 of each against a plain FP64 run
    at (2n)^2 averaged over 2 x 2 blocks.  `value` is the learned path's
    Mcell-updates/s."""
    import numpy as np

    from gks_amd import learn
    from gks_amd.params import BC_FREE, planar_shock_params
    from gks_amd.problems import riemann_prim, riemann_states

    def log(m):
        print(m, file=sys.stderr, flush=True)

    n, tstop = args.n, 0.25
    fast = 1 if args.variant == "fast" else 0
    p = planar_shock_params(n, fast_math=fast)
    prim = riemann_prim(p, riemann_states(args.seed))
    bc = [BC_FREE] * 4
    P, thr, _ = learn.load(args.classifier or learn.DEFAULT_PARAMS)
    legs = {}
    # plain: HIP events on the flux launches (the dominant kernel, k_tanflux)
    legs["plain"] = _leg(p, prim, bc, tstop, phases=["flux_x", "flux_y"], warmup=args.warmup, log=log, tag="plain")
    # learned: HIP events on the launches over the listed cells and faces
    legs["learned"] = _leg(p, prim, bc, tstop, cls=(P, thr), phases=LEARNED_PHASES, warmup=args.warmup, log=log,
                           tag="learned")
    if not args.no_batched:
        legs["learned_batched"] = _leg_batched(p, prim, bc, tstop, (P, thr), warmup=args.warmup, log=log,
                                               tag="learned+batched")
    ref = None
    if not args.no_l1:
        pr = planar_shock_params(2 * n, fast_math=fast)
        r = _leg(pr, riemann_prim(pr, riemann_states(args.seed)), bc, tstop, log=log, tag=f"plain {2 * n}^2")
        ref = r["u"].reshape(4, n, 2, n, 2).mean(axis=(2, 4))
        legs["reference_2n"] = r
    res = {}
    for tag, r in legs.items():
        m = (2 * n if tag == "reference_2n" else n)
        e = {"value": round(m * m * r["steps"] / r["wall"] / 1e6, 3), "steps": r["steps"],
             "wall_s": round(r["wall"], 3), "ms_per_step": round(r["wall"] / r["steps"] * 1e3, 4),
             "t_final": r["t"]}
        if ref is not None and tag != "reference_2n":
            e["l1"] = [float(x) for x in np.abs(r["u"] - ref).mean(axis=(1, 2))]
        res[tag] = e
    L = legs["learned"]
    c = L["counts"]
    nfaces = c["tangential_total"] // 2  # faces swept, x and y (one count per face side)
    listed = c["listed_faces_x"] + c["listed_faces_y"]
    res["learned"].update({
        "linear_share_normal": round(c["normal_linear"] / max(c["normal_total"], 1), 4),
        "linear_share_tangential": round(c["tangential_linear"] / max(c["tangential_total"], 1), 4),
        "listed_face_share": round(listed / max(nfaces, 1), 4),
        "logit_threshold": thr})
    plain_v = res["plain"]["value"]
    value = res["learned"]["value"]

    # roofline of the plain path's dominant kernel: k_tanflux<1> (y faces:
    # tangential WENO5-AO-DF of both sides + 2 flux points per face)
    tim = legs["plain"]["timing"]
    units, fpu = kernel_units("flux_y", n, fused=True)
    avg_ms = tim["flux_y"][0] / max(tim["flux_y"][1], 1)
    achieved = units * fpu / (avg_ms * 1e-3) / 1e12
    exact = n == 2048 and args.variant == "fast"
    roof = {"bound": "fp64-valu", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP64_TFLOPS, 4),
            "traffic": TRAFFIC_2048_FAST["flux_y"] if exact else None,
            "kernel": "k_tanflux<1> (plain leg: y faces, tangential reconstruction + 2 GKS2D flux points per face)",
            "avg_ms": round(avg_ms, 4), "launches": tim["flux_y"][1], "flops_per_launch": units * fpu}
    if exact:
        roof["executed_flops_per_launch"] = EXECUTED_FLOPS_2048_FAST["flux_y"]
        roof["executed_frac"] = round(EXECUTED_FLOPS_2048_FAST["flux_y"] / (avg_ms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS,
                                      4)
    # the learned leg's full-work kernels, each priced by its own units
    roof_learned = learned_rooflines(c, L["timing"])
    out = {
        "metric": "Mcell-updates/s FP64, 2D 2048\u00b2 WENO5-AO GKS: learned vs plain at equal L1",
        "value": value,
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": L["steps"],
        "warmup": args.warmup,
        "window": f"whole run from the initial condition to t={tstop} ({L['steps']} S2O4 steps of the learned path; "
                  f"each leg also ran {args.warmup} untimed warm-up steps on a separate solver first)",
        "ms_per_step": round(L["wall"] / L["steps"] * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"four-quadrant Riemann {n}x{n} seed {args.seed} to t={tstop}, WENO5-AO-DF "
                               f"characteristic, GKS2D gks2nd, S2O4, CFL 0.5, free BCs; learned path (packaged "
                               f"troubled-interface classifier, BF16 MFMA inference) against the plain path "
                               f"({args.variant} kernels)",
                   "n": n, "path": "learned", "kernels": args.variant},
        "plain": res["plain"],
        "learned": res["learned"],
        "learned_batched": res.get("learned_batched"),
        "speedup_learned_vs_plain": round(value / plain_v, 4),
        "speedup_learned_batched_vs_plain": (round(res["learned_batched"]["value"] / plain_v, 4)
                                             if "learned_batched" in res else None),
    }
    if ref is not None:
        out["l1_ratio_density"] = round(res["learned"]["l1"][0] / res["plain"]["l1"][0], 5)
        if "learned_batched" in res:
            out["l1_ratio_density_batched"] = round(res["learned_batched"]["l1"][0] / res["plain"]["l1"][0], 5)
        out["equal_l1"] = bool(out["l1_ratio_density"] <= 1.05)
        out["l1_reference"] = (f"plain FP64 run at {2 * n}^2 to t={tstop} ({res['reference_2n']['steps']} steps, "
                               f"{res['reference_2n']['value']} Mcell-updates/s), averaged over 2x2 blocks; "
                               "l1 = mean |difference| per conservative variable over the interior")
    out["roofline"] = roof
    out["roofline_learned"] = roof_learned
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, threads, args.cpu_steps, args.cpu_warm)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30,
                    help="timed steps of the --window measurement (the default measurement times whole runs)")
    ap.add_argument("--warmup", type=int, default=3, help="untimed warm-up steps before each timed leg")
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--variant", choices=["ieee", "fast"], default="fast")
    ap.add_argument("--cpu-n", type=int, default=2048)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--cpu-warm", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--path", choices=["plain", "learned"], default="plain",
                    help="learned: the trained troubled-interface classifier picks linear weights for smooth stencils")
    ap.add_argument("--classifier", default=None, help="trained classifier (.npz); default: the packaged one")
    ap.add_argument("--workload", choices=["single", "batch", "tgv3d", "tgv3d-dissipation"], default="single",
                    help="batch: --batch uneven problems stepped together, against running them one by one; "
                         "tgv3d: the 3D Taylor-Green cube at --n3d^3; tgv3d-dissipation: plain vs learned to --t3d "
                         "with the kinetic-energy dissipation rate")
    ap.add_argument("--n3d", type=int, default=256)
    ap.add_argument("--t3d", type=float, default=10.0, help="end time of --workload tgv3d-dissipation")
    ap.add_argument("--ek-every", type=int, default=50, help="steps between kinetic-energy samples")
    ap.add_argument("--cpu-n3d", type=int, default=48)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--window", action="store_true",
                    help="time --steps steps of one --path after --warmup steps from t=0, instead of the whole-run "
                         "learned-versus-plain measurement")
    ap.add_argument("--no-l1", action="store_true", help="skip the (2n)^2 reference run and the L1 comparison")
    ap.add_argument("--no-batched", action="store_true", help="skip the learned+batched leg")
    ap.add_argument("--full-run", action="store_true",
                    help="time the whole run to t=0.25 instead of a window of --steps steps (the learned path's "
                         "cost depends on how much of the flow is smooth, which changes along the run)")
    ap.add_argument("--equal-l1", action="store_true",
                    help="also run plain and learned to t=0.25 and a plain (2n)^2 reference, and report density L1")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print("bench.py: this build runs on one GPU; multi-GPU runs are not part of it "
              f"(asked for --gpus {args.gpus}, WORLD_SIZE {world})", file=sys.stderr)
        sys.exit(3)
    if args.workload == "batch":
        bench_batch(args)
        return
    if args.workload == "tgv3d":
        bench_tgv3d(args)
        return
    if args.workload == "tgv3d-dissipation":
        bench_tgv_dissipation(args)
        return
    if args.window:
        bench_window(args)
        return
    bench_headline(args)



if __name__ == "__main__":
    main()
