"""Batched execution (include/gks2d_batch.h) on the GPU: a batched problem's
state, time and step count equal those of the same problem run alone with
gks2d_run, for uneven grids, per-problem tstop and step limits, after the
autotuner, with the learned classifier, and when one problem blows up; a
batch run to t = 0.2 against the oracle; and the learned batch's accuracy
against plain runs.  The batched kernels run the single-problem kernels'
device bodies with the same dt arithmetic, so in the IEEE build the runs
agree bit for bit.  The fast build contracts multiply-adds per kernel: a
body inlined into another kernel can round differently, so fast-build
comparisons across kernels take the same step count and a rounding
tolerance."""
import numpy as np
import pytest

import gks_amd
from gks_amd import GksError, learn
from gks_amd.batch import Batch
from gks_amd.params import BC_FREE, BC_PERIODIC, BC_REFLECT, planar_shock_params, vortex_params
from gks_amd.problems import riemann_prim, riemann_states, vortex_convar

pytestmark = pytest.mark.gpu
SIZES = [24, 40, 33, 48]


def say(*a):
    print(*a, flush=True)


def setup(s, seed):
    s.init_prim(riemann_prim(s.p, riemann_states(seed)))


def single(p, seed, bc, tstop, max_steps=2**31 - 1, classifier=None, init=None):
    s = gks_amd.Solver2D(p)
    (init or setup)(s, seed)
    if classifier is not None:
        learn.set_classifier(s, *classifier)
    n = s.run(bc, tstop, max_steps)
    out = (s.get_convar(), s.t, s.step_count, n)
    s.close()
    return out


def vortex_setup(s, seed):
    s.init_convar(vortex_convar(s.p))


@pytest.mark.parametrize("fast", [0, 1])
def test_batch_equals_single_runs(fast):
    """Three Riemann problems of uneven size (free, reflective and mixed
    boundaries) and an isentropic vortex on its own periodic [0,10]^2 domain
    in one batch."""
    params = [planar_shock_params(n, fast_math=fast) for n in SIZES[:3]] + [
        vortex_params(SIZES[3], weno_order=5, fast_math=fast)]
    bcs = [[BC_FREE] * 4, [BC_REFLECT] * 4, [BC_FREE, BC_REFLECT, BC_REFLECT, BC_FREE], [BC_PERIODIC] * 4]
    inits = [setup, setup, setup, vortex_setup]
    tstop = [0.06, 0.06, 0.06, 0.5]
    seeds = [0, 1, 3, 4]  # (seed 2's states blow up the scheme, on the CPU reference too)
    say(f"[batch] sizes {SIZES}, fast_math {fast}")
    b = Batch(params)
    for k, s in enumerate(b.problems):
        inits[k](s, seeds[k])
    steps = b.run(bcs, tstop)
    for k, p in enumerate(params):
        u, t, n, _ = single(p, seeds[k], bcs[k], tstop[k], init=inits[k])
        ub = b.problems[k].get_convar()
        say(f"   problem {k} ({SIZES[k]}^2): steps {steps[k]} vs {n}, t {b.problems[k].t} vs {t}")
        assert steps[k] == n
        if fast:
            err = np.max(np.abs(ub - u)) / np.max(np.abs(u))
            say(f"   max relative difference {err:.2e}")
            assert abs(b.problems[k].t - t) <= 1e-12 * t and err <= 1e-10
        else:
            assert b.problems[k].t == t and np.array_equal(ub, u)
    b.close()


def test_batch_tstop_and_step_limits_per_problem():
    params = [planar_shock_params(n, fast_math=0) for n in SIZES]
    tstop = [0.02, 0.05, 0.03, 0.04]
    bc = [BC_FREE] * 4
    b = Batch(params)
    for k, s in enumerate(b.problems):
        setup(s, 10 + k)
    first = b.run(bc, tstop, max_steps=3)
    assert np.all(first <= 3)
    b.run(bc, tstop)
    for k, p in enumerate(params):
        s = gks_amd.Solver2D(p)
        setup(s, 10 + k)
        s.run(bc, tstop[k], 3)
        s.run(bc, tstop[k])
        say(f"   problem {k}: t {b.problems[k].t} vs {s.t}, steps {b.problems[k].step_count} vs {s.step_count}")
        assert b.problems[k].t == s.t and b.problems[k].step_count == s.step_count
        assert np.array_equal(b.problems[k].get_convar(), s.get_convar())
        s.close()
    b.close()


def test_autotune_leaves_the_run_unchanged():
    params = [planar_shock_params(n, fast_math=0) for n in SIZES]
    bc = [BC_FREE] * 4
    out = []
    for tune in (False, True):
        b = Batch(params)
        for k, s in enumerate(b.problems):
            setup(s, 20 + k)
        if tune:
            tf, lf, ms = b.autotune(bc, steps=2)
            say(f"[batch] autotune: {tf} faces per flux workgroup, largest first {lf}, {ms:.3f} ms/step")
            assert tf in (64, 128) and ms > 0
        b.run(bc, 0.04)
        out.append([(s.get_convar(), s.t, s.step_count) for s in b.problems])
        b.close()
    for a, c in zip(*out):
        assert a[1] == c[1] and a[2] == c[2] and np.array_equal(a[0], c[0])


@pytest.mark.parametrize("shape", [(64, False), (128, True)])
def test_tile_shapes_give_the_same_states(shape):
    params = [planar_shock_params(n, fast_math=0) for n in SIZES]
    b = Batch(params)
    b.set_shape(*shape)
    for k, s in enumerate(b.problems):
        setup(s, 30 + k)
    b.run([BC_FREE] * 4, 0.03)
    for k, p in enumerate(params):
        u, t, n, _ = single(p, 30 + k, [BC_FREE] * 4, 0.03)
        assert b.problems[k].step_count == n and np.array_equal(b.problems[k].get_convar(), u)
    b.close()


@pytest.mark.parametrize("fast", [0, 1])
def test_batch_with_the_classifier_equals_single_learned_runs(fast):
    """The batched learned kernels (kb_normal, kb_normal_list,
    kb_tanflux_cheap, kb_tanflux_list) run the same device bodies as the
    single-problem ones in other kernels.  IEEE build: bit for bit.  Fast
    build: FMA contraction may round a body differently in another kernel,
    so the same step count and the state to 1e-9 relative."""
    rng = np.random.default_rng(3)
    P = rng.normal(0, 0.4, learn.NPARAM).astype(np.float32)
    thr = 0.0
    params = [planar_shock_params(n, fast_math=fast) for n in SIZES]
    b = Batch(params)
    for k, s in enumerate(b.problems):
        setup(s, 40 + k)
        learn.set_classifier(s, P, thr)
    b.run([BC_FREE] * 4, 0.03)
    lin = [learn.classifier_stats(s) for s in b.problems]
    say("[batch] learned: linear / total per problem", lin)
    for k, p in enumerate(params):
        u, t, n, _ = single(p, 40 + k, [BC_FREE] * 4, 0.03, classifier=(P, thr))
        ub = b.problems[k].get_convar()
        assert b.problems[k].step_count == n
        if fast:
            err = np.max(np.abs(ub - u)) / np.max(np.abs(u))
            say(f"   problem {k}: max relative difference {err:.2e}")
            assert err <= 1e-9
        else:
            assert np.array_equal(ub, u)
    assert sum(a for a, _ in lin) > 0
    b.close()


def test_a_problem_that_blows_up_stops_alone():
    params = [planar_shock_params(n, fast_math=0) for n in SIZES]
    b = Batch(params)
    for k, s in enumerate(b.problems):
        setup(s, 50 + k)
    # problem 1 starts from a state with negative density in one cell
    s1 = b.problems[1]
    u = s1.get_convar()
    u[0, 20, 20] = -1.0
    s1.set_convar(u)
    with pytest.raises(GksError) as e:
        b.run([BC_FREE] * 4, 0.03)
    say("[batch] blow-up:", e.value)
    assert e.value.code == 4 and "1" in str(e.value)
    assert s1.step_count == 0 and s1.t == 0.0
    for k in (0, 2, 3):
        u, t, n, _ = single(params[k], 50 + k, [BC_FREE] * 4, 0.03)
        assert b.problems[k].step_count == n and np.array_equal(b.problems[k].get_convar(), u)
    b.close()


ORACLE_SIZES = [24, 40, 33, 48, 64, 56]
ORACLE_BCS = [[BC_FREE] * 4, [BC_REFLECT] * 4, [BC_FREE, BC_REFLECT, BC_REFLECT, BC_FREE], [BC_PERIODIC] * 4,
              [BC_FREE] * 4, [BC_PERIODIC, BC_PERIODIC, BC_FREE, BC_FREE]]


@pytest.mark.parametrize("fast", [0, 1])
def test_batch_to_final_time_matches_the_oracle(fast):
    """Six uneven four-quadrant Riemann problems (24^2 to 64^2; free,
    reflective, mixed and periodic boundaries; seeds 60-65) in one batch to
    the batched workload's end time t = 0.2, each against the oracle run
    alone on the same inputs: the same step count, t to 1e-14 and the state
    to L1 1e-9 (the plain path's parity bar), in both kernel builds."""
    from oracle_lib import oracle
    from test_gpu_parity import L1_TOL, l1
    params = [planar_shock_params(n, fast_math=fast) for n in ORACLE_SIZES]
    say(f"[batch vs oracle] sizes {ORACLE_SIZES}, fast_math {fast}")
    b = Batch(params)
    for k, s in enumerate(b.problems):
        setup(s, 60 + k)
    steps = b.run(ORACLE_BCS, 0.2)
    G = 5
    for k, p in enumerate(params):
        o = oracle().solver(p)
        o.set_prim(riemann_prim(p, riemann_states(60 + k)))
        rc, _ = o.run(ORACLE_BCS[k], 0.2, 100000)
        assert rc == 0
        st = o.get()
        s = b.problems[k]
        e = l1(s.get_convar()[:, G:-G, G:-G], st["convar"][:, G:-G, G:-G], p)
        say(f"   problem {k} ({p.nodex}^2): steps {steps[k]} vs {st['step']}, t {s.t} vs {st['t']}, L1 {e}")
        assert steps[k] == st["step"] and abs(s.t - st["t"]) <= 1e-14
        assert np.all(e <= L1_TOL)
    b.close()


def test_learned_batch_density_error_against_plain_runs():
    """The batched workload's accuracy check on a small batch: five uneven
    problems (96^2 to 192^2, seeds 1000-1004 as the benchmark draws them)
    with the packaged classifier (BF16 MFMA inference), fast build, to
    t = 0.2.  Per problem, the density L1 of the learned batched run and of
    the plain run alone against a plain run at (2n)^2 averaged over 2x2
    blocks: the learned error stays within 5% of the plain one (the
    equal-L1 bar), and the two solutions differ by less than the plain
    error itself."""
    sizes = [96, 128, 160, 192, 128]
    P, thr, _ = learn.load()
    bc = [BC_FREE] * 4
    params = [planar_shock_params(n, fast_math=1) for n in sizes]
    b = Batch(params)
    for k, s in enumerate(b.problems):
        setup(s, 1000 + k)
        learn.set_classifier(s, P, thr)
    b.run(bc, 0.2)
    G = 5
    ratios = []
    for k, p in enumerate(params):
        n = p.nodex
        ul = b.problems[k].get_convar()[0, G:-G, G:-G]
        up = single(p, 1000 + k, bc, 0.2)[0][0, G:-G, G:-G]
        pr = planar_shock_params(2 * n, fast_math=1)
        ur = single(pr, 1000 + k, bc, 0.2)[0][0, G:-G, G:-G].reshape(n, 2, n, 2).mean(axis=(1, 3))
        el, ep = np.abs(ul - ur).mean(), np.abs(up - ur).mean()
        diff = np.abs(ul - up).mean()
        ratios.append(el / ep)
        say(f"   problem {k} ({n}^2): density L1 learned {el:.4e}, plain {ep:.4e}, ratio {el / ep:.4f}, "
            f"learned - plain {diff:.3e}")
        assert el <= 1.05 * ep
        assert diff < ep
    b.close()
