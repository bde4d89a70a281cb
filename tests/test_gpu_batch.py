"""Batched execution (include/gks2d_batch.h) on the GPU: a batched problem's
state, time and step count equal those of the same problem run alone with
gks2d_run (bit for bit: same kernels, same dt arithmetic on the device), for
uneven grids, both builds, per-problem tstop and step limits, after the
autotuner, with the learned classifier, and when one problem blows up."""
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
        assert steps[k] == n and b.problems[k].t == t
        assert np.array_equal(ub, u)
    b.close()


def test_batch_tstop_and_step_limits_per_problem():
    params = [planar_shock_params(n, fast_math=1) for n in SIZES]
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
    params = [planar_shock_params(n, fast_math=1) for n in SIZES]
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
    params = [planar_shock_params(n, fast_math=1) for n in SIZES]
    b = Batch(params)
    b.set_shape(*shape)
    for k, s in enumerate(b.problems):
        setup(s, 30 + k)
    b.run([BC_FREE] * 4, 0.03)
    for k, p in enumerate(params):
        u, t, n, _ = single(p, 30 + k, [BC_FREE] * 4, 0.03)
        assert b.problems[k].step_count == n and np.array_equal(b.problems[k].get_convar(), u)
    b.close()


def test_batch_with_the_classifier_equals_single_learned_runs():
    """IEEE build (the batched and the single-problem learned paths launch
    different kernels; only the IEEE build rounds alike in any kernel)."""
    rng = np.random.default_rng(3)
    P = rng.normal(0, 0.4, learn.NPARAM).astype(np.float32)
    thr = 0.0
    params = [planar_shock_params(n, fast_math=0) for n in SIZES]
    b = Batch(params)
    for k, s in enumerate(b.problems):
        setup(s, 40 + k)
        learn.set_classifier(s, P, thr)
    b.run([BC_FREE] * 4, 0.03)
    lin = [learn.classifier_stats(s) for s in b.problems]
    say("[batch] learned: linear / total per problem", lin)
    for k, p in enumerate(params):
        u, t, n, _ = single(p, 40 + k, [BC_FREE] * 4, 0.03, classifier=(P, thr))
        assert b.problems[k].step_count == n and np.array_equal(b.problems[k].get_convar(), u)
    assert sum(a for a, _ in lin) > 0
    b.close()


def test_a_problem_that_blows_up_stops_alone():
    params = [planar_shock_params(n, fast_math=1) for n in SIZES]
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
