"""The learned troubled-interface classifier on the GPU (include/gks_learn.h):
training kernels against the CPU oracle, the BF16 matrix-core inference
against the oracle's BF16 emulation (this also fixes the MFMA lane layout),
the solver-generated samples against host features, and the learned solver
against the plain one (identical where the classifier never fires; equal
L1 error where the trained network decides)."""
import os
import sys

import numpy as np
import pytest

import gks_amd
from gks_amd import learn
from gks_amd.params import BC_FREE, BC_PERIODIC, planar_shock_params, vortex_params
from gks_amd.problems import riemann_prim, riemann_states, vortex_convar

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as mo  # noqa: E402

pytestmark = pytest.mark.gpu
G = 5


def say(*a):
    print(*a, flush=True)


def rand_params(seed=0, scale=0.4):
    rng = np.random.default_rng(seed)
    return rng.normal(0, scale, mo.NPARAM).astype(np.float32)


def rand_features(n, seed=1):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1, 1, (n, 16)).astype(np.float32)
    X[:, 10:12] = rng.uniform(-1, 0, (n, 2))
    X[:, 12:] = 0
    return X


def test_loss_grad_matches_oracle():
    P = rand_params(2)
    X = rand_features(3000, 3)
    rng = np.random.default_rng(4)
    y = (rng.uniform(size=3000) > 0.6).astype(np.float32)
    w = rng.uniform(0.5, 2, 3000).astype(np.float32)
    say("[learn] loss/grad, 3000 samples")
    loss, g = learn.loss_grad(P, X, y, w)
    lo, go = mo.loss_grad(P.astype(np.float64), X.astype(np.float64), y, w)
    say("   loss", loss, lo, "max grad diff", np.abs(g - go).max(), "scale", np.abs(go).max())
    assert abs(loss - lo) <= 1e-5 * max(1.0, abs(lo))
    assert np.max(np.abs(g - go)) <= 1e-4 * np.max(np.abs(go))


def test_mfma_inference_matches_bf16_emulation():
    P = rand_params(5, 0.3)
    X = rand_features(10000, 6)
    say("[learn] BF16 MFMA inference, 10000 samples")
    z = learn.infer(P, X)
    ze = mo.forward_bf16(P, X)
    zf, _ = mo.forward(P, X)
    err = np.abs(z - ze) / (1.0 + np.abs(ze))
    say("   vs BF16 emulation: max", err.max(), "median", np.median(err),
        " vs FP64: max", (np.abs(z - zf) / (1 + np.abs(zf))).max())
    assert np.median(err) < 1e-5
    assert err.max() < 2e-2  # a hidden unit on a BF16 rounding boundary can flip one ulp
    assert np.max(np.abs(z - zf) / (1 + np.abs(zf))) < 0.1


def test_training_reduces_loss_and_is_reproducible():
    rng = np.random.default_rng(7)
    X = rand_features(20000, 8)
    y = ((X[:, 0] - X[:, 5] + 0.3 * X[:, 2]) > 0.1).astype(np.float32)
    P0 = learn.init_params(0)
    l0, _ = learn.loss_grad(P0, X, y)
    say("[learn] train 20000 samples x 8 epochs")
    P1, hist = learn.train(X, y, P=P0, epochs=8, batch=1024, lr=3e-3, seed=1)
    P2, _ = learn.train(X, y, P=P0, epochs=8, batch=1024, lr=3e-3, seed=1)
    l1, _ = learn.loss_grad(P1, X, y)
    acc = np.mean((learn.infer(P1, X) > 0) == (y > 0.5))
    say("   loss", l0, "->", l1, "per epoch", hist, "accuracy", acc)
    assert l1 < 0.3 * l0
    assert acc > 0.95
    assert np.array_equal(P1, P2)
    del rng


def test_samples_match_host_features():
    n = 64
    p = planar_shock_params(n)
    s = gks_amd.Solver2D(p)
    s.init_prim(riemann_prim(p, riemann_states(0)))
    s.run([BC_FREE] * 4, 0.25, 20)
    for side in range(4):
        s.boundary(side, BC_FREE)
    u = s.get_convar()
    r = u[0]
    gm1 = 1.4 - 1.0
    # Convar_to_primvar_2D in the IEEE build's operation order, x-like frame
    # of each sweep (the y sweep sees (rho, rho v, -rho u, E))
    prs = [gm1 * (u[3] - 0.5 * u[1] * u[1] / u[0] - 0.5 * u[2] * u[2] / u[0]),
           gm1 * (u[3] - 0.5 * u[2] * u[2] / u[0] - 0.5 * (-u[1]) * (-u[1]) / u[0])]
    say("[learn] solver samples vs host features")
    for direction in (0, 1):
        X, y = learn.samples(s, direction, 1, 1e-3)
        pr = prs[direction]
        if direction == 0:
            i = np.arange(G - 1, n + G + 1)[:, None] * np.ones((1, p.ny), int)
            j = np.ones((n + 2, 1), int) * np.arange(p.ny)[None, :]
            q = np.stack([np.stack([r[i + o, j] for o in range(-2, 3)], -1),
                          np.stack([pr[i + o, j] for o in range(-2, 3)], -1)], -2)
        else:
            i = np.arange(p.nx)[:, None] * np.ones((1, n + 2), int)
            j = np.ones((p.nx, 1), int) * np.arange(G - 1, n + G + 1)[None, :]
            q = np.stack([np.stack([r[i, j + o] for o in range(-2, 3)], -1),
                          np.stack([pr[i, j + o] for o in range(-2, 3)], -1)], -2)
        Xh = mo.features(q.reshape(-1, 2, 5))
        assert X.shape == Xh.shape
        d = np.abs(X - Xh)
        say("   dir", direction, "samples", len(y), "troubled", y.mean(), "max feature diff", d.max())
        assert np.max(d) < 1e-5
        assert 0.0 < y.mean() < 0.5
    s.close()


def test_classifier_that_never_fires_gives_the_plain_run():
    """IEEE build: its kernels round every operation as written, so the
    learned kernels' WENO-AO branch must reproduce the plain kernels bit for
    bit (the fast build contracts multiply-adds per kernel as the compiler
    sees fit, so two different kernels need not round alike there)."""
    n = 48
    p = planar_shock_params(n, fast_math=0)
    prim = riemann_prim(p, riemann_states(1))
    P = rand_params(9)
    P[1632] = 1e9  # logit bias: nothing is smooth
    out = []
    for use in (False, True):
        s = gks_amd.Solver2D(p)
        s.init_prim(prim)
        if use:
            learn.set_classifier(s, P, 0.0)
        s.run([BC_FREE] * 4, 0.25, 30)
        out.append(s.get_convar())
        if use:
            lin, tot = learn.classifier_stats(s)
            say("[learn] never-firing classifier: linear", lin, "of", tot)
            assert lin == 0 and tot > 0
        s.close()
    assert np.array_equal(out[0], out[1])


@pytest.mark.skipif(not os.path.exists(learn.DEFAULT_PARAMS), reason="trained classifier not present")
@pytest.mark.parametrize("n", [48, 37])
def test_one_classifier_launch_matches_the_per_direction_sweeps(n):
    """IEEE build (every operation rounded as written): the classifier
    launch that takes both directions at once (rho and p of each staged
    cell formed once, 16 x 16-cell tiles of the whole frame) calls exactly
    as many stencils linear as the two per-direction sweeps do (the path
    without the smooth-flow flux), over each sweep's whole range: (n + 2)
    rows of the x sweep, (n + 2) columns of the y sweep, ghost cells
    included.  37^2 leaves partly filled tiles at the frame's edges."""
    P, thr, _ = learn.load()
    p = planar_shock_params(n, fast_math=0)
    counts = []
    for smooth in (True, False):
        s = gks_amd.Solver2D(p)
        s.init_prim(riemann_prim(p, riemann_states(0)))
        s.run([BC_FREE] * 4, 0.25, 12)
        learn.set_classifier(s, P, thr)
        learn.classifier_options(s, smooth)
        for side in range(4):
            s.boundary(side, BC_FREE)
        s.reconstruct()
        c = learn.classifier_counts(s)
        counts.append((c["normal_linear"], c["normal_total"]))
        s.close()
    say(f"[learn] {n}^2 classifier counts (linear, total): one launch {counts[0]}, per direction {counts[1]}")
    assert counts[0] == counts[1]
    assert counts[0][1] == (n + 2) * p.ny + p.nx * (n + 2)
    assert 0 < counts[0][0] < counts[0][1]


@pytest.mark.parametrize("n", [64, 40])
def test_linear_everywhere_on_a_smooth_vortex(n):
    """A classifier that always fires: every stencil of the smooth vortex
    (DF sensor zero) takes the linear weights; the run stays conservative
    and at least as accurate as WENO5-AO there.  At 40^2 the smooth-face
    flux tiles (16 faces wide) leave partly filled tiles at the grid's
    edge in both directions."""
    from gks_amd.drivers import vortex_l1
    p = vortex_params(n, weno_order=5, fast_math=1)
    u0 = vortex_convar(p)
    P = rand_params(10)
    P[1632] = -1e9
    errs = []
    for use in (False, True):
        s = gks_amd.Solver2D(p)
        s.init_convar(u0)
        if use:
            learn.set_classifier(s, P, 0.0)
        s.run([BC_PERIODIC] * 4, 2.0)
        u = s.get_convar()
        errs.append(vortex_l1(u, p, s.t))
        if use:
            lin, tot = learn.classifier_stats(s)
            say("[learn] always-firing classifier on the vortex: linear", lin, "of", tot)
            assert lin == tot
            assert np.allclose(u[:, G:-G, G:-G].sum(axis=(1, 2)), u0[:, G:-G, G:-G].sum(axis=(1, 2)), rtol=1e-12)
        s.close()
    say("   L1 rho plain", errs[0][0], "linear", errs[1][0])
    assert errs[1][0] <= 1.05 * errs[0][0]


@pytest.mark.skipif(not os.path.exists(learn.DEFAULT_PARAMS), reason="trained classifier not present")
def test_trained_classifier_keeps_the_l1_error():
    """The trained network on an unseen seed at 512^2: the learned run's
    density L1 error against a 1024^2 plain run (averaged to 512^2) is within
    5% of the plain run's, and most stencils take the linear weights."""
    P, thr, meta = learn.load()
    seed, n = 5, 512  # training used seeds 100-131
    say(f"[learn] trained classifier, seed {seed}, {n}^2 vs 1024^2 plain, threshold {thr:.3f}")
    from gks_amd.drivers import run_riemann
    ref = run_riemann(2 * n, seed)["convar"][:, G:-G, G:-G]
    ref = ref.reshape(4, n, 2, n, 2).mean(axis=(2, 4))
    res = {}
    for use in (False, True):
        p = planar_shock_params(n, fast_math=1)
        s = gks_amd.Solver2D(p)
        s.init_prim(riemann_prim(p, riemann_states(seed)))
        if use:
            learn.set_classifier(s, P, thr)
        s.run([BC_FREE] * 4, 0.25)
        u = s.get_convar()[:, G:-G, G:-G]
        res[use] = np.abs(u - ref).mean(axis=(1, 2))
        if use:
            lin, tot = learn.classifier_stats(s)
            res["share"] = lin / tot
        s.close()
    say("   L1 plain", res[False], "learned", res[True], "linear share", res["share"])
    assert np.all(res[True] <= 1.05 * res[False])
    assert res["share"] > 0.5


@pytest.mark.parametrize("fast", [0, 1])
def test_smooth_flux_leaf_matches_oracle(fast):
    """GPU smooth-flow flux against the numpy restatement (pinned to GKS2D in
    test_smooth_flux.py), and against the GPU's own GKS2D when both sides
    agree."""
    import smooth_flux as sf
    from test_smooth_flux import states
    p = planar_shock_params(16, fast_math=fast).copy(tau_type=1, mu=2e-3, prandtl_fix=1, pr=0.72)
    L, R = states(3000, 5), states(3000, 6)
    R = L + 0.05 * (R - L)
    dt = 1e-3
    f, d = learn.flux_points_smooth(p, L, R, dt)
    fo, do = sf.smooth_flux(L, R, dt, K=p.K, tau_type=p.tau_type, mu=p.mu, nu=p.nu, pr=p.pr, prandtl_fix=1)
    ef = np.max(np.abs(f - fo) / np.abs(fo).max(axis=1, keepdims=True))
    ed = np.max(np.abs(d - do) / np.abs(do).max(axis=1, keepdims=True))
    _, fg, dg = gks_amd.flux_points(p, L, L, dt)
    fe, de = learn.flux_points_smooth(p, L, L, dt)
    eg = np.max(np.abs(fe - fg) / np.abs(fg).max(axis=1, keepdims=True))
    say(f"[learn] smooth flux, fast {fast}: vs numpy f {ef:.2e} derf {ed:.2e}; vs GKS2D at equal sides {eg:.2e}")
    assert ef < 1e-12 and ed < 1e-10 and eg < 1e-12


@pytest.mark.parametrize("fast", [0, 1])
def test_linear_weight_reconstructions_are_exact_for_quartics(fast):
    """gks2d_linear_weights_points (the linear-weight reconstructions of the
    learned 2D and 3D paths): on cell averages of random quartics the
    fifth-order candidate is exact, so the values and derivatives at the
    centre cell's faces (-h/2, +h/2) and Gauss points (-+h/(2 sqrt 3))
    equal the polynomial's; the fast build (compile-time combined
    coefficients, reciprocal multiplication) agrees with the IEEE one to
    1e-13 of the values' scale."""
    from gks_amd.learn import linear_points
    rng = np.random.default_rng(5)
    n, h = 4000, 0.37
    c = rng.normal(0, 1, (n, 5))  # P(x) = sum c_k x^k
    Pv = lambda x: sum(c[:, k] * x ** k for k in range(5))
    dP = lambda x: sum(k * c[:, k] * x ** (k - 1) for k in range(1, 5))
    Q = lambda x: sum(c[:, k] * x ** (k + 1) / (k + 1) for k in range(5))
    edges = (np.arange(6) - 2.5) * h
    W = np.stack([(Q(edges[s + 1]) - Q(edges[s])) / h for s in range(5)], axis=1)
    p = planar_shock_params(16, fast_math=fast)
    out = linear_points(p, W, h)
    g = h / (2 * np.sqrt(3.0))
    exact = np.stack([Pv(-h / 2), dP(-h / 2), Pv(h / 2), dP(h / 2), Pv(-g), dP(-g), Pv(g), dP(g)], axis=1)
    scale = np.max(np.abs(W), axis=1, keepdims=True) / h
    err = np.max(np.abs(out - exact) / scale)
    print(f"[linear weights] fast {fast}: max error / scale {err:.2e}", flush=True)
    assert err < 1e-12
    if fast:
        ref = linear_points(planar_shock_params(16, fast_math=0), W, h)
        assert np.max(np.abs(out - ref) / scale) < 1e-13

