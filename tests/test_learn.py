"""The classifier's CPU oracle (oracle/mlp_oracle.py) and C ABI, without a
GPU: back-propagation against finite differences (this pins the oracle that
checks the GPU's training kernels), the BF16 rounding, the features'
invariances, and the exported symbols of include/gks_learn.h."""
import os
import re
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as mo  # noqa: E402


def rand_params(seed=0):
    rng = np.random.default_rng(seed)
    return rng.normal(0, 0.4, mo.NPARAM)


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(1)
    P = rand_params(1)
    X = rng.normal(0, 1, (40, 16))
    X[:, 12:] = 0
    y = (rng.uniform(size=40) > 0.5).astype(float)
    w = rng.uniform(0.5, 2.0, 40)
    _, g = mo.loss_grad(P, X, y, w)
    h = 1e-6
    for k in list(range(0, 512, 37)) + list(range(512, 544, 5)) + list(range(544, 1568, 97)) + \
            list(range(1568, 1633, 11)) + [1632]:
        e = np.zeros_like(P)
        e[k] = h
        fd = (mo.loss_grad(P + e, X, y, w)[0] - mo.loss_grad(P - e, X, y, w)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 0.0, 3.0e38], dtype=np.float32)
    r = mo.bf16(x)
    assert r[0] == 1.0 and r[3] == -2.5 and r[4] == 0.0
    assert r[1] == 1.0                   # exactly half an ulp: ties to the even neighbour
    assert r[2] == np.float32(1.0 + 2 ** -7)  # above half: rounds up
    assert np.all(np.abs(r - x) <= np.abs(x) * 2 ** -8)


def test_features_are_shift_and_scale_invariant():
    rng = np.random.default_rng(3)
    q = rng.uniform(0.5, 2.0, (50, 2, 5))
    f = mo.features(q)
    f2 = mo.features(q * 7.0)  # the shape features and relative spreads ignore scale
    assert np.allclose(f, f2, atol=1e-6)
    assert np.all(np.abs(f[:, :10]) <= 1.0) and np.all(f[:, 12:] == 0)
    flat = mo.features(np.ones((1, 2, 5)))
    assert np.all(flat[0, :10] == 0) and np.allclose(flat[0, 10:12], -1.0, atol=1e-6)


def test_learn_abi_exports():
    import gks_amd
    from gks_amd import learn
    text = open(os.path.join(ROOT, "include", "gks_learn.h")).read()
    names = sorted(set(re.findall(r"\b((?:gks2d|gks_mlp)_[a-z_0-9]+)\s*\(", text)))
    assert len(names) >= 7
    lib = gks_amd.lib()
    assert not [n for n in names if not hasattr(lib, n)]
    assert set(names) <= set(learn._SIGS)
    assert learn.NPARAM == mo.NPARAM


def test_threshold_calibration_rule(monkeypatch):
    """gks_amd.learn.calibrate on the host logic alone (the solver runs
    replaced by a stand-in whose learned error grows with the threshold):
    the largest threshold whose worst per-variable L1 ratio over the
    calibration problems stays within the margin and that blows up on none
    of them; problems whose plain run or reference blows up are left out."""
    sys.path.insert(0, os.path.join(ROOT, "high-order-solver_amd"))
    from gks_amd import learn

    base = np.ones((4, 8, 8))

    def fake_run(p, prim, tstop, cls=None):
        seed = int(prim)
        if p.nodex == 8 and seed == 3:  # the 2n reference of seed 3 blows up
            return None, None
        if p.nodex == 16:  # the reference: zero error against itself
            return np.zeros((4, 16, 16)), None
        if cls is None:
            return base, None
        t = cls[1]
        if seed == 1 and t >= 1.0:  # blows up only on the learned path
            return None, None
        # error ratio 1 + 0.01 * (t + 2) on seed 0, 1 + 0.005 * (t + 2) on the others
        k = 0.01 if seed == 0 else 0.005
        return base * (1 + k * (t + 2)), (9, 10)

    monkeypatch.setattr(learn, "_run_convar", fake_run)
    monkeypatch.setattr(learn, "riemann_prim", lambda p, st: st)
    monkeypatch.setattr(learn, "riemann_states", lambda seed: seed)
    grid = [-2.0, -1.5, -1.0, -0.5, 0.0, 0.5, 1.0]
    # with n = 8 the calibration runs at 8 and its references at 16
    thr, table = learn.calibrate(None, grid, n=8, seeds=range(4), margin=1.02)
    rows = {r["threshold"]: r for r in table}
    assert thr == 0.0  # seed 0: 1 + 0.01 * 2 = 1.02 at t = 0, above it at 0.5
    assert rows[0.5]["worst_l1_ratio"] > 1.02 and rows[0.0]["worst_l1_ratio"] <= 1.02
    assert rows[1.0]["blowups"] == 1 and rows[0.0]["blowups"] == 0
    # the reference blow-up of seed 3 removed it: three problems contribute
    assert rows[-2.0]["linear_share"] == 0.9
    # no threshold meets a margin below every ratio
    thr2, _ = learn.calibrate(None, grid, n=8, seeds=range(4), margin=0.999)
    assert thr2 is None
