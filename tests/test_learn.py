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
