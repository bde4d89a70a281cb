"""The isentropic-vortex workload on the oracle (no GPU): its initial
condition, the exact solution it is judged against, the oracle run's
accuracy and conservation, equality with the compiled reference's solver on
the same initial state, and the reference's text output formats."""
import os

import numpy as np
import pytest

from gks_amd.drivers import vortex_l1, write_output1d, write_output2d
from gks_amd.params import BC_PERIODIC, sod_params, vortex_params
from gks_amd.problems import vortex_convar, vortex_prim_point
from oracle_lib import have_ref, oracle, reference

G = 5


def test_vortex_point_values_are_isentropic_and_centred():
    x = np.linspace(2, 8, 41)[:, None] * np.ones((1, 41))
    y = x.T
    rho, u, v, p = vortex_prim_point(x, y)
    assert np.allclose(p, rho ** 1.4, rtol=1e-14)          # p = rho^gamma everywhere
    assert np.allclose(u[20, 20], 1.0) and np.allclose(v[20, 20], 1.0)  # no velocity jump at the centre
    assert rho.min() == rho[20, 20] and rho[20, 20] < 0.6  # density dip at the centre
    # the swirl is antisymmetric about the centre
    assert np.allclose(u - 1.0, -(u[::-1, ::-1] - 1.0), atol=1e-14)


def test_vortex_cell_averages_converge_with_quadrature():
    p = vortex_params(32)
    a = vortex_convar(p, quad=5)
    b = vortex_convar(p, quad=8)
    assert np.max(np.abs(a - b)) < 1e-7
    # the exact solution returns after one period of the (1, 1) advection
    assert np.allclose(vortex_convar(p, shift=(10.0, 10.0)), a, atol=1e-13)


def run(lib, p, u0, tstop):
    s = lib.solver(p)
    s.init_convar(u0)
    rc, hist = s.run([BC_PERIODIC] * 4, tstop, 100000)
    assert rc == 0
    return s.get()


def test_oracle_vortex_accuracy_and_conservation():
    p = vortex_params(32)
    u0 = vortex_convar(p)
    st = run(oracle(), p, u0, 10.0)
    e = vortex_l1(st["convar"], p, st["t"])
    assert e[0] < 5e-4, e  # WENO7 at 32^2 after one period (2.6e-4 measured)
    tot0 = u0[:, G:-G, G:-G].sum(axis=(1, 2))
    tot1 = st["convar"][:, G:-G, G:-G].sum(axis=(1, 2))
    assert np.allclose(tot1, tot0, rtol=1e-12)


@pytest.mark.skipif(not have_ref(), reason="compiled reference (oracle/_ref) not built")
def test_oracle_vortex_equals_compiled_reference():
    p = vortex_params(16)
    u0 = vortex_convar(p)
    a = run(oracle(), p, u0, 10.0)
    b = run(reference(), p, u0, 10.0)
    assert a["step"] == b["step"]
    assert np.array_equal(a["convar"], b["convar"])


def test_output_formats(tmp_path):
    p = sod_params(8)
    u = np.zeros((3, p.nx))
    u[0, G:G + 8] = np.arange(8) / 3.0
    write_output1d(os.path.join(tmp_path, "R.txt"), u, p)
    lines = open(os.path.join(tmp_path, "R.txt")).read().split()
    assert lines == ["0", "0.333333", "0.666667", "1", "1.33333", "1.66667", "2", "2.33333"]
    p2 = vortex_params(3)
    u2 = np.zeros((4, p2.nx, p2.ny))
    u2[0, G:G + 3, G:G + 3] = np.arange(9).reshape(3, 3)  # [i, j] = 3 i + j
    path = write_output2d(tmp_path, u2, p2, 17)
    assert os.path.basename(path) == "R-17.txt"
    vals = [float(x) for x in open(path).read().split()]
    assert vals == [0, 3, 6, 1, 4, 7, 2, 5, 8]  # x (i) fastest, as output2d writes
