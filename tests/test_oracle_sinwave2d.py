"""The 2D accuracy test (reference sinwave_2d / accuracy_sinwave_2d,
src/2Dproblem.cpp:1323-1603) on the CPU: the oracle against the compiled
reference's golden run (tests/golden/sinwave2d.npz, written by
tests/golden/make_golden.py), the initial condition against the mathematics
of the cell averages, and the error table format of output_error_form."""
import os

import numpy as np
import pytest

from gks_amd.drivers import convergence_orders, sinwave2d_errors, write_error_form
from gks_amd.params import BC_PERIODIC, sinwave2d_params
from gks_amd.problems import sinwave2d_prim, sinwave2d_rho
from oracle_lib import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sinwave2d.npz")
G = 5


def golden_errors(g, n):
    p = sinwave2d_params(n)
    u = np.zeros((4, p.nx, p.ny))
    u[0, G:-G, G:-G] = g[f"n{n}_rho"]
    return sinwave2d_errors(u, p, float(g[f"n{n}_t"]), 2.0)


def test_initial_cell_averages():
    """The cell averages of 1 + 0.2 sin(pi (x + y)): their mean over the
    periodic [0, 2]^2 is 1, they differ from the point values at the cell
    centres by the exact factor (sin(h)/h)^2 with h = pi dx / 2 (dx = dy),
    and from a 6-point Gauss quadrature of the cell by rounding only."""
    p = sinwave2d_params(20)
    rho = sinwave2d_rho(p)[G:-G, G:-G]
    assert abs(rho.mean() - 1.0) < 1e-14
    dx = 2.0 / 20
    xc = (np.arange(20) + 0.5) * dx
    pt = 0.2 * np.sin(np.pi * (xc[:, None] + xc[None, :]))
    h = np.pi * dx / 2
    np.testing.assert_allclose(rho - 1.0, pt * (np.sin(h) / h) ** 2, atol=1e-13)
    xg, wg = np.polynomial.legendre.leggauss(6)
    xq = (np.arange(20)[:, None] + 0.5 + 0.5 * xg[None, :]) * dx
    q = np.einsum("a,b,iajb->ij", wg / 2, wg / 2,
                  1 + 0.2 * np.sin(np.pi * (xq[:, :, None, None] + xq[None, None, :, :])))
    np.testing.assert_allclose(rho, q, atol=1e-13)


def test_oracle_matches_the_reference_golden_run_at_10():
    g = np.load(GOLD)
    p = sinwave2d_params(10)
    o = oracle().solver(p)
    o.set_prim(sinwave2d_prim(p))
    rc, hist = o.run([BC_PERIODIC] * 4, 2.0, 10**6)
    assert rc == 0
    st = o.get()
    assert st["step"] == int(g["n10_steps"]) and st["t"] == float(g["n10_t"])
    np.testing.assert_array_equal(hist[:st["step"]], g["n10_dt"])
    np.testing.assert_array_equal(st["convar"][0, G:-G, G:-G], g["n10_rho"])


def test_reference_converges_at_ninth_order():
    """The golden runs' density errors fall by 2^8.7 and more per mesh
    halving (WENO9-AO with 4 Gauss points and S2O4 at CFL 0.05)."""
    g = np.load(GOLD)
    errs = [golden_errors(g, n) for n in (10, 20, 40)]
    orders = np.array(convergence_orders(errs))
    assert np.all(orders >= 8.5), orders
    assert errs[2][0] < 1e-10


def test_error_table_format(tmp_path):
    """output_error_form's layout (output.cpp:17-61)."""
    path = write_error_form(str(tmp_path / "error.txt"), 0.05, 1.0, [10, 20],
                            [[1.5e-5, 2e-5, 3e-5], [3e-8, 4e-8, 6e-8]])
    lines = open(path).read().splitlines()
    assert lines[0] == "the CFL number is 0.05" and lines[1] == "the dt ratio over dx is 1"
    assert lines[2] == "1/10 & 1.500000e-05 & ~ & 2.000000e-05 & ~ & 3.000000e-05 & ~ \\\\ "
    assert lines[3] == "1/20 & 3.000000e-08 & 8.97 & 4.000000e-08 & 8.97 & 6.000000e-08 & 8.97 \\\\ "
