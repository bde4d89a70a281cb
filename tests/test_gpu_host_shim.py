"""The reference-shaped C++ driver (high-order-solver_amd/host/examples/
riemann2d.cpp: PlanarShock's structure, built against the shim headers)
runs on the GPU and matches the oracle on the same seeded problem."""
import os
import subprocess

import numpy as np
import pytest

from gks_amd.params import BC_FREE, planar_shock_params
from gks_amd.problems import riemann_prim, riemann_states
from oracle_lib import oracle

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "high-order-solver_amd", "bin", "riemann2d")
G = 5


@pytest.mark.parametrize("n,seed,steps", [(48, 4, 40), (64, 0, 1000000)])
def test_reference_style_driver_matches_oracle(tmp_path, n, seed, steps):
    out = tmp_path / "u.bin"
    print("driver", n, seed, steps, flush=True)
    r = subprocess.run([EXE, str(n), str(seed), "0.25", str(steps), str(out)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(out, dtype=np.float64)
    t, k = raw[0], int(raw[1])
    u = raw[2:].reshape(4, n, n)
    p = planar_shock_params(n)
    o = oracle().solver(p)
    o.set_prim(riemann_prim(p, riemann_states(seed)))
    rc, _ = o.run([BC_FREE] * 4, 0.25, steps)
    assert rc == 0
    st = o.get()
    assert k == st["step"] and abs(t - st["t"]) <= 1e-14
    l1 = np.abs(u - st["convar"][:, G:-G, G:-G]).sum(axis=(1, 2)) / (n * n)
    print("  steps", k, "L1", l1, flush=True)
    assert np.all(l1 <= 1e-9)


EXE3 = os.path.join(ROOT, "high-order-solver_amd", "bin", "cubic_tube")


@pytest.mark.parametrize("n,steps,periodic", [(12, 5, False), (10, 4, True)])
def test_reference_style_3d_driver_matches_oracle(tmp_path, n, steps, periodic):
    """cubic_tube.cpp: CubicTube's loop (3Dproblem.cpp:3-172) compiled against
    the shim, the cubic Sod problem, against the 3D oracle."""
    from gks_amd.params import BC_PERIODIC, cubic_params
    from gks_amd.problems import cubic_sod_prim, prim_to_convar_3d
    from oracle_lib import oracle3
    out = tmp_path / "u3.bin"
    print("3d driver", n, steps, periodic, flush=True)
    args = [EXE3, str(n), "1.0", str(steps), str(out)] + (["periodic"] if periodic else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(out, dtype=np.float64)
    t, k = raw[0], int(raw[1])
    u = raw[2:].reshape(5, n, n, n)
    p = cubic_params(n)
    prim = cubic_sod_prim(p)
    bc = [BC_PERIODIC if periodic else BC_FREE] * 6
    ou, ot, on, rc, _ = oracle3().run(p, prim_to_convar_3d(prim, 1.4), prim, bc, 1.0, steps)
    assert rc == 0 and k == on == steps and abs(t - ot) <= 1e-13 * ot
    b = ou[:, 3:-3, 3:-3, 3:-3]
    err = np.max(np.abs(u - b)) / np.max(np.abs(b))
    print("  t", t, "max rel diff", err, flush=True)
    assert err <= 1e-11


EXE1 = os.path.join(ROOT, "high-order-solver_amd", "bin", "sod1d")


@pytest.mark.parametrize("n,steps,df", [(100, 1000000, False), (200, 60, True)])
def test_reference_style_1d_driver_matches_oracle(tmp_path, n, steps, df):
    """sod1d.cpp: SodTubeProblem's loop (1Dproblem.cpp:3-148) compiled against
    the shim, against the 1D oracle (the DF variant switches to
    WENO5_AO_with_DF from step 1, as the driver does)."""
    from gks_amd.params import sod_params
    from gks_amd.problems import sod_prim
    from oracle_lib import oracle1d
    out = tmp_path / "u1.bin"
    print("1d driver", n, steps, df, flush=True)
    args = [EXE1, str(n), "0.2", str(steps), str(out)] + (["df"] if df else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(out, dtype=np.float64)
    t, k = raw[0], int(raw[1])
    u = raw[2:2 + 3 * n].reshape(3, n)
    alpha = raw[2 + 3 * n:]
    p = sod_params(n, use_df=1 if df else 0)
    o = oracle1d().solver(p)
    o.set_prim(sod_prim(p))
    rc, _ = o.run([BC_FREE, BC_FREE], 0.2, steps)
    assert rc == 0
    st = o.get()
    assert k == st["step"] and abs(t - st["t"]) <= 1e-14
    l1 = np.abs(u - st["convar"][:, G:-G]).sum(axis=1) / n
    print("  steps", k, "L1", l1, flush=True)
    assert np.all(l1 <= 1e-10)  # device erfc / exp against the CPU's over ~70 steps: ~1e-12
    if df:
        assert np.max(np.abs(alpha - st["alpha"][G:-G])) <= 1e-9


EXE_V = os.path.join(ROOT, "high-order-solver_amd", "bin", "vortex2d")


@pytest.mark.parametrize("n,order,steps,fast", [(48, 7, 60, False), (32, 9, 40, True)])
def test_reference_style_vortex_driver_matches_oracle(tmp_path, n, order, steps, fast):
    """vortex2d.cpp: the smooth-flow loop written the reference's way with
    WENO7_AO_with_df_normal/_tangent (gausspoint 3) or WENO9 (gausspoint 4)
    assigned to the scheme pointers, periodic boundaries, the kernel build
    chosen through gks_fast_kernels; the isentropic vortex's cell averages,
    against the oracle on the same inputs (L1 1e-9, the same steps and t)."""
    from gks_amd.params import BC_PERIODIC, vortex_params
    from gks_amd.problems import vortex_convar
    p = vortex_params(n, weno_order=order)
    u0 = vortex_convar(p)
    src = tmp_path / "u0.bin"
    np.ascontiguousarray(u0[:, G:-G, G:-G]).tofile(src)
    out = tmp_path / "uv.bin"
    print("vortex driver", n, order, steps, fast, flush=True)
    args = [EXE_V, str(n), str(order), "10.0", str(steps), str(src), str(out)] + (["fast"] if fast else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(out, dtype=np.float64)
    t, k = raw[0], int(raw[1])
    u = raw[2:].reshape(4, n, n)
    o = oracle().solver(p)
    o.init_convar(u0)
    rc, _ = o.run([BC_PERIODIC] * 4, 10.0, steps)
    assert rc == 0
    st = o.get()
    assert k == st["step"] == steps and abs(t - st["t"]) <= 1e-13
    dx = 10.0 / n
    l1 = np.abs(u - st["convar"][:, G:-G, G:-G]).sum(axis=(1, 2)) * dx * dx
    print("  steps", k, "L1", l1, flush=True)
    assert np.all(l1 <= 1e-9)
