"""C-ABI checks that need no GPU: the library loads, exports every function
include/gks2d.h declares, the ctypes mirror of gks2d_params matches the C
layout, and argument validation happens on the host before any device call."""
import ctypes
import os
import re

import numpy as np
import pytest

import gks_amd
from gks_amd.params import Params, planar_shock_params
from gks_amd.problems import MT19937_64, riemann_states
from oracle_lib import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gks2d.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(gks2d_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_abi():
    names = declared_functions()
    assert "gks2d_create" in names and "gks2d_run" in names and len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = gks_amd.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python wrapper binds every one of them
    assert set(declared_functions()) <= set(gks_amd._SIGS)


def test_params_layout_matches_c():
    off = (ctypes.c_int64 * 32)()
    f = oracle().lib.orc2d_params_layout
    f.restype = ctypes.c_int64
    f.argtypes = [ctypes.POINTER(ctypes.c_int64)]
    size = f(off)
    assert size == ctypes.sizeof(Params)
    for k, (name, _) in enumerate(Params._fields_):
        assert off[k] == getattr(Params, name).offset, name


@pytest.mark.parametrize("bad,code", [
    (dict(ghost=3), 2), (dict(weno_order=6), 2), (dict(nodex=2), 1), (dict(right=-1.0), 1),
    (dict(use_df=0), 2), (dict(recon_variable=7), 1),
])
def test_create_rejects_bad_params_on_host(bad, code):
    p = planar_shock_params(16).copy(**bad)
    h = ctypes.c_void_p()
    rc = gks_amd.lib().gks2d_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == code
    assert not h.value
    assert gks_amd.lib().gks2d_last_error()


def test_null_context_is_an_argument_error():
    L = gks_amd.lib()
    assert L.gks2d_copy_old(None) == 1
    assert L.gks2d_primvar(None) == 1
    assert L.gks2d_destroy(None) == 0


def test_mt19937_64_matches_standard_first_outputs():
    # the C++ standard fixes the 10000th output of mt19937_64 with default seed
    g = MT19937_64(5489)
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042


def test_riemann_states_in_range_and_seeded():
    a = riemann_states(0)
    b = riemann_states(0)
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(a, riemann_states(1))
    assert np.all((a[:, 0] >= 0.1) & (a[:, 0] <= 1.5))
    assert np.all(np.abs(a[:, 1:3]) <= 0.8)
    assert np.all((a[:, 3] >= 0.02) & (a[:, 3] <= 1.5))


# ------------------------------------------------------------------------ 1D
HEADER1 = os.path.join(ROOT, "include", "gks1d.h")


def declared_functions_1d():
    text = open(HEADER1).read()
    return sorted(set(re.findall(r"\b(gks1d_[a-z_0-9]+)\s*\(", text)))


def test_1d_library_exports_every_declared_symbol():
    from gks_amd import solver1d
    names = declared_functions_1d()
    assert "gks1d_run" in names and len(names) >= 8
    lib = solver1d.lib1()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) <= set(solver1d.SIGS1)


def test_1d_params_layout_matches_c():
    from gks_amd.params import Params1D
    off = (ctypes.c_int64 * 32)()
    f = oracle().lib.orc1d_params_layout
    f.restype = ctypes.c_int64
    f.argtypes = [ctypes.POINTER(ctypes.c_int64)]
    size = f(off)
    assert size == ctypes.sizeof(Params1D)
    for k, (name, _) in enumerate(Params1D._fields_):
        assert off[k] == getattr(Params1D, name).offset, name


@pytest.mark.parametrize("bad,code", [
    (dict(ghost=4), 2), (dict(nodex=3), 1), (dict(right=-1.0), 1), (dict(use_df=2), 1),
    (dict(recon_variable=3), 1), (dict(cfl=0.0), 1), (dict(tau_type=5), 1),
])
def test_1d_create_rejects_bad_params_on_host(bad, code):
    from gks_amd import solver1d
    from gks_amd.params import sod_params
    p = sod_params(16).copy(**bad)
    h = ctypes.c_void_p()
    rc = solver1d.lib1().gks1d_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == code
    assert not h.value
    assert solver1d.lib1().gks1d_last_error()


def test_1d_null_arguments_are_argument_errors():
    from gks_amd import solver1d
    L = solver1d.lib1()
    bc = (ctypes.c_int32 * 2)(0, 0)
    t = ctypes.c_double(0.0)
    s = ctypes.c_int32(0)
    assert L.gks1d_run(None, bc, 1.0, 1, ctypes.byref(t), ctypes.byref(s), None) == 1
    assert L.gks1d_get_face_fluxes(None, 0, None) == 1
    assert L.gks1d_destroy(None) == 0
    assert L.gks1d_flux_points(None, 0, None, None, 0.0, None, None, None) == 1
    assert L.gks1d_boundary(None, 0, 0) == 1
    assert L.gks1d_copy_old(None) == 1
    assert L.gks1d_cfl(None, 0.0, 1.0, None) == 1
    assert L.gks1d_reconstruct(None, 0) == 1
    assert L.gks1d_flux(None, 0, 0.1) == 1
    assert L.gks1d_update(None, 0) == 1
    assert L.gks1d_update_alpha(None) == 1


def test_1d_noalloc_checks_arguments_on_host():
    """The allocation-free run rejects bad parameters, null pointers and
    tubes beyond its scratch limit (16384 interior cells) before any device
    call (the fake state pointer is never dereferenced)."""
    from gks_amd import solver1d
    from gks_amd.params import sod_params
    L = solver1d.lib1()
    bc = (ctypes.c_int32 * 2)(0, 0)
    t = ctypes.c_double(0.0)
    s = ctypes.c_int32(0)
    fake = ctypes.c_void_p(4096)
    args = (bc, 0.2, 10, ctypes.byref(t), ctypes.byref(s), None, None)
    assert L.gks1d_run_noalloc(None, fake, None, *args) == 1
    p = sod_params(400)
    assert L.gks1d_run_noalloc(ctypes.byref(p), None, None, *args) == 1
    big = sod_params(16385)
    assert L.gks1d_run_noalloc(ctypes.byref(big), fake, None, *args) == 1
    assert b"16384" in L.gks1d_last_error()
    bad = sod_params(400).copy(ghost=4)
    assert L.gks1d_run_noalloc(ctypes.byref(bad), fake, None, *args) == 2


HEADER_B = os.path.join(ROOT, "include", "gks2d_batch.h")


def test_batch_library_exports_every_declared_symbol():
    from gks_amd import batch
    names = sorted(set(re.findall(r"\b(gks2d_batch_[a-z_0-9]+)\s*\(", open(HEADER_B).read())))
    assert len(names) == 6
    lib = gks_amd.lib()
    assert not [n for n in names if not hasattr(lib, n)]
    assert set(names) <= set(batch._SIGS)


def test_batch_rejects_bad_arguments_on_host():
    from gks_amd import batch
    L = batch.L()
    h = ctypes.c_void_p()
    arr = (Params * 2)(planar_shock_params(16), planar_shock_params(16).copy(weno_order=7, gausspoint=3))
    assert L.gks2d_batch_create(arr, 2, ctypes.byref(h)) == 2 and not h.value
    arr = (Params * 2)(planar_shock_params(16), planar_shock_params(16, fast_math=1))
    assert L.gks2d_batch_create(arr, 2, ctypes.byref(h)) == 1 and not h.value
    assert L.gks2d_batch_create(arr, 0, ctypes.byref(h)) == 1
    assert L.gks2d_batch_set_shape(None, 64, 1) == 1
    assert L.gks2d_batch_run(None, None, None, 1, 1, None, None) == 1
    assert L.gks2d_batch_destroy(None) == 0


# ------------------------------------------------------------------------ 3D
HEADER3 = os.path.join(ROOT, "include", "gks3d.h")


def test_3d_library_exports_every_declared_symbol():
    from gks_amd import solver3d
    names = sorted(set(re.findall(r"\b(gks3d_[a-z_0-9]+)\s*\(", open(HEADER3).read())))
    assert "gks3d_run" in names and len(names) >= 9
    lib = solver3d.lib3()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) <= set(solver3d.SIGS3)


@pytest.mark.parametrize("bad,code", [
    (dict(ghost=2), 2), (dict(ghost=9), 2), (dict(nodex=0), 1), (dict(right=-1.0), 1), (dict(front=-2.0), 1),
    (dict(recon_variable=3), 1), (dict(cfl=0.0), 1), (dict(tau_type=5), 1), (dict(gamma=1.0), 1),
])
def test_3d_create_rejects_bad_params_on_host(bad, code):
    from gks_amd import solver3d
    from gks_amd.params import cubic_params
    p = cubic_params(8).copy(**bad)
    h = ctypes.c_void_p()
    rc = solver3d.lib3().gks3d_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == code
    assert not h.value
    assert solver3d.lib3().gks3d_last_error()


def test_3d_null_arguments_are_argument_errors():
    from gks_amd import solver3d
    L = solver3d.lib3()
    bc = (ctypes.c_int32 * 6)(0, 0, 0, 0, 0, 0)
    t = ctypes.c_double(0.0)
    s = ctypes.c_int32(0)
    assert L.gks3d_run(None, bc, 1.0, 1, ctypes.byref(t), ctypes.byref(s)) == 1
    assert L.gks3d_set_convar(None, None, 8, 8) == 1
    assert L.gks3d_get_convar(None, None, 8, 8) == 1
    assert L.gks3d_set_cfl_primvar(None, None, 8, 8) == 1
    assert L.gks3d_sync(None) == 1
    assert L.gks3d_boundary(None, 0, 0) == 1
    assert L.gks3d_primvar(None) == 1
    assert L.gks3d_copy_old(None) == 1
    assert L.gks3d_cfl(None, 0.0, 1.0, None) == 1
    assert L.gks3d_flux(None, 0, 0.1) == 1
    assert L.gks3d_update(None, 0) == 1
    assert L.gks3d_destroy(None) == 0
    assert L.gks3d_flux_points(None, 0, None, None, 0.0, None, None, None) == 1
