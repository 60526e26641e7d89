"""The C ABI library loads and exports every symbol include/dyphur_hip.h declares;
host-side entry points behave as specified.  No kernel is launched here."""
import ctypes as C
import os
import re

import numpy as np

from dyphur_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    txt = open(os.path.join(ROOT, "include", "dyphur_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(dy_\w+)\s*\(", txt, flags=re.M)))


def test_all_declared_symbols_exported(capi):
    names = _declared_functions()
    assert len(names) >= 45
    lib = capi.load()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # The ctypes binding covers every declared entry point.
    assert set(names) <= set(capi._SIGS)


def test_status_strings(capi):
    L = capi.load()
    assert L.dy_status_string(0) == b"success"
    assert b"unsupported" in L.dy_status_string(capi.DY_ERR_UNSUPPORTED)
    assert L.dy_status_string(-99) == b"unknown status"


def test_body_dynamics_matches_oracle(capi, oracle):
    sc = scenes.random_mixed(500, seed=7)
    inv, iI, flags = capi.body_dynamics(sc.mass, sc.inertia, sc.flags)
    ob = oracle.Bodies(sc)
    assert np.array_equal(inv.view(np.uint32), ob.arr["inv_mass"].view(np.uint32))
    for j, k in enumerate(("iI_xx", "iI_yy", "iI_zz", "iI_xy", "iI_xz", "iI_yz")):
        assert np.array_equal(iI[:, j].view(np.uint32), ob.arr[k].view(np.uint32)), k
    assert np.array_equal(flags, ob.flags)


def test_full_inertia_inverse(capi, oracle):
    rng = np.random.default_rng(3)
    n = 200
    A = rng.normal(size=(n, 3, 3))
    I = (A @ A.transpose(0, 2, 1) + 0.5 * np.eye(3)).astype(np.float32).reshape(n, 9)
    mass = rng.uniform(0.1, 5, n).astype(np.float32)
    inv, iI, _ = capi.body_dynamics(mass, I, np.zeros(n, np.uint32))
    sc = scenes.random_mixed(n - 1, seed=1)
    sc.mass[:] = mass
    sc.inertia[:] = I
    sc.flags[:] = 0
    ob = oracle.Bodies(sc)
    for j, k in enumerate(("iI_xx", "iI_yy", "iI_zz", "iI_xy", "iI_xz", "iI_yz")):
        assert np.array_equal(iI[:, j].view(np.uint32), ob.arr[k].view(np.uint32)), k
    # And it is the inverse (float64 check of the adjugate formula).
    full = np.linalg.inv(I.reshape(n, 3, 3).astype(np.float64))
    assert np.allclose(iI[:, 0], full[:, 0, 0], rtol=1e-3)
    assert np.allclose(iI[:, 3], full[:, 0, 1], rtol=1e-3, atol=1e-5)


def test_invalid_arguments_rejected(capi):
    L = capi.load()
    assert L.dy_malloc(None, 16) == capi.DY_ERR_INVALID_ARG
    assert L.dy_broadphase_create(None, 4, 4, None) == capi.DY_ERR_INVALID_ARG
    assert L.dy_integrate(None, None, 0, 0, 0, 0.1, 1) == capi.DY_ERR_INVALID_ARG
    n = C.c_int(-1)
    assert L.dy_device_count(C.byref(n)) == 0 and n.value >= 0
