"""The CPU oracle against the reference's own known-answer tests and golden hash.

Each test restates one reference test case (file:line cited) with the same
scene, tolerance and expected value, run through the oracle.  These pin the
oracle before it is trusted as the checker for the gfx950 kernels.
"""
import numpy as np
import pytest

from dyphur_amd import scenes

F32 = np.float32


def _run_frames(rc, sc, frames):
    b, s = rc.Bodies(sc), rc.Shapes(sc)
    w = rc.World.for_scene(sc)
    for _ in range(frames):
        w.frame(b, s, sc)
    return b


# core/tests/sim_determinism_test.cpp:218-226,458-471 (OMP golden)
def test_golden_box_hash(oracle):
    sc = scenes.determinism_boxes()
    h1 = _run_frames(oracle, sc, 20).fnv()
    h2 = _run_frames(oracle, sc, 20).fnv()
    assert h1 == h2
    assert h1 == 0xCBCD209C3665C818


# core/tests/sim_determinism_test.cpp:478-486
def test_256_box_stack_run_to_run(oracle):
    sc = scenes.determinism_boxes(8, 4, 8)
    assert _run_frames(oracle, sc, 20).fnv() == _run_frames(oracle, sc, 20).fnv()


def _one_body(**kw):
    sc = scenes.two_body(scenes.SHAPE_SPHERE, scenes.SHAPE_SPHERE, (1, 1, 1), (1, 1, 1),
                         (0, 0, 0), (100, 0, 0))
    for k, v in kw.items():
        getattr(sc, k)[0] = v
    return sc


# core/tests/integrator_test.cpp:26-50
def test_integrator_static_body(oracle):
    sc = _one_body(pos=(1, 2, 3), mass=0.0, flags=scenes.FLAG_STATIC)
    b = oracle.Bodies(sc)
    oracle.integrate(b, (0, -9.81, 0), 1 / 60, 1)
    assert abs(b.arr["pos_x"][0] - 1) < 1e-6 and abs(b.arr["pos_y"][0] - 2) < 1e-6
    assert abs(b.arr["pos_z"][0] - 3) < 1e-6


# core/tests/integrator_test.cpp:52-83: y_N = y0 + g (N+1) / (2N)
def test_integrator_free_fall_closed_form(oracle):
    sc = _one_body(pos=(0, 10, 0))
    b = oracle.Bodies(sc)
    oracle.integrate(b, (0, -9.81, 0), 1.0, 60)
    expected = F32(10.0) + F32(-9.81) * F32(61) / F32(120)
    assert abs(b.arr["pos_y"][0] - expected) < 1e-3
    assert abs(b.arr["pos_x"][0]) < 1e-5 and abs(b.arr["pos_z"][0]) < 1e-5


# core/tests/integrator_test.cpp:85-111
def test_integrator_constant_velocity(oracle):
    sc = _one_body(vel=(3, 0, 0))
    b = oracle.Bodies(sc)
    oracle.integrate(b, (0, 0, 0), 1.0, 1)
    assert abs(b.arr["pos_x"][0] - 3) < 1e-5 and abs(b.arr["pos_y"][0]) < 1e-5


# core/tests/integrator_test.cpp:113-145
def test_integrator_quaternion_norm_under_spin(oracle):
    sc = _one_body(ang=(0, 1, 0))
    b = oracle.Bodies(sc)
    for _ in range(60):
        oracle.integrate(b, (0, 0, 0), 1 / 60, 4)
    q = np.array([b.arr[k][0] for k in ("rot_w", "rot_x", "rot_y", "rot_z")], np.float32)
    assert abs(float(q @ q) - 1.0) < 1e-4
    assert abs(q[2]) > 0.1


# core/tests/body_store_test.cpp:21-110
def test_body_store_dynamics(oracle):
    sc = scenes.two_body(0, 0, (1, 1, 1), (1, 1, 1), (1, 2, 3), (0, 10, 0), mass=(2.0, 1.0),
                         flags=(0, scenes.FLAG_STATIC), vel_a=(0.5, 0, 0), inertia=(4.0, 1.0))
    b = oracle.Bodies(sc)
    assert b.arr["inv_mass"][0] == F32(0.5) and b.arr["inv_mass"][1] == 0
    assert b.flags[0] & 1 == 0 and b.flags[1] & 1 == 1
    sc2 = _one_body()
    sc2.inertia[0] = np.array([2, 0, 0, 0, 4, 0, 0, 0, 8], np.float32)
    b2 = oracle.Bodies(sc2)
    got = [b2.arr[k][0] for k in ("iI_xx", "iI_yy", "iI_zz", "iI_xy", "iI_xz", "iI_yz")]
    assert np.allclose(got, [0.5, 0.25, 0.125, 0, 0, 0], atol=1e-5)


def _spheres(positions, r=1.0, bounds=(-2, -2, -2, 2, 2, 2)):
    n = len(positions)
    sc = scenes.two_body(1, 1, (r, r, r), (r, r, r), (0, 0, 0), (0, 0, 0))
    b = scenes._alloc(n)
    b["pos"][:] = positions
    b["shape"][:] = 0
    for k, v in b.items():
        setattr(sc, k, v)
    sc.shape_type = np.array([1], np.uint32)
    sc.shape_half = np.array([[r, r, r]], np.float32)
    sc.bounds = np.array(bounds, np.float32)
    sc.max_pairs = 32
    return sc


# core/tests/broadphase_test.cpp:53-87
def test_broadphase_three_spheres(oracle):
    sc = _spheres([(0, 0, 0), (1, 0, 0), (100, 0, 0)], bounds=(-2, -2, -2, 102, 2, 2))
    w = oracle.World(4, 32, 16)
    n = w.build_and_query(oracle.Bodies(sc), oracle.Shapes(sc), sc.bounds)
    assert n == 1
    assert w.pairs(1).tolist() == [[0, 1]]


# core/tests/broadphase_test.cpp:90-131
def test_broadphase_four_spheres_all_pairs(oracle):
    sc = _spheres([(0, 0, 0), (0.5, 0, 0), (0, 0.5, 0), (0, 0, 0.5)])
    w = oracle.World(4, 32, 16)
    n = w.build_and_query(oracle.Bodies(sc), oracle.Shapes(sc), sc.bounds)
    assert n == 6
    p = w.pairs(n)
    assert (p[:, 0] < p[:, 1]).all()
    assert len({tuple(x) for x in p.tolist()}) == 6


# core/tests/broadphase_test.cpp:134-155
def test_broadphase_disjoint(oracle):
    sc = _spheres([(0, 0, 0), (5, 0, 0)], r=0.4, bounds=(-1, -1, -1, 6, 1, 1))
    w = oracle.World(4, 32, 16)
    assert w.build_and_query(oracle.Bodies(sc), oracle.Shapes(sc), sc.bounds) == 0


def _np_contact(oracle, sc):
    w = oracle.World(2, 32, 16)
    b, s = oracle.Bodies(sc), oracle.Shapes(sc)
    n = w.narrowphase(np.array([[0, 1]], np.uint32), b, s)
    return n, w.contacts()


# core/tests/narrowphase_test.cpp:58-88
def test_np_sphere_sphere(oracle):
    n, c = _np_contact(oracle, scenes.two_body(1, 1, (1, 1, 1), (1, 1, 1), (0, 0, 0), (1.5, 0, 0)))
    assert n == 1
    for k, v in (("depth", 0.5), ("norm_x", -1), ("norm_y", 0), ("norm_z", 0), ("pos_x", 0.75),
                 ("pos_y", 0), ("pos_z", 0)):
        assert abs(c[k][0] - v) < 1e-4, k


# core/tests/narrowphase_test.cpp:91-113
def test_np_sphere_sphere_none(oracle):
    n, _ = _np_contact(oracle, scenes.two_body(1, 1, (1, 1, 1), (1, 1, 1), (0, 0, 0), (3, 0, 0)))
    assert n == 0


# core/tests/narrowphase_test.cpp:117-151
def test_np_sphere_box(oracle):
    n, c = _np_contact(oracle, scenes.two_body(0, 1, (1, 1, 1), (0.5, 0.5, 0.5), (0, 0, 0), (1.3, 0, 0)))
    assert n == 1
    for k, v in (("depth", 0.2), ("norm_x", -1), ("norm_y", 0), ("norm_z", 0), ("pos_x", 1),
                 ("pos_y", 0), ("pos_z", 0)):
        assert abs(c[k][0] - v) < 1e-4, k


# core/tests/narrowphase_test.cpp:155-184
def test_np_box_box_face(oracle):
    n, c = _np_contact(oracle, scenes.two_body(0, 0, (1, 1, 1), (1, 1, 1), (0, 0, 0), (1.8, 0, 0)))
    assert n == 1
    for k, v in (("depth", 0.2), ("norm_x", -1), ("norm_y", 0), ("norm_z", 0), ("pos_x", 1)):
        assert abs(c[k][0] - v) < 1e-4, k


# core/tests/narrowphase_test.cpp:187-210
def test_np_box_box_none(oracle):
    n, _ = _np_contact(oracle, scenes.two_body(0, 0, (1, 1, 1), (1, 1, 1), (0, 0, 0), (5, 0, 0)))
    assert n == 0


def _solve_pair(oracle, sc, n_iters=1, dt=1 / 60):
    w = oracle.World(2, 32, 16, n_iters=n_iters)
    b, s = oracle.Bodies(sc), oracle.Shapes(sc)
    assert w.narrowphase(np.array([[0, 1]], np.uint32), b, s) == 1
    w.solve(b, dt)
    return b


# core/tests/xpbd_solver_test.cpp:37-70
def test_solver_sphere_sphere_separates(oracle):
    b = _solve_pair(oracle, scenes.two_body(1, 1, (1, 1, 1), (1, 1, 1), (0, 0, 0), (0, 1.5, 0)))
    assert b.arr["pos_y"][1] - b.arr["pos_y"][0] >= 2 - 1e-3


# core/tests/xpbd_solver_test.cpp:74-109
def test_solver_approaching_velocity_damped(oracle):
    b = _solve_pair(oracle, scenes.two_body(1, 1, (1, 1, 1), (1, 1, 1), (0, 0, 0), (0, 1.5, 0),
                                            vel_b=(0, -2, 0)))
    assert (b.arr["vel_y"][0] - b.arr["vel_y"][1]) * -1 >= -1e-3


# core/tests/xpbd_solver_test.cpp:112-149
def test_solver_static_unmoved(oracle):
    b = _solve_pair(oracle, scenes.two_body(1, 1, (1, 1, 1), (1, 1, 1), (0, 0, 0), (0, 1.5, 0),
                                            mass=(0.0, 1.0), flags=(scenes.FLAG_STATIC, 0)))
    assert abs(b.arr["pos_y"][0]) < 1e-3 and b.arr["pos_y"][1] >= 1 - 1e-3


def independent_pairs_scene(n_pairs=300):
    n = 2 * n_pairs
    sc = _spheres([(0, 0, 0)] * n, r=0.5)
    for k in range(n_pairs):
        sc.pos[2 * k] = (0, 10 * k, 0)
        sc.pos[2 * k + 1] = (0, 10 * k + 0.8, 0)
    sc.bounds = np.array([-5, -5, -5, 5, 10 * n_pairs + 5, 5], np.float32)
    return sc


# core/tests/xpbd_solver_test.cpp:157-230
def test_solver_colored_equals_serial_independent_pairs(oracle):
    sc = independent_pairs_scene()
    pairs = np.array([[2 * k, 2 * k + 1] for k in range(300)], np.uint32)
    res = []
    for serial in (False, True, False):
        w = oracle.World(600, 1024, 308)
        b, s = oracle.Bodies(sc), oracle.Shapes(sc)
        assert w.narrowphase(pairs, b, s) == 300
        st = w.solve(b, 1 / 60, force_serial=serial)
        assert st["used_serial"] == int(serial)
        res.append(b.arr["pos_y"].copy())
    assert np.array_equal(res[0].view(np.uint32), res[1].view(np.uint32))
    assert np.array_equal(res[0].view(np.uint32), res[2].view(np.uint32))


# compute/tests/smoke_test.cpp:74-101 (multi-launch bitonic path)
def test_bitonic_sort_reverse(oracle):
    import ctypes as C
    n = 1 << 10
    keys = np.arange(n - 1, -1, -1, dtype=np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    oracle.lib().rc_bitonic_sort_u32_u32(keys.ctypes.data, vals.ctypes.data, n)
    assert np.array_equal(keys, np.arange(n, dtype=np.uint32))
    assert np.array_equal(vals, np.arange(n - 1, -1, -1, dtype=np.uint32))


# Stress scene load at settle (REPORT.md:333-334: ~3 700 contacts / ~7 500 pairs)
def test_stress_scene_load_matches_report(oracle):
    sc = scenes.stress_blocks()
    b, s = oracle.Bodies(sc), oracle.Shapes(sc)
    w = oracle.World.for_scene(sc)
    for _ in range(150):
        st = w.frame(b, s, sc)
    assert 5000 < st["n_pairs"] < 12000
    assert 2500 < st["n_contacts"] < 6000
    assert st["used_serial"] == 0 and 4 <= st["n_colors"] <= 64


# core/src/xpbd_solver.cpp:639-651: a contact whose higher-priority neighbours
# hold all 64 colours sends the whole solve to the serial path.  The star_plate
# contacts all share the plate, so they form one clique of 301 and greedy
# colouring needs 301 colours; the oracle must report overflow and solve in
# contact order, i.e. give the same bits as a forced serial solve.
def test_star_plate_overflows_to_serial(oracle):
    sc = scenes.star_plate()
    res = []
    for serial in (False, True):
        b, s = oracle.Bodies(sc), oracle.Shapes(sc)
        oracle.integrate(b, sc.gravity, sc.dt, sc.n_substeps)
        w = oracle.World.for_scene(sc)
        n_pairs = w.build_and_query(b, s, sc.bounds)
        w.sort_pairs(n_pairs)
        nc = w.narrowphase(w.pairs(n_pairs), b, s)
        assert nc == 301
        st = w.solve(b, sc.dt, force_serial=serial)
        assert st["used_serial"] == 1
        if not serial:
            assert st["overflow"] == 1
        res.append(b.state())
    for f in ("pos_x", "pos_y", "vel_y", "ang_x", "rot_w"):
        assert np.array_equal(res[0][f].view(np.uint32), res[1][f].view(np.uint32)), f
