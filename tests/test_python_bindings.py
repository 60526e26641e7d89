"""The dyphur_py module (the reference's Python binding surface,
bindings/python/dyphur_py.cpp:29-163) over the gfx950 C++ layer.  The GPU
cases restate bindings/python/tests/test_contact_sensor.py, with
Device.default_gpu() in place of Device.default_cpu() (this build has no CPU
device)."""
import os
import sys

import numpy as np
import pytest

from dyphur_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def dp():
    pkg = os.path.join(ROOT, "dyphur_amd")
    if not any(f.startswith("dyphur_py") and f.endswith(".so") for f in os.listdir(pkg)):
        from dyphur_amd import build
        build.build_python_module()
    if pkg not in sys.path:
        sys.path.insert(0, pkg)
    import dyphur_py
    return dyphur_py


REFERENCE_NAMES = ["Device", "Stream", "AABB", "ShapeType", "ShapeParams", "ShapeStore", "ShapeView",
                   "BodyParams", "BodyView", "BodyStore", "Broadphase", "Narrowphase", "ContactSensor"]


def test_module_exposes_reference_surface(dp):
    for n in REFERENCE_NAMES + ["XpbdSolver", "IntegratorParams", "integrate", "step", "state_hash"]:
        assert hasattr(dp, n), n
    assert int(dp.ShapeType.Box) == 0 and int(dp.ShapeType.Sphere) == 1
    assert dp.XpbdSolver.kSerialThreshold == 256
    for m in ("set_position", "set_velocity", "mass", "shape_handle", "flags"):
        assert hasattr(dp.BodyParams(), m)
    a = dp.AABB(0, 0, 0, 1, 1, 1)
    assert a.overlaps(dp.AABB(1, 0, 0, 2, 1, 1)) and not a.overlaps(dp.AABB(1.5, 0, 0, 2, 1, 1))


def test_default_cpu_raises(dp):
    with pytest.raises(dp.Error):
        dp.Device.default_cpu()


@pytest.fixture(scope="module")
def two_sphere_scene(dp, gpu):
    dev = dp.Device.default_gpu()
    s = dev.make_stream()
    ss = dp.ShapeStore(s, 4)
    sp = dp.ShapeParams()
    sp.type = dp.ShapeType.Sphere
    sp.half_x = 1.0
    ss.add(sp)
    ss.upload()
    bs = dp.BodyStore(s, 4)
    p = dp.BodyParams()
    p.mass = 1.0
    p.shape_handle = 0
    p.set_position(0.0, 0.0, 0.0)
    bs.add(p)
    p.set_position(1.5, 0.0, 0.0)
    bs.add(p)
    bs.upload()
    s.wait()
    scene = dp.AABB(-2.0, -2.0, -2.0, 4.0, 2.0, 2.0)
    bp = dp.Broadphase(s, 4, 32)
    bp.build_and_query(s, bs.view(), ss.view(), scene)
    s.wait()
    n_pairs = bp.download_count(s)
    np_ = dp.Narrowphase(s, 32)
    np_.run(s, bp, n_pairs, bs.view(), ss.view())
    s.wait()
    return s, bp, np_, bs, ss, n_pairs


@pytest.mark.gpu
def test_broadphase_finds_pair(two_sphere_scene):
    assert two_sphere_scene[5] == 1


@pytest.mark.gpu
def test_narrowphase_finds_contact(two_sphere_scene):
    s, _, np_, _, _, _ = two_sphere_scene
    assert np_.download_count(s) == 1


@pytest.mark.gpu
def test_contact_sensor_count_depth_normal(dp, two_sphere_scene):
    s, _, np_, bs, _, _ = two_sphere_scene
    sensor = dp.ContactSensor(0)
    assert sensor.query(s, np_, bs.body_shapes()) == 1 and sensor.count() == 1
    assert float(sensor.depth()[0]) > 0.0
    n = np.array([sensor.norm_x()[0], sensor.norm_y()[0], sensor.norm_z()[0]], np.float64)
    assert abs(np.linalg.norm(n) - 1.0) < 1e-4


@pytest.mark.gpu
def test_contact_sensor_numpy_zero_copy(dp, two_sphere_scene):
    s, _, np_, bs, _, _ = two_sphere_scene
    sensor = dp.ContactSensor(0)
    sensor.query(s, np_, bs.body_shapes())
    assert not np.asarray(sensor.depth()).flags["OWNDATA"]
    assert not np.asarray(sensor.body_a()).flags["OWNDATA"]


@pytest.mark.gpu
def test_contact_sensor_no_hit_wrong_shape(dp, two_sphere_scene):
    s, _, np_, bs, _, _ = two_sphere_scene
    assert dp.ContactSensor(99).query(s, np_, bs.body_shapes()) == 0


@pytest.mark.gpu
def test_step_golden_hash_and_oracle_trajectory(dp, gpu, oracle):
    """The reference's determinism scene stepped through the Python API: the
    golden free-fall hash after 20 frames, and the oracle's state."""
    sc = scenes.determinism_boxes()
    s = dp.Device.default_gpu().make_stream()
    ss = dp.ShapeStore(s, sc.n_shapes)
    for t, h in zip(sc.shape_type, sc.shape_half):
        sp = dp.ShapeParams()
        sp.type = dp.ShapeType(int(t))
        sp.half_x, sp.half_y, sp.half_z = (float(v) for v in h)
        ss.add(sp)
    ss.upload()
    bs = dp.BodyStore(s, sc.n)
    for i in range(sc.n):
        p = dp.BodyParams()
        p.set_position(*(float(v) for v in sc.pos[i]))
        p.mass = float(sc.mass[i])
        p.set_inertia([float(v) for v in sc.inertia[i]])
        p.shape_handle = int(sc.shape[i])
        p.flags = int(sc.flags[i])
        bs.add(p)
    bs.upload()
    bounds = dp.AABB(*(float(v) for v in sc.bounds))
    bp = dp.Broadphase(s, sc.n, sc.max_pairs)
    np_ = dp.Narrowphase(s, sc.max_contacts)
    solver = dp.XpbdSolver(s, sc.n_iters, sc.max_contacts, sc.mu)
    ip = dp.IntegratorParams()
    ip.set_gravity(*sc.gravity)
    ip.dt = sc.dt
    ip.n_substeps = 1
    for _ in range(20):
        dp.step(s, bs.view(), ss.view(), bp, np_, solver, bounds, ip)
    assert dp.state_hash(s, bs.view()) == 0xCBCD209C3665C818
    ob, osh = oracle.Bodies(sc), oracle.Shapes(sc)
    ow = oracle.World.for_scene(sc)
    for _ in range(20):
        ow.frame(ob, osh, sc)
    got, ref = bs.state(s), ob.state()
    for f in ("pos_x", "pos_y", "pos_z", "rot_w", "vel_y"):
        assert np.array_equal(got[f].view(np.uint32), ref[f].view(np.uint32)), f
