"""XPBD joints (core/src/xpbd_solver.cpp:60-407, 523-556, 559-733;
core/include/core/articulation.hpp, joint_store.hpp).

CPU: the oracle's joint restatement against the reference's golden
articulation hash (core/tests/sim_determinism_test.cpp:258-270) and its
articulation cases (core/tests/articulation_test.cpp:139-320), closed forms
for the parts those leave open (prismatic slide, limits, motors, compliance,
momentum, hinge angle), the Jones-Plassmann colours over contacts and joints
against a round-by-round simulation, and the kernels' joint arithmetic
(joint_math.hpp, host build) bit for bit against the oracle.

GPU: dy_xpbd_solve with joints bit for bit against the oracle: the
articulated arm, joint chains with contacts on the reference-order and the
coloured paths, every sweep variant, joint validation.
"""
import ctypes as C

import numpy as np
import pytest

from dyphur_amd import scenes
from tests import _hostmath

F32 = np.float32
GOLDEN_ARTIC = 0x44D086AF21F338CF  # core/tests/sim_determinism_test.cpp:246 (OpenMP build)
EMPTY_CONTACTS = {k: np.zeros(0) for k in ("body_a", "body_b", "pos_x", "pos_y", "pos_z", "norm_x", "norm_y",
                                           "norm_z", "depth")}


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _arm_run(oracle, frames=20):
    """integrate + solve with the arm's joints and no contacts, as the reference's run_artic_sim."""
    sc = scenes.articulated_arm()
    b = oracle.Bodies(sc)
    w = oracle.World(sc.n, 1, 1, 10, 0.5)
    w.set_contacts(EMPTY_CONTACTS)
    js = oracle.Joints(sc.joints)
    L = oracle.lib()
    for _ in range(frames):
        L.rc_integrate(C.byref(b.view()), 0.0, -9.81, 0.0, sc.dt, 1)
        w.solve(b, sc.dt, joints=js)
    return b


@pytest.fixture
def libc_atan2(oracle):
    oracle.lib().rc_set_libc_atan2(1)
    yield
    oracle.lib().rc_set_libc_atan2(0)


def test_oracle_articulated_arm_golden_hash(oracle, libc_atan2):
    """The reference's own golden: with the C library's float atan2 (what
    sycl::atan2 is on its OpenMP backend) the oracle reproduces it exactly."""
    h1 = _arm_run(oracle).fnv()
    h2 = _arm_run(oracle).fnv()
    assert h1 == h2 == GOLDEN_ARTIC


def test_oracle_arm_with_correctly_rounded_atan2_stays_within_ulps(oracle):
    """The default (correctly rounded) atan2 changes the arm's last bits only."""
    a = _arm_run(oracle)
    oracle.lib().rc_set_libc_atan2(1)
    try:
        b = _arm_run(oracle)
    finally:
        oracle.lib().rc_set_libc_atan2(0)
    for k in ("pos_x", "pos_y", "pos_z", "rot_w", "rot_x", "rot_y", "rot_z"):
        assert np.allclose(a.arr[k], b.arr[k], rtol=0, atol=2e-7), k


# ---- core/tests/articulation_test.cpp:120-320 ------------------------------
def _two(oracle, pos_b, static_a, jtype, n_iters, axis=(0.0, 1.0, 0.0)):
    """Two unit-mass bodies with identity inertia (two_bodies, :124-137), one
    joint with anchors (0, 0.5, 0) / (0, -0.5, 0), one solve at 1/60 s."""
    sc = scenes.two_body(1, 1, (0.1, 0.1, 0.1), (0.1, 0.1, 0.1), (0, 0, 0), pos_b,
                         mass=(0.0 if static_a else 1.0, 1.0), flags=(1 if static_a else 0, 0))
    j = scenes.joint_arrays(1)
    j["body_parent"][0], j["body_child"][0] = 0, 1
    j["anchor_py"][0], j["anchor_cy"][0] = 0.5, -0.5
    j["axis_px"][0], j["axis_py"][0], j["axis_pz"][0] = axis
    j["type"][0] = jtype
    b = oracle.Bodies(sc)
    w = oracle.World(2, 1, 1, n_iters, 0.5)
    w.set_contacts(EMPTY_CONTACTS)
    w.solve(b, 1.0 / 60.0, joints=oracle.Joints(j))
    return b.arr


def test_reference_fixed_joint_holds_two_bodies(oracle):
    a = _two(oracle, (0, 1.2, 0), False, scenes.JOINT_FIXED, 20)
    assert abs((a["pos_y"][0] + 0.5) - (a["pos_y"][1] - 0.5)) <= 1e-3


def test_reference_fixed_joint_static_parent(oracle):
    a = _two(oracle, (0.5, 1.0, 0), True, scenes.JOINT_FIXED, 100)
    assert abs(a["pos_x"][0]) <= 1e-5 and abs(a["pos_y"][0]) <= 1e-5
    assert abs(a["pos_x"][1]) <= 1e-3 and abs(a["pos_y"][1] - 1.0) <= 1e-3


def test_reference_revolute_joint_satisfied(oracle):
    a = _two(oracle, (0, 1.0, 0), True, scenes.JOINT_REVOLUTE, 10)
    assert abs(a["pos_x"][1]) <= 1e-3 and abs(a["pos_y"][1] - 0.5 - 0.5) <= 1e-3 and abs(a["pos_z"][1]) <= 1e-3


def test_reference_ball_joint_violation_corrected(oracle):
    a = _two(oracle, (0, 1.3, 0), True, scenes.JOINT_BALL, 20)
    assert abs(a["pos_x"][1]) <= 1e-3 and abs(a["pos_y"][1] - 1.0) <= 1e-3 and abs(a["pos_z"][1]) <= 1e-3


# ---- closed forms ------------------------------------------------------------
def _pair(oracle, jd, pos_b, rot_b=(1, 0, 0, 0), vel_b=(0, 0, 0), ang_b=(0, 0, 0), mass_b=1.0, inertia_b=1.0,
          static_a=True, n_iters=1, dt=1.0 / 60.0):
    sc = scenes.two_body(1, 1, (0.1, 0.1, 0.1), (0.1, 0.1, 0.1), (0, 0, 0), pos_b,
                         mass=(0.0 if static_a else 1.0, mass_b), flags=(1 if static_a else 0, 0),
                         rot_b=rot_b, vel_b=vel_b, inertia=(1.0, inertia_b))
    sc.ang[1] = ang_b
    b = oracle.Bodies(sc)
    w = oracle.World(2, 1, 1, n_iters, 0.5)
    w.set_contacts(EMPTY_CONTACTS)
    j = scenes.joint_arrays(1)
    j["body_parent"][0], j["body_child"][0] = 0, 1
    for k, v in jd.items():
        j[k][0] = v
    w.solve(b, dt, joints=oracle.Joints(j))
    return b.arr


def test_prismatic_removes_transverse_offset_and_keeps_the_slide(oracle):
    """Anchors at the centres: the transverse error (0.2, -, -0.1) is removed,
    the offset along the axis (y) is untouched, orientation stays identity."""
    a = _pair(oracle, {"type": scenes.JOINT_PRISMATIC, "axis_px": 0, "axis_py": 1, "axis_pz": 0}, (0.2, -0.7, -0.1))
    assert abs(a["pos_x"][1]) <= 1e-6 and abs(a["pos_z"][1]) <= 1e-6
    assert a["pos_y"][1] == F32(-0.7)
    assert a["rot_w"][1] == 1.0


def test_prismatic_limit_moves_along_the_axis_with_the_reference_sign(oracle):
    """apply_limit (:338-345) on a prismatic joint: with offset v = (p_b -
    p_a).u below lo, viol = v - lo < 0 and dl = -viol / w; the parent moves
    by +im_a dl u and the child by -im_b dl u, so v falls by dl: a static
    parent and unit inverse mass take v = -0.6 (lo = -0.5) to -0.7.  The
    violation doubles; the reference's tests never reach this branch."""
    a = _pair(oracle, {"type": scenes.JOINT_PRISMATIC, "axis_py": 1, "limit_lo": -0.5, "limit_hi": -0.3},
              (0.0, -0.6, 0.0))
    assert abs(a["pos_y"][1] - (-0.7)) <= 1e-6


def test_revolute_limit_brings_the_angle_back_to_the_bound(oracle):
    """Child turned by 0.8 rad about z, limits [-0.5, 0.5]: after the limit's
    angular correction (static parent, isotropic child) the hinge angle is
    0.5 to first order; the angle is measured with the motor's stiffness term
    below, here through the resulting orientation."""
    th = 0.8
    q = (np.cos(th / 2), 0.0, 0.0, np.sin(th / 2))
    a = _pair(oracle, {"type": scenes.JOINT_REVOLUTE, "axis_px": 0, "axis_py": 0, "axis_pz": 1,
                       "limit_lo": -0.5, "limit_hi": 0.5}, (0.0, -1.0, 0.0), rot_b=q)
    ang = 2 * np.arctan2(a["rot_z"][1], a["rot_w"][1])
    assert abs(ang - 0.5) <= 2e-2


def test_revolute_motor_damping_closed_form(oracle):
    """Damping-only motor, static parent, isotropic child: the relative
    angular velocity along the axis becomes w + h d (target - w)."""
    h, d, tv, w0 = 1.0 / 60.0, 12.0, 1.5, -0.4
    a = _pair(oracle, {"type": scenes.JOINT_REVOLUTE, "axis_px": 1, "axis_py": 0, "damping": d, "target_vel": tv},
              (0.0, -1.0, 0.0), ang_b=(w0, 0, 0), inertia_b=2.0, n_iters=0)
    assert abs(a["ang_x"][1] - (w0 + h * d * (tv - w0))) <= 1e-6


def test_revolute_motor_stiffness_reads_the_hinge_angle(oracle):
    """Stiffness-only motor: the angular velocity change along the axis is
    h k (target - theta), so it reads the hinge angle theta of a child turned
    by theta about the axis (joint_current_value, :278-318)."""
    h, k, tp = 1.0 / 60.0, 40.0, 0.1
    for th in (0.3, -1.2, 2.5):
        q = (np.cos(th / 2), np.sin(th / 2), 0.0, 0.0)
        a = _pair(oracle, {"type": scenes.JOINT_REVOLUTE, "axis_px": 1, "axis_py": 0, "stiffness": k,
                           "target_pos": tp}, (0.0, -1.0, 0.0), rot_b=q, n_iters=0)
        theta = tp - a["ang_x"][1] / (h * k)
        assert abs(theta - th) <= 1e-4, (th, theta)


def test_prismatic_motor_damping_closed_form(oracle):
    h, d, tv, v0 = 1.0 / 60.0, 8.0, 0.25, -1.0
    a = _pair(oracle, {"type": scenes.JOINT_PRISMATIC, "axis_py": 1, "damping": d, "target_vel": tv},
              (0.0, -1.0, 0.0), vel_b=(0, v0, 0), mass_b=2.0, n_iters=0)
    assert abs(a["vel_y"][1] - (v0 + h * d * (tv - v0))) <= 1e-6


def test_compliance_leaves_the_xpbd_fraction(oracle):
    """Ball joint, anchors at the centres, unit inverse mass: one projection
    with compliance c removes the fraction 1 / (1 + c / h^2) of the gap."""
    h, c, gap = 1.0 / 60.0, 2e-4, 0.3
    a = _pair(oracle, {"type": scenes.JOINT_BALL, "compliance_pos": c}, (0.0, -gap, 0.0), dt=h)
    alpha = c / (h * h)
    assert abs(-a["pos_y"][1] - gap * alpha / (1 + alpha)) <= 1e-6


def test_joint_conserves_linear_momentum(oracle):
    """Two dynamic bodies, random anchors and poses: every positional impulse
    is equal and opposite, so sum m dx = 0 and sum m dv = 0."""
    rng = np.random.default_rng(4)
    for _ in range(6):
        sc = scenes.two_body(1, 1, (0.1,) * 3, (0.1,) * 3, (0, 0, 0), tuple(rng.normal(size=3) * 0.3 + [0, -1, 0]),
                             mass=(1.5, 0.7), rot_b=tuple(np.r_[1.0, rng.normal(size=3) * 0.2]))
        b = oracle.Bodies(sc)
        x0 = {k: b.arr[k].astype(np.float64).copy() for k in ("pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z")}
        w = oracle.World(2, 1, 1, 1, 0.5)
        w.set_contacts(EMPTY_CONTACTS)
        j = scenes.joint_arrays(1)
        j["body_parent"][0], j["body_child"][0] = 0, 1
        j["anchor_px"][0], j["anchor_py"][0], j["anchor_pz"][0] = rng.normal(size=3) * 0.2
        j["anchor_cx"][0], j["anchor_cy"][0], j["anchor_cz"][0] = rng.normal(size=3) * 0.2
        j["type"][0] = scenes.JOINT_BALL
        w.solve(b, 1.0 / 60.0, joints=oracle.Joints(j))
        m = np.array([1.5, 0.7])
        for k in x0:
            assert abs(np.sum(m * (b.arr[k].astype(np.float64) - x0[k]))) <= 1e-5, k


# ---- colouring over contacts and joints ---------------------------------------
def test_jp_colours_with_joints_match_round_simulation(oracle):
    """Constraint k >= nc is joint k - nc (xpbd_solver.cpp:611-614,649-660):
    the colours over the combined endpoint list equal the reference's round
    rule simulated directly."""
    from tests.test_oracle_pins import _fmix32
    sc = scenes.joint_chains(40, 8)
    b, s = oracle.Bodies(sc), oracle.Shapes(sc)
    w = oracle.World.for_scene(sc)
    js = oracle.Joints(sc.joints)
    for _ in range(60):
        w.frame(b, s, sc, joints=js)
        c = w.contacts()
        if c["n"] >= 8:
            break
    nc, nj = c["n"], js.n
    assert nc > 0 and nc + nj > 256
    n_con = nc + nj
    im = b.arr["inv_mass"]
    A = np.r_[c["body_a"], sc.joints["body_parent"]].astype(np.int64)
    B = np.r_[c["body_b"], sc.joints["body_child"]].astype(np.int64)
    # recoloured below by a solve over the same contacts
    pri = (_fmix32(np.arange(n_con)) << np.uint64(32)) | np.arange(n_con, dtype=np.uint64)
    ends = [[x for x in (A[k], B[k]) if im[x] != 0] for k in range(n_con)]
    colour = np.full(n_con, -1, np.int64)
    mask = np.zeros(sc.n, np.uint64)
    left = set(range(n_con))
    while left:
        best = {}
        for k in left:
            for x in ends[k]:
                best[x] = max(best[x], pri[k]) if x in best else pri[k]
        winners = [k for k in left if all(best[x] == pri[k] for x in ends[k])]
        for k in winners:
            used = np.uint64(0)
            for x in ends[k]:
                used |= mask[x]
            bit = 0
            while (int(used) >> bit) & 1:
                bit += 1
            colour[k] = bit
            for x in ends[k]:
                mask[x] |= np.uint64(1) << np.uint64(bit)
        left -= set(winners)
    bb = b.copy()
    st = w.solve(bb, sc.dt, joints=js)
    assert st["used_serial"] == 0 and st["n_constraints"] == n_con
    assert np.array_equal(w.colors(n_con).astype(np.int64), colour)


def test_coloured_equals_serial_for_independent_joints(oracle):
    """Disjoint two-body joints (no shared body): colour order and index
    order give the same bits (cf. core/tests/xpbd_solver_test.cpp:157-230)."""
    n_pairs = 300  # > 256 constraints: coloured path
    n = 2 * n_pairs
    rng = np.random.default_rng(9)
    b0 = scenes._alloc(n)
    b0["pos"][:] = rng.normal(size=(n, 3)).astype(F32)
    b0["pos"][1::2] += F32(0.6)
    b0["mass"][:] = 1.0
    sc = scenes.Scene(name="pairs", shape_type=np.array([1], np.uint32), shape_half=np.array([[0.1] * 3], F32),
                      bounds=np.array([-9, -9, -9, 9, 9, 9], F32), max_pairs=1, max_contacts=1, **b0)
    j = scenes.joint_arrays(n_pairs)
    j["body_parent"][:] = np.arange(0, n, 2)
    j["body_child"][:] = np.arange(1, n, 2)
    j["type"][:] = rng.integers(0, 4, n_pairs)
    j["anchor_py"][:], j["anchor_cy"][:] = 0.3, -0.3
    j["axis_px"][:], j["axis_py"][:], j["axis_pz"][:] = 0, 0, 1
    j["damping"][:] = 5.0
    j["target_vel"][:] = 1.0
    res = []
    for fs in (False, True):
        b = oracle.Bodies(sc)
        w = oracle.World(n, 1, 1, 10, 0.5)
        w.set_contacts(EMPTY_CONTACTS)
        st = w.solve(b, 1 / 60, force_serial=fs, joints=oracle.Joints(j))
        assert st["used_serial"] == int(fs)
        res.append(b.state())
    for k in scenes_fields():
        assert np.array_equal(bits(res[0][k]), bits(res[1][k])), k


def scenes_fields():
    return ("pos_x", "pos_y", "pos_z", "rot_w", "rot_x", "rot_y", "rot_z", "vel_x", "vel_y", "vel_z",
            "ang_x", "ang_y", "ang_z")


# ---- the kernels' arithmetic (host build) against the oracle ----------------------
def test_atan2_cr_is_the_correctly_rounded_atan2(oracle):
    """joint_math.hpp's atan2_cr equals float64 atan2 rounded once, and the
    oracle's default atan2, on random and special inputs."""
    L = _hostmath.lib()
    rng = np.random.default_rng(2)
    n = 400_000
    y = np.r_[rng.normal(size=n), rng.normal(size=n) * np.exp(rng.uniform(-60, 60, n))].astype(F32)
    x = np.r_[rng.normal(size=n), rng.normal(size=n) * np.exp(rng.uniform(-60, 60, n))].astype(F32)
    sp = np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.nan, 1e-45, -1e-45, 3e38], F32)
    Y, X = np.meshgrid(sp, sp)
    y, x = np.r_[y, Y.ravel()], np.r_[x, X.ravel()]
    out = np.empty_like(y)
    L.hm_atan2_cr_array(y.ctypes.data, x.ctypes.data, out.ctypes.data, len(y))
    want = np.arctan2(y.astype(np.float64), x.astype(np.float64)).astype(F32)
    same = (out.view(np.uint32) == want.view(np.uint32)) | (np.isnan(out) & np.isnan(want))
    assert same.all(), np.nonzero(~same)[0][:5]
    o = oracle.lib()
    for yy, xx, got in zip(y[:2000], x[:2000], out[:2000]):
        assert bits(np.float32(o.rc_joint_atan2(float(yy), float(xx)))) == bits(got)


def _random_joint(rng, jtype):
    jp = np.zeros(17, F32)
    jp[0:6] = rng.normal(size=6) * 0.3
    ax = rng.normal(size=3)
    jp[6:9] = ax / np.linalg.norm(ax) if rng.uniform() < 0.5 else np.eye(3)[rng.integers(3)]
    if rng.uniform() < 0.6:
        w = rng.uniform(0.05, 1.0)
        jp[9:11] = (-w, w) if jtype != scenes.JOINT_PRISMATIC else (-0.3 - w, -0.3 + w)
    else:
        jp[9:11] = (1.0, -1.0)
    if rng.uniform() < 0.7:
        jp[11:15] = (rng.normal() * 0.5, rng.normal(), rng.uniform(0, 200), rng.uniform(0, 30))
    if rng.uniform() < 0.5:
        jp[15:17] = rng.uniform(1e-6, 1e-3, 2)
    return jp


def _random_body(rng, static):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    q *= 1 + rng.normal() * 1e-3
    b = np.zeros(20, F32)
    b[0:3] = rng.normal(size=3)
    b[3:7] = q
    b[7:10] = rng.normal(size=3)
    b[10:13] = rng.normal(size=3)
    if not static:
        b[13] = rng.uniform(0.2, 3.0)
        d = rng.uniform(0.5, 5.0, 3)
        b[14:17] = d
        b[17:20] = rng.normal(size=3) * 0.1
    return b


def test_kernel_joint_arithmetic_equals_oracle_bitwise(oracle):
    """hm_joint_pair (joint_math.hpp compiled for the host: motor, then
    iterations of solve_joint and renormalisation) against the oracle's
    serial solve of the same two bodies and joint, every type, static and
    dynamic parents, limits, motors, compliance."""
    L = _hostmath.lib()
    rng = np.random.default_rng(11)
    names = ("pos_x", "pos_y", "pos_z", "rot_w", "rot_x", "rot_y", "rot_z", "vel_x", "vel_y", "vel_z",
             "ang_x", "ang_y", "ang_z", "inv_mass", "iI_xx", "iI_yy", "iI_zz", "iI_xy", "iI_xz", "iI_yz")
    for case in range(400):
        jtype = case % 4
        iters = int(rng.integers(1, 4))
        a = _random_body(rng, static=rng.uniform() < 0.4)
        bb = _random_body(rng, static=False)
        jp = _random_joint(rng, jtype)
        dt = 1.0 / 60.0
        ga, gb = a.copy(), bb.copy()
        L.hm_joint_pair(ga.ctypes.data, gb.ctypes.data, jp.ctypes.data, jtype, dt, iters)
        ob = oracle.Bodies(n=2)
        for i, k in enumerate(names):
            ob.arr[k][:] = (a[i], bb[i])
        w = oracle.World(2, 1, 1, iters, 0.5)
        w.set_contacts(EMPTY_CONTACTS)
        j = scenes.joint_arrays(1)
        j["body_parent"][0], j["body_child"][0], j["type"][0] = 0, 1, jtype
        for i, k in enumerate(scenes.JOINT_F32):
            j[k][0] = jp[i]
        w.solve(ob, dt, joints=oracle.Joints(j))
        for i, k in enumerate(names[:13]):
            assert bits(ob.arr[k][0]) == bits(ga[i]) and bits(ob.arr[k][1]) == bits(gb[i]), (case, jtype, k)


# ---- GPU: dy_xpbd_solve with joints against the oracle -----------------------
def _gpu_arm(gpu, frames=20):
    sc = scenes.articulated_arm()
    s = gpu.Stream()
    db = gpu.DeviceBodies(sc, s)
    dj = gpu.DeviceJoints(sc.joints, s)
    npz = gpu.Narrowphase(s, 4)
    npz.run(s, 0, 0, db, gpu.DeviceShapes(sc, s))  # an empty contact set
    solver = gpu.XpbdSolver(s, 10, 16, 0.5)
    for _ in range(frames):
        gpu.integrate(s, db, sc.gravity, sc.dt, 1)
        solver.solve(s, npz.contacts_view(), db, sc.dt, dj)
    st = solver.stats()
    return db.download(s), st


@pytest.mark.gpu
def test_gpu_articulated_arm_equals_oracle(gpu, oracle):
    print("[joints] articulated arm, 20 frames", flush=True)
    got, st = _gpu_arm(gpu)
    assert st["n_constraints"] == 2 and st["used_serial"] == 1
    want = _arm_run(oracle)
    for k in scenes_fields():
        assert np.array_equal(bits(got[k]), bits(want.arr[k])), k


def _world_vs_oracle(gpu, oracle, sc, frames, force_serial=False):
    w = gpu.World(sc)
    ob, os_ = oracle.Bodies(sc), oracle.Shapes(sc)
    ow = oracle.World.for_scene(sc)
    js = oracle.Joints(sc.joints)
    paths = set()
    for f in range(frames):
        g = w.step(force_serial=force_serial)
        o = ow.frame(ob, os_, sc, force_serial=force_serial, joints=js)
        paths.add(g["used_serial"])
        assert g["n_pairs"] == o["n_pairs"] and g["n_contacts"] == o["n_contacts"], (f, g, o)
        st = w.state()
        for k in scenes_fields():
            assert np.array_equal(bits(st[k]), bits(ob.arr[k])), (f, k)
    return paths


@pytest.mark.gpu
@pytest.mark.parametrize("n_chains,links,serial", [(6, 5, 1), (40, 8, 0)])
def test_gpu_joint_chains_frames_equal_oracle(gpu, oracle, n_chains, links, serial):
    """Full frames (integrate, broadphase, narrowphase, contacts + joints
    solve) bit for bit: 30 constraints on the reference-order path, 320 + on
    the coloured one."""
    sc = scenes.joint_chains(n_chains, links)
    print(f"[joints] chains {n_chains}x{links} ({sc.n} bodies), 40 frames", flush=True)
    paths = _world_vs_oracle(gpu, oracle, sc, 40)
    assert paths == {serial}


@pytest.mark.gpu
def test_gpu_joint_chains_force_serial_equals_oracle(gpu, oracle):
    sc = scenes.joint_chains(40, 8, seed=3)
    print("[joints] chains 40x8 forced serial, 20 frames", flush=True)
    assert _world_vs_oracle(gpu, oracle, sc, 20, force_serial=True) == {1}


@pytest.mark.gpu
@pytest.mark.parametrize("env", [{"DYPHUR_SWEEP_MODE": "0"}, {"DYPHUR_SWEEP_MODE": "1"},
                                 {"DYPHUR_SWEEP_MODE": "3"}, {"DYPHUR_SOLVER_PATH": "1"}],
                         ids=["phases_one_xcd", "phases_grid", "dataflow_lane_cursors", "launch_per_class"])
def test_gpu_joint_sweep_variants_equal_oracle(gpu, oracle, monkeypatch, env):
    """Every sweep variant (read when the solver is created) gives the
    oracle's bits with joints on the coloured path."""
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    sc = scenes.joint_chains(40, 8, seed=7)
    print(f"[joints] sweep variant {env}, 15 frames", flush=True)
    assert _world_vs_oracle(gpu, oracle, sc, 15) == {0}


@pytest.mark.gpu
def test_gpu_invalid_joint_is_rejected(gpu):
    """A joint naming a body out of range: the solve runs without joints and
    the error comes back as an invalid argument (no device fault)."""
    sc = scenes.articulated_arm()
    sc.joints["body_child"][1] = 7
    s = gpu.Stream()
    db = gpu.DeviceBodies(sc, s)
    dj = gpu.DeviceJoints(sc.joints, s)
    npz = gpu.Narrowphase(s, 4)
    npz.run(s, 0, 0, db, gpu.DeviceShapes(sc, s))
    solver = gpu.XpbdSolver(s, 10, 16, 0.5)
    solver.solve(s, npz.contacts_view(), db, sc.dt, dj)
    with pytest.raises(gpu.DyError) as e:
        solver.stats()
    assert e.value.status == -1


_SECOND_PROCESS = r"""
import sys
sys.path.insert(0, sys.argv[1])
from dyphur_amd import capi, scenes
capi.load()
w = capi.World(scenes.joint_chains(40, 8, seed=7))
for _ in range(8):
    w.step()
print(w.fnv())
"""


@pytest.mark.gpu
def test_gpu_second_process_takes_the_launch_per_round_path(gpu, tmp_path, monkeypatch):
    """While this process holds the device's cooperative launches, another
    process's are refused; its solver then launches per round and per class
    and gives the same bits.  Both processes ask for cooperative launches
    (DYPHUR_COOP_LAUNCH=1; the default launches are ordinary ones)."""
    import os
    import subprocess
    import sys
    monkeypatch.setenv("DYPHUR_COOP_LAUNCH", "1")
    w = gpu.World(scenes.joint_chains(40, 8, seed=7))
    for _ in range(8):
        w.step()
    assert w.last["used_serial"] == 0
    here = w.fnv()
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    script = tmp_path / "second.py"
    script.write_text(_SECOND_PROCESS)
    print("[joints] second process on the device", flush=True)
    out = subprocess.run([sys.executable, str(script), root], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr[-2000:]
    assert int(out.stdout.strip().splitlines()[-1]) == here
