"""gfx950 kernels vs the CPU oracle, through the C ABI, on identical inputs.

Bar: bit-exact for every integer/index output (Morton keys, sorted order, BVH
topology, fp16 node boxes, pair list, contact order) and for the float outputs
too (contacts, solver state, whole-frame state hash): both sides run the same
IEEE operation sequence with no contraction, so any difference is a bug.
"""
import numpy as np
import pytest

from dyphur_amd import scenes

pytestmark = pytest.mark.gpu

U32 = np.uint32


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _settled_state(oracle, sc, frames):
    """Advance the oracle `frames` frames; returns (oracle Bodies, shapes)."""
    b, s = oracle.Bodies(sc), oracle.Shapes(sc)
    w = oracle.World.for_scene(sc)
    for _ in range(frames):
        w.frame(b, s, sc)
    return b, s


def _gpu_bodies(gpu, sc, s, state):
    d = gpu.DeviceBodies(sc, s)
    d.upload_state(state, s)
    return d


CASES = [
    ("stress_f0", lambda: scenes.stress_blocks(), 0),
    ("stress_f90", lambda: scenes.stress_blocks(), 90),
    ("stress_f140", lambda: scenes.stress_blocks(), 140),
    ("mixed_3000", lambda: scenes.random_mixed(3000, seed=11, extent=(12, 10, 12)), 0),
    ("mixed_3000_f30", lambda: scenes.random_mixed(3000, seed=12, extent=(12, 10, 12)), 30),
    ("mixed_12000", lambda: scenes.random_mixed(12000, seed=13, extent=(24, 16, 24)), 0),
    # the 32 768-body workload at its stated size: the initial drop and frame 40
    ("c2_32768_f0", lambda: scenes.random_mixed(32768, seed=20260002), 0),
    ("c2_32768_f40", lambda: scenes.random_mixed(32768, seed=20260002), 40),
]


@pytest.fixture(scope="module")
def stage_inputs(oracle):
    out = {}
    for name, mk, frames in CASES:
        sc = mk()
        b, sh = _settled_state(oracle, sc, frames)
        out[name] = (sc, b, sh)
    return out


@pytest.mark.parametrize("case", [c[0] for c in CASES])
def test_broadphase_bitexact(gpu, oracle, stage_inputs, case):
    sc, ob, osh = stage_inputs[case]
    print(f"[broadphase] {case} n={sc.n}", flush=True)
    w = oracle.World.for_scene(sc)
    n_ref = w.build_and_query(ob, osh, sc.bounds)
    ref = w.bvh(sc.n)
    w.sort_pairs(n_ref)
    ref_pairs = w.pairs(n_ref)

    s = gpu.Stream()
    db = _gpu_bodies(gpu, sc, s, ob.state())
    ds = gpu.DeviceShapes(sc, s)
    bp = gpu.Broadphase(s, sc.n, sc.max_pairs)
    bp.build_and_query(s, db, ds, sc.bounds)
    n = bp.download_count(s)
    got = bp.bvh(s, sc.n)
    assert np.array_equal(got["morton"], ref["morton"])
    assert np.array_equal(got["sorted_idx"], ref["sorted_idx"])
    assert got["root"][0] == ref["root"][0] == 0
    assert np.array_equal(got["left"], ref["left"])
    assert np.array_equal(got["right"], ref["right"])
    assert np.array_equal(got["parent"], ref["parent"])
    assert np.array_equal(got["aabb"], ref["aabb"])
    assert n == n_ref
    assert np.array_equal(bp.pairs(s, n), ref_pairs)


@pytest.mark.parametrize("case", [c[0] for c in CASES])
def test_narrowphase_bitexact(gpu, oracle, stage_inputs, case):
    sc, ob, osh = stage_inputs[case]
    print(f"[narrowphase] {case}", flush=True)
    w = oracle.World.for_scene(sc)
    n_pairs = w.build_and_query(ob, osh, sc.bounds)
    w.sort_pairs(n_pairs)
    pairs = w.pairs(n_pairs)
    nc_ref = w.narrowphase(pairs, ob, osh)
    ref = w.contacts()

    feat_ref = w.features()

    s = gpu.Stream()
    db = _gpu_bodies(gpu, sc, s, ob.state())
    ds = gpu.DeviceShapes(sc, s)
    dp = gpu.DeviceArray.from_numpy(pairs.ravel(), s)
    np_ = gpu.Narrowphase(s, sc.max_contacts)
    np_.run(s, dp.ptr, len(pairs), db, ds)
    got = np_.contacts(s)
    assert got["n"] == nc_ref == ref["n"]
    assert np_.last_count(s) == ref["last_count"]
    for f, _ in gpu.CONTACT_FIELDS:
        assert np.array_equal(bits(got[f]), bits(ref[f])), f
    # the device-count overload over a larger launch bound, with feature words
    cnt = gpu.DeviceArray.from_numpy(np.array([len(pairs)], np.uint32), s)
    bound = max(len(pairs) + 1000, 2 * len(pairs))
    dp2 = gpu.DeviceArray.from_numpy(np.concatenate([pairs.ravel(), np.zeros(2 * (bound - len(pairs)), np.uint32)]),
                                     s)
    np_.set_features(True)
    np_.run_device_count(s, dp2.ptr, cnt.ptr, bound, db, ds)
    got2 = np_.contacts(s)
    assert got2["n"] == nc_ref
    for f, _ in gpu.CONTACT_FIELDS:
        assert np.array_equal(bits(got2[f]), bits(ref[f])), f
    assert np.array_equal(np_.features(s), feat_ref)


def _contacts_for(oracle, sc, ob, osh):
    w = oracle.World.for_scene(sc)
    n_pairs = w.build_and_query(ob, osh, sc.bounds)
    w.sort_pairs(n_pairs)
    w.narrowphase(w.pairs(n_pairs), ob, osh)
    return w


@pytest.mark.parametrize("case", ["stress_f90", "stress_f140", "mixed_3000", "mixed_3000_f30", "c2_32768_f40"])
@pytest.mark.parametrize("force_serial", [False, True])
def test_solver_bitexact(gpu, oracle, stage_inputs, case, force_serial):
    sc, ob0, osh = stage_inputs[case]
    print(f"[solver] {case} serial={force_serial}", flush=True)
    ob = ob0.copy()
    w = _contacts_for(oracle, sc, ob, osh)
    c = w.contacts()
    st_ref = w.solve(ob, sc.dt, force_serial=force_serial)
    colors_ref = w.colors(c["n"]) if not st_ref["used_serial"] else None

    s = gpu.Stream()
    db = _gpu_bodies(gpu, sc, s, ob0.state())
    ds = gpu.DeviceShapes(sc, s)
    np_ = gpu.Narrowphase(s, sc.max_contacts)
    pairs = np.stack([c["body_a"], c["body_b"]], 1).astype(U32)
    # Feed the identical contact list (bit-identical narrowphase is tested above).
    dp = gpu.DeviceArray.from_numpy(pairs.ravel(), s)
    np_.run(s, dp.ptr, len(pairs), db, ds)
    assert np_.download_count(s) == c["n"]
    solver = gpu.XpbdSolver(s, sc.n_iters, sc.max_contacts, sc.mu)
    solver.set_force_serial(force_serial)
    solver.solve(s, np_.contacts_view(), db, sc.dt)
    st = solver.stats()
    assert st["used_serial"] == st_ref["used_serial"]
    if colors_ref is not None:
        assert st["n_classes"] == st_ref["n_colors"]
        assert np.array_equal(solver.classes(s, c["n"]), colors_ref)
    got = db.download(s)
    ref = ob.state()
    for f in ("pos_x", "pos_y", "pos_z", "rot_w", "rot_x", "rot_y", "rot_z",
              "vel_x", "vel_y", "vel_z", "ang_x", "ang_y", "ang_z"):
        assert np.array_equal(bits(got[f]), bits(ref[f])), f


def test_c2_frames_match_oracle(gpu, oracle, stage_inputs):
    """The 32 768-body workload, frames 40..43 through the whole frame step
    (integrate, broadphase, device-count narrowphase, coloured solve): pair and
    contact counts and the state hash equal the oracle's after every frame."""
    sc, ob0, osh = stage_inputs["c2_32768_f40"]
    ob = ob0.copy()
    ow = oracle.World.for_scene(sc)
    gw = gpu.World(sc)
    gw.bodies.upload_state(ob0.state(), gw.s)
    for f in range(3):
        print(f"[c2 frames] frame {40 + f}", flush=True)
        ost = ow.frame(ob, osh, sc)
        gst = gw.step()
        assert (gst["n_pairs"], gst["n_contacts"], gst["n_colors"]) == \
            (ost["n_pairs"], ost["n_contacts"], ost["n_colors"]), f
        assert gw.fnv() == ob.fnv(), f"frame {40 + f}"
    assert gst["n_contacts"] > 5000


def test_golden_box_hash_on_gpu(gpu):
    # core/tests/sim_determinism_test.cpp:224 (OMP golden), 20 frames.
    sc = scenes.determinism_boxes()
    hs = []
    for _ in range(2):
        w = gpu.World(sc)
        for _ in range(20):
            w.step()
        hs.append(w.fnv())
    assert hs[0] == hs[1] == 0xCBCD209C3665C818


@pytest.mark.parametrize("sweep_mode", ["2", "2-single", "3", "0"])
def test_stress_frames_match_oracle(gpu, oracle, monkeypatch, sweep_mode):
    """150 frames of the stress scene against the oracle under every sweep:
    the dataflow sweep with a pair of lanes per record (the default) and with
    one lane per record, the per-lane-cursor form and the class phases."""
    monkeypatch.setenv("DYPHUR_SWEEP_MODE", sweep_mode.split("-")[0])
    if sweep_mode.endswith("-single"):
        monkeypatch.setenv("DYPHUR_SWEEP_PAIR", "0")
    print(f"[stress 150 frames] sweep mode {sweep_mode}", flush=True)
    sc = scenes.stress_blocks()
    frames = 150
    b, s = oracle.Bodies(sc), oracle.Shapes(sc)
    ow = oracle.World.for_scene(sc)
    gw = gpu.World(sc)
    for f in range(frames):
        ost = ow.frame(b, s, sc)
        gst = gw.step()
        assert gst["n_pairs"] == ost["n_pairs"], f
        assert gst["n_contacts"] == ost["n_contacts"], f
        if f % 10 == 9 or f == frames - 1:
            assert gw.fnv() == b.fnv(), f"frame {f}"


def test_coloured_run_to_run_identical(gpu):
    sc = scenes.random_mixed(4000, seed=5, extent=(14, 12, 14))
    hs = []
    for _ in range(2):
        w = gpu.World(sc)
        for _ in range(40):
            w.step()
        hs.append(w.fnv())
    assert hs[0] == hs[1]


@pytest.mark.parametrize("n", [1, 63, 1024, 4097, 32769, 36864, 36865, 65536, 65537, 200003])
def test_exclusive_scan_sizes(gpu, n):
    """dy_scan_exclusive_u32 over every path (LDS-staged workgroup up to
    36 864 values, tiled workgroup up to 65 536, three launches beyond): the
    exclusive prefix sums and the total, exact (values small enough that no
    sum wraps)."""
    print(f"[scan] n={n}", flush=True)
    rng = np.random.default_rng(n)
    a = rng.integers(0, 64, size=n, dtype=np.uint32)
    s = gpu.Stream()
    din = gpu.DeviceArray.from_numpy(a, s)
    dout = gpu.DeviceArray(n, np.uint32)
    dtot = gpu.DeviceArray.from_numpy(np.array([0xDEADBEEF], np.uint32), s)
    gpu.check(gpu.lib().dy_scan_exclusive_u32(s.h, gpu.vp(din.ptr), gpu.vp(dout.ptr), n, gpu.vp(dtot.ptr)),
              "dy_scan_exclusive_u32")
    got = dout.download(s)
    ref = np.concatenate([[0], np.cumsum(a, dtype=np.uint64)[:-1]]).astype(np.uint32)
    assert np.array_equal(got, ref)
    assert int(dtot.download(s)[0]) == int(a.sum(dtype=np.uint64))


def test_cooperative_launch_path_identical(gpu, monkeypatch):
    """The grid-synchronised kernels (Morton sort, class assignment, sweep)
    give the same frames launched as ordinary launches (the default) and
    through the runtime's cooperative path (DYPHUR_COOP_LAUNCH=1)."""
    sc = scenes.random_mixed(12000, seed=11, extent=(24, 14, 24))
    hs = []
    for coop in ("0", "1"):
        monkeypatch.setenv("DYPHUR_COOP_LAUNCH", coop)
        print(f"[launch path] cooperative={coop}", flush=True)
        w = gpu.World(sc)
        for _ in range(30):
            w.step()
        hs.append(w.fnv())
    assert hs[0] == hs[1]


def test_colour_overflow_falls_back_to_serial(gpu, oracle):
    """301 contacts sharing one dynamic plate: JP runs out of its 64 colours and
    the reference solves in contact order; the GPU must take the same path."""
    sc = scenes.star_plate()
    b, s = oracle.Bodies(sc), oracle.Shapes(sc)
    ow = oracle.World.for_scene(sc)
    gw = gpu.World(sc)
    for f in range(4):
        print(f"[star_plate] frame {f}", flush=True)
        ost = ow.frame(b, s, sc)
        gst = gw.step()
        assert (gst["n_pairs"], gst["n_contacts"]) == (ost["n_pairs"], ost["n_contacts"]), f
        assert bool(gst["used_serial"]) == bool(ost["used_serial"]), f
        if f == 0:
            assert gst["n_contacts"] > 256 and gst["used_serial"]
            assert gw.solver.stats()["overflow"] == 1
        assert gw.fnv() == b.fnv(), f"frame {f}"


LAUNCH_SHAPES = [
    ("per_round_launches", {"DYPHUR_SOLVER_PATH": "1"}),
    ("wg1024", {"DYPHUR_ASSIGN_THREADS": "1024", "DYPHUR_SWEEP_THREADS": "1024"}),
    ("one_block", {"DYPHUR_ASSIGN_BLOCKS": "1", "DYPHUR_SWEEP_BLOCKS": "1"}),
    ("sweep_phases_whole_grid", {"DYPHUR_SWEEP_MODE": "1"}),
    ("sweep_phases_one_xcd", {"DYPHUR_SWEEP_MODE": "0"}),
    ("sweep_phases_one_block", {"DYPHUR_SWEEP_MODE": "0", "DYPHUR_SWEEP_BLOCKS": "1"}),
    ("sweep_dataflow_lane_cursors", {"DYPHUR_SWEEP_MODE": "3"}),
    ("sweep_dataflow_lane_cursors_one_block", {"DYPHUR_SWEEP_MODE": "3", "DYPHUR_SWEEP_BLOCKS": "1"}),
]


@pytest.mark.parametrize("shape", [s[0] for s in LAUNCH_SHAPES])
@pytest.mark.parametrize("force_serial", [False, True])
def test_solver_launch_shapes_bitexact(gpu, oracle, stage_inputs, monkeypatch, shape, force_serial):
    """The class assignment and sweeps give the same bits under every launch
    shape the solver supports (one launch per round/class, 1024-thread
    workgroups, a single workgroup, and the class-phase sweeps over one XCD or
    the whole grid beside the default dataflow sweep)."""
    for k, v in dict(LAUNCH_SHAPES)[shape].items():
        monkeypatch.setenv(k, v)
    sc, ob0, osh = stage_inputs["mixed_3000_f30"]
    print(f"[solver shape] {shape} serial={force_serial}", flush=True)
    ob = ob0.copy()
    w = _contacts_for(oracle, sc, ob, osh)
    c = w.contacts()
    st_ref = w.solve(ob, sc.dt, force_serial=force_serial)
    s = gpu.Stream()
    db = _gpu_bodies(gpu, sc, s, ob0.state())
    ds = gpu.DeviceShapes(sc, s)
    np_ = gpu.Narrowphase(s, sc.max_contacts)
    pairs = np.stack([c["body_a"], c["body_b"]], 1).astype(U32)
    dp = gpu.DeviceArray.from_numpy(pairs.ravel(), s)
    np_.run(s, dp.ptr, len(pairs), db, ds)
    solver = gpu.XpbdSolver(s, sc.n_iters, sc.max_contacts, sc.mu)
    solver.set_force_serial(force_serial)
    solver.solve(s, np_.contacts_view(), db, sc.dt)
    st = solver.stats()
    assert st["used_serial"] == st_ref["used_serial"]
    if not st_ref["used_serial"]:
        assert np.array_equal(solver.classes(s, c["n"]), w.colors(c["n"]))
    got, ref = db.download(s), ob.state()
    for f in ("pos_x", "pos_y", "pos_z", "rot_w", "rot_x", "rot_y", "rot_z",
              "vel_x", "vel_y", "vel_z", "ang_x", "ang_y", "ang_z"):
        assert np.array_equal(bits(got[f]), bits(ref[f])), f


@pytest.mark.parametrize("sort_blocks", ["0", "1", "7", "32", "64"])
def test_morton_sort_workgroup_counts(gpu, oracle, stage_inputs, monkeypatch, sort_blocks):
    """The Morton sort gives the reference order with one workgroup (0) and
    with the cooperative multi-workgroup form at any number of participating
    workgroups (on one XCD)."""
    monkeypatch.setenv("DYPHUR_SORT_BLOCKS", sort_blocks)
    sc, ob, osh = stage_inputs["mixed_12000"]
    print(f"[morton sort] blocks={sort_blocks} n={sc.n}", flush=True)
    w = oracle.World.for_scene(sc)
    n_ref = w.build_and_query(ob, osh, sc.bounds)
    ref = w.bvh(sc.n)
    s = gpu.Stream()
    db = _gpu_bodies(gpu, sc, s, ob.state())
    ds = gpu.DeviceShapes(sc, s)
    bp = gpu.Broadphase(s, sc.n, sc.max_pairs)
    bp.build_and_query(s, db, ds, sc.bounds)
    assert bp.download_count(s) == n_ref
    got = bp.bvh(s, sc.n)
    assert np.array_equal(got["morton"], ref["morton"])
    assert np.array_equal(got["sorted_idx"], ref["sorted_idx"])


@pytest.mark.parametrize("variant", [{}, {"DYPHUR_EMIT_COOP": "0"}, {"DYPHUR_TRAVERSE_TEAM": "0"},
                                     {"DYPHUR_TRAVERSE_TEAM": "16"}, {"DYPHUR_TEAM_STACK": "3"}])
def test_broadphase_long_rows(gpu, oracle, monkeypatch, variant):
    """Dense clusters, so rows run past the 16 kept in registers and past
    16 + 128 (the team traversal's partner list; the emission's re-walks);
    every traversal and emission variant gives the reference's pair list.  A
    3-entry team stack sends most queries to the per-lane fallback walk,
    whose overflow entries the team then orders."""
    for k, v in variant.items():
        monkeypatch.setenv(k, v)
    for n_dyn, seed, ext in ((400, 7, (2.0, 2.0, 2.0)), (1500, 8, (6.0, 5.0, 6.0))):
        sc = scenes.random_mixed(n_dyn, seed=seed, extent=ext)
        sc.max_pairs = sc.n * sc.n
        ob, osh = oracle.Bodies(sc), oracle.Shapes(sc)
        w = oracle.World.for_scene(sc)
        n_ref = w.build_and_query(ob, osh, sc.bounds)
        w.sort_pairs(n_ref)
        ref_pairs = w.pairs(n_ref)
        rows = np.bincount(ref_pairs[:, 0] if ref_pairs.ndim == 2 else ref_pairs["a"], minlength=sc.n)
        print(f"[long rows] {variant} n={sc.n} pairs={n_ref} longest row {rows.max()} "
              f"rows 17-64: {int(((rows > 16) & (rows <= 64)).sum())} 65-144: {int(((rows > 64) & (rows <= 144)).sum())} "
              f"> 144: {int((rows > 144).sum())}", flush=True)
        s = gpu.Stream()
        db = _gpu_bodies(gpu, sc, s, ob.state())
        ds = gpu.DeviceShapes(sc, s)
        bp = gpu.Broadphase(s, sc.n, sc.max_pairs)
        bp.build_and_query(s, db, ds, sc.bounds)
        n = bp.download_count(s)
        assert n == n_ref
        assert np.array_equal(bp.pairs(s, n), ref_pairs)
        if n_dyn == 400:
            assert rows.max() > 144 and ((rows > 64) & (rows <= 144)).any() and ((rows > 16) & (rows <= 64)).any()


def test_broadphase_100k_bodies(gpu, oracle):
    """100 001 bodies: more than 1 024 tree nodes cross the 1 024-leaf refit
    chunks, so the top of the tree is merged by refit_top's global-memory
    climb instead of its LDS form; tree, boxes and pairs equal the
    reference's."""
    sc = scenes.random_mixed(100000, seed=21, extent=(60.0, 40.0, 60.0))
    ob, osh = oracle.Bodies(sc), oracle.Shapes(sc)
    w = oracle.World.for_scene(sc)
    n_ref = w.build_and_query(ob, osh, sc.bounds)
    ref = w.bvh(sc.n)
    w.sort_pairs(n_ref)
    ref_pairs = w.pairs(n_ref)
    s = gpu.Stream()
    db = _gpu_bodies(gpu, sc, s, ob.state())
    ds = gpu.DeviceShapes(sc, s)
    bp = gpu.Broadphase(s, sc.n, sc.max_pairs)
    bp.build_and_query(s, db, ds, sc.bounds)
    n = bp.download_count(s)
    got = bp.bvh(s, sc.n)
    for k in ("morton", "sorted_idx", "left", "right", "parent", "aabb"):
        assert np.array_equal(got[k], ref[k]), k
    assert n == n_ref
    assert np.array_equal(bp.pairs(s, n), ref_pairs)
