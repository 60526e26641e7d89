"""Batched scenes (dy_batch_*) vs the CPU oracle, scene by scene.

Every scene of a batch must evolve exactly as a stand-alone world over the
same bodies does in the oracle (examples/stress_test_blocks/main.cpp:264-273
per frame): the pair list, the contact list, the solver path and the body
state, bit for bit, after every frame.  The batch mixes small piles
(reference-order solves), a large settled pile and a dense mixed scene
(coloured solves), so both solver paths and the contact capacity are used.
"""
import numpy as np
import pytest

from dyphur_amd import scenes

pytestmark = pytest.mark.gpu

STATE = ("pos_x", "pos_y", "pos_z", "rot_w", "rot_x", "rot_y", "rot_z", "vel_x", "vel_y", "vel_z",
         "ang_x", "ang_y", "ang_z")
CONTACT = ("body_a", "body_b", "pos_x", "pos_y", "pos_z", "norm_x", "norm_y", "norm_z", "depth")


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _mixed_batch():
    rng = scenes.Pcg32(20261207)
    counts = scenes.lognormal_counts(14, 48.0, 0.75, 8, 160, rng)
    out = [scenes.pile_scene(int(c), rng) for c in counts]
    out.append(scenes.pile_scene(300, rng))
    out.append(scenes.random_mixed(500, seed=31, extent=(6, 5, 6)))
    out.append(scenes.pile_scene(8, rng))
    return out


def _valid_pairs(batch):
    """Each scene's stored pairs (the rest of the buffer is stale)."""
    allp, st = batch.pairs_all(), batch.stats()
    out = []
    for i, sc in enumerate(batch.scenes):
        p0 = int(batch.pair_offset[i])
        out.append(allp[p0:p0 + min(int(st["n_pairs"][i]), sc.max_pairs)])
    return out


def _same_pairs(a, b):
    pa, pb = _valid_pairs(a), _valid_pairs(b)
    return all(np.array_equal(x, y) for x, y in zip(pa, pb))


def _oracle_worlds(oracle, scs):
    return [(oracle.World.for_scene(sc), oracle.Bodies(sc), oracle.Shapes(sc)) for sc in scs]


def _compare_frame(batch, ref, frame, stats=None):
    got = batch.state()
    pairs = batch.pairs_all()
    con = batch.contacts_all()
    st = batch.stats() if stats is None else stats
    for i, (w, b, _sh, fs) in enumerate(ref):
        sl = batch.scene_slice(i)
        tag = f"frame {frame} scene {i} ({batch.scenes[i].name}, n={batch.scenes[i].n})"
        assert st["error"][i] == 0, tag
        assert st["n_pairs"][i] == fs["n_pairs"], tag
        assert st["n_contacts"][i] == w.contacts()["last_count"], tag
        assert st["used_serial"][i] == fs["used_serial"], tag
        if not fs["used_serial"]:
            assert st["n_classes"][i] == fs["n_colors"], tag
        npairs = min(int(fs["n_pairs"]), batch.scenes[i].max_pairs)
        p0 = int(batch.pair_offset[i])
        assert np.array_equal(pairs[p0:p0 + npairs], w.pairs(npairs)), tag
        rc = w.contacts()
        c0 = int(batch.contact_offset[i])
        for k in CONTACT:
            assert np.array_equal(bits(con[k][c0:c0 + rc["n"]]), bits(rc[k])), (tag, k)
        ref_state = b.state()
        for k in STATE:
            assert np.array_equal(bits(got[k][sl]), bits(ref_state[k])), (tag, k)


def _run_oracle_frame(ref, scs, force_serial=False):
    out = []
    for (w, b, sh), sc in zip(ref, scs):
        fs = w.frame(b, sh, sc, force_serial)
        out.append((w, b, sh, fs))
    return out


@pytest.mark.parametrize("shape", [(64, 0), (256, 128), (128, 64)])
def test_batch_matches_oracle_every_frame(gpu, oracle, shape):
    scs = _mixed_batch()
    print(f"[batch] {len(scs)} scenes shape={shape} frames=40", flush=True)
    batch = gpu.Batch(scs)
    batch.set_shape(*shape)
    ref = _oracle_worlds(oracle, scs)
    saw_jp = saw_serial = False
    sums = np.zeros((len(scs), 4), np.int64)
    for f in range(40):
        batch.step(1)
        fr = _run_oracle_frame(ref, scs)
        _compare_frame(batch, fr, f)
        saw_jp |= any(not x[3]["used_serial"] and x[3]["n_contacts"] > 0 for x in fr)
        saw_serial |= any(x[3]["used_serial"] for x in fr)
        for i, (w, b, _sh, fs) in enumerate(fr):
            c = w.contacts()
            im = b.arr["inv_mass"]
            sums[i] += (fs["n_pairs"], c["n"],
                        int((im[c["body_a"]] != 0).sum() + (im[c["body_b"]] != 0).sum()), int(c["n"] > 0))
    assert saw_jp and saw_serial
    st = batch.stats()
    assert np.array_equal(st["sum_pairs"], sums[:, 0])
    assert np.array_equal(st["sum_contacts"], sums[:, 1])
    assert np.array_equal(st["sum_dyn_endpoints"], sums[:, 2])
    assert np.array_equal(st["solves"], sums[:, 3])
    assert (st["substeps"] == 40).all()


@pytest.mark.parametrize("lds_limit", [0, 65536])
def test_batch_large_and_global_scenes_match_oracle(gpu, oracle, lds_limit):
    """A 700-body dense scene (78 KB of state) runs in the launch of its own
    with the device's LDS limit, and from global memory at 64 KiB, beside
    small piles in the LDS-resident launch; every frame equals the oracle."""
    rng = scenes.Pcg32(20261208)
    scs = [scenes.random_mixed(700, seed=41, extent=(6, 5, 6)), scenes.pile_scene(40, rng),
           scenes.pile_scene(12, rng)]
    print(f"[batch] 700-body scene beside piles, lds limit {lds_limit}", flush=True)
    batch = gpu.Batch(scs)
    batch.set_lds_limit(lds_limit)
    ref = _oracle_worlds(oracle, scs)
    for f in range(10):
        batch.step(1)
        _compare_frame(batch, _run_oracle_frame(ref, scs), f)


def test_batch_lds_limit_below_64k_reads_as_64k(gpu):
    """A limit under 64 KiB is taken as 64 KiB: the same launch plan, so the
    same state bits as an explicit 64 KiB after the same frames."""
    rng = scenes.Pcg32(20261210)
    scs = [scenes.random_mixed(700, seed=47, extent=(6, 5, 6)), scenes.pile_scene(20, rng)]
    got = []
    for limit in (1, 65536):
        b = gpu.Batch(scs)
        b.set_lds_limit(limit)
        b.step(4)
        got.append(b.state())
    for k in STATE:
        assert np.array_equal(bits(got[0][k]), bits(got[1][k])), k


def test_batch_grid_pairs_match_oracle(gpu, oracle):
    """A sparse 600-body scene stepped by a workgroup finds its pairs on the
    x-z grid (the statistics count those substeps) and still equals the
    oracle's pair list, contacts and state every frame; a compact pile of the
    same batch tests every pair."""
    rng = scenes.Pcg32(20261209)
    scs = [scenes.random_mixed(600, seed=43, extent=(20, 4, 20)), scenes.pile_scene(300, rng),
           scenes.pile_scene(30, rng)]
    batch = gpu.Batch(scs)
    batch.set_shape(256, 128)
    ref = _oracle_worlds(oracle, scs)
    for f in range(8):
        batch.step(1)
        _compare_frame(batch, _run_oracle_frame(ref, scs), f)
    st = batch.stats()
    print(f"[batch] grid substeps per scene: {list(st['grid_substeps'])}", flush=True)
    assert st["grid_substeps"][0] == 8
    assert st["grid_substeps"][2] == 0  # 31 bodies: a wavefront tests every pair


def test_batch_force_serial_matches_oracle(gpu, oracle):
    scs = _mixed_batch()[-3:]
    batch = gpu.Batch(scs)
    batch.set_force_serial(True)
    ref = _oracle_worlds(oracle, scs)
    for f in range(12):
        batch.step(1)
        fr = _run_oracle_frame(ref, scs, force_serial=True)
        _compare_frame(batch, fr, f)


def test_batch_one_launch_many_frames_equals_frame_by_frame(gpu):
    scs = _mixed_batch()
    a, b = gpu.Batch(scs), gpu.Batch(scs)
    a.step(25)
    for _ in range(25):
        b.step(1)
    sa, sb = a.state(), b.state()
    for k in STATE:
        assert np.array_equal(bits(sa[k]), bits(sb[k])), k
    assert _same_pairs(a, b)
    ta, tb = a.stats(), b.stats()
    for k in ta:
        assert np.array_equal(ta[k], tb[k]), k


def test_batch_launch_shapes_agree_and_autotune_restores(gpu):
    scs = _mixed_batch()
    ref = gpu.Batch(scs)
    ref.set_shape(128, 128)
    ref.step(10)
    want = ref.state()
    tuned = gpu.Batch(scs)
    before = tuned.state()
    chosen, ms = tuned.autotune(3)
    assert chosen in gpu.Batch.TUNE_SHAPES and all(m > 0 for m in ms)
    assert tuned.shape() == chosen
    after = tuned.state()
    for k in STATE:
        assert np.array_equal(bits(before[k]), bits(after[k])), k
    assert int(tuned.stats()["substeps"].max()) == 0
    tuned.step(10)
    got = tuned.state()
    for k in STATE:
        assert np.array_equal(bits(got[k]), bits(want[k])), k
    for shp in gpu.Batch.TUNE_SHAPES:
        other = gpu.Batch(scs)
        other.set_shape(*shp)
        other.step(10)
        st = other.state()
        for k in STATE:
            assert np.array_equal(bits(st[k]), bits(want[k])), (shp, k)


def test_batch_stage_timing_leaves_results_unchanged(gpu):
    scs = _mixed_batch()
    plain, timed = gpu.Batch(scs), gpu.Batch(scs)
    timed.set_timing(True)
    plain.step(6)
    timed.step(6)
    sp, st = plain.state(), timed.state()
    for k in STATE:
        assert np.array_equal(bits(sp[k]), bits(st[k])), k
    tm = timed.timing()
    assert all(v.shape == (len(scs),) for v in tm.values())
    assert (tm["pairs"] > 0).all() and tm["sweep"].sum() > 0


def test_batch_graph_replay_equals_eager(gpu):
    scs = _mixed_batch()[:8]
    eager, graphed = gpu.Batch(scs), gpu.Batch(scs, s=None)
    with gpu.Graph.capture(graphed.s) as g:
        graphed.step(1)
        graphed.step(1)
    for rep in range(4):
        eager.step(1)
        eager.step(1)
        g.launch(graphed.s)
        graphed.s.wait()
        se, sg = eager.state(), graphed.state()
        for k in STATE:
            assert np.array_equal(bits(se[k]), bits(sg[k])), (rep, k)
        assert _same_pairs(eager, graphed), rep


def test_batch_pair_and_contact_capacity_clamp_in_order(gpu):
    sc = scenes.random_mixed(400, seed=41, extent=(4, 3, 4))
    wide = gpu.Batch([sc])
    wide.step(1)
    st_w = wide.stats()
    n_pairs, n_con = int(st_w["n_pairs"][0]), int(st_w["n_contacts"][0])
    assert n_pairs > 100 and n_con > 40
    tight_sc = scenes.random_mixed(400, seed=41, extent=(4, 3, 4))
    tight_sc.max_pairs = n_pairs // 2
    tight = gpu.Batch([tight_sc])
    tight.step(1)
    st_t = tight.stats()
    assert int(st_t["n_pairs"][0]) == n_pairs
    pw, pt = wide.pairs_all(), tight.pairs_all()
    assert np.array_equal(pt[:n_pairs // 2], pw[:n_pairs // 2])
    # contacts of the kept pairs: the wide run's contacts whose pair is kept
    cw, ct = wide.contacts_all(), tight.contacts_all()
    kept = {tuple(p) for p in pw[:n_pairs // 2].tolist()}
    sel = [j for j in range(n_con) if (int(cw["body_a"][j]), int(cw["body_b"][j])) in kept]
    assert int(st_t["n_contacts"][0]) == len(sel)
    for k in CONTACT:
        assert np.array_equal(bits(ct[k][:len(sel)]), bits(cw[k][sel])), k


def test_batch_contact_capacity_keeps_first_contacts(gpu):
    sc = scenes.random_mixed(400, seed=41, extent=(4, 3, 4))
    wide = gpu.Batch([sc])
    wide.step(1)
    n_con = int(wide.stats()["n_contacts"][0])
    tight_sc = scenes.random_mixed(400, seed=41, extent=(4, 3, 4))
    tight_sc.max_contacts = n_con // 3
    tight = gpu.Batch([tight_sc])
    tight.step(1)
    st = tight.stats()
    assert int(st["n_contacts"][0]) == n_con
    cw, ct = wide.contacts_all(), tight.contacts_all()
    for k in CONTACT:
        assert np.array_equal(bits(ct[k][:n_con // 3]), bits(cw[k][:n_con // 3])), k


def test_batch_rejects_hull_shapes_loudly(gpu):
    sc = scenes.two_body(0, 0, (0.5, 0.5, 0.5), (0.5, 0.5, 0.5), (0, 0, 0), (0.8, 0, 0))
    sc.shape_type = np.array([0, 2], np.uint32)  # ConvexHull on body 1
    b = gpu.Batch([sc, scenes.pile_scene(12, scenes.Pcg32(5))])
    b.step(1)
    with pytest.raises(gpu.DyError) as e:
        b.stats()
    assert e.value.status == gpu.DY_ERR_UNSUPPORTED
    st = b.stats(raise_on_error=False)
    assert int(st["error"][0]) & 1 and int(st["error"][1]) == 0
