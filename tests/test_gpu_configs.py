"""The batched and differentiable workloads at their stated sizes.

Scenes of a batch are independent, so checking a seeded sample of scenes
inside the full launch against their own oracle worlds (or host forward-mode
references) is exact for those scenes:

* the 4 096-scene pile workload (seed 20260003): every sampled scene's pairs,
  contacts, solver path and state equal its oracle world's after every frame
  of the first 10, and its state again after 40 more frames stepped in one
  launch; a second batch stepping the same 50 frames in one launch ends in
  the same state bits;
* the 256 thrown stacks over 120 frames (seed 20260004, checkpoint every 10):
  for sampled scenes the trajectory equals the host float forward mode's bit
  for bit and the reverse sweep's directional derivative equals forward
  mode's within 1e-3 over the whole horizon;
* the 1 024 mixed scenes of 16-1 024 bodies over 240 frames of 4 substeps
  (seed 20260005, checkpoint every 20, 4 contacts per body on the tapes): the
  same check for the two smallest scenes and a median one.
"""
import numpy as np
import pytest

from dyphur_amd import scenes
from tests.test_gpu_diff import _host_directional, _substep_scene, hm  # noqa: F401

pytestmark = pytest.mark.gpu

STATE = ("pos_x", "pos_y", "pos_z", "rot_w", "rot_x", "rot_y", "rot_z", "vel_x", "vel_y", "vel_z",
         "ang_x", "ang_y", "ang_z")
CONTACT = ("body_a", "body_b", "pos_x", "pos_y", "pos_z", "norm_x", "norm_y", "norm_z", "depth")


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def test_batched_piles_4096_scenes(gpu, oracle):
    scs = scenes.batched_piles()
    assert len(scs) == 4096
    rng = np.random.default_rng(20260003)
    sample = sorted(rng.choice(len(scs), 24, replace=False).tolist()) + [int(np.argmax([s.n for s in scs]))]
    print(f"[c3] {len(scs)} scenes, {sum(s.n for s in scs)} bodies; sample {sample}", flush=True)
    b = gpu.Batch(scs)
    ref = {i: (oracle.World.for_scene(scs[i]), oracle.Bodies(scs[i]), oracle.Shapes(scs[i])) for i in sample}
    frames = 10
    saw_jp = False
    for f in range(frames):
        b.step(1)
        st, got = b.stats(), b.state()
        pairs, con = b.pairs_all(), b.contacts_all()
        for i in sample:
            w, ob, osh = ref[i]
            fs = w.frame(ob, osh, scs[i])
            tag = f"frame {f} scene {i} (n={scs[i].n})"
            assert st["error"][i] == 0, tag
            assert (st["n_pairs"][i], st["used_serial"][i]) == (fs["n_pairs"], fs["used_serial"]), tag
            np_ = min(int(fs["n_pairs"]), scs[i].max_pairs)
            p0 = int(b.pair_offset[i])
            assert np.array_equal(pairs[p0:p0 + np_], w.pairs(np_)), tag
            rc = w.contacts()
            c0 = int(b.contact_offset[i])
            for k in CONTACT:
                assert np.array_equal(bits(con[k][c0:c0 + rc["n"]]), bits(rc[k])), (tag, k)
            sl = b.scene_slice(i)
            rs = ob.state()
            for k in STATE:
                assert np.array_equal(bits(got[k][sl]), bits(rs[k])), (tag, k)
            saw_jp |= not fs["used_serial"] and fs["n_contacts"] > 0
    # 40 more frames in one launch; the sampled scenes' states again
    more = 40
    b.step(more)
    got, st = b.state(), b.stats()
    for i in sample:
        w, ob, osh = ref[i]
        for _ in range(more):
            fs = w.frame(ob, osh, scs[i])
            saw_jp |= not fs["used_serial"] and fs["n_contacts"] > 0
        sl = b.scene_slice(i)
        rs = ob.state()
        tag = f"frame {frames + more - 1} scene {i} (n={scs[i].n})"
        assert st["error"][i] == 0 and st["n_pairs"][i] == fs["n_pairs"], tag
        for k in STATE:
            assert np.array_equal(bits(got[k][sl]), bits(rs[k])), (tag, k)
    assert saw_jp, "no sampled scene reached the coloured (more than 256 contacts) path"
    # run to run: a second batch stepping the same 50 frames in one launch
    b2 = gpu.Batch(scs)
    b2.step(frames + more)
    s1, s2 = b.state(), b2.state()
    for k in STATE:
        assert np.array_equal(bits(s1[k]), bits(s2[k])), k


def _grads_and_checks(gpu, hm, scs, frames, every, sample, sub=1, **tape):
    b = gpu.Batch(scs)
    d = gpu.Diff(b, frames, every, n_substeps=sub, **tape)
    d.forward(frames)
    st = b.state()
    adj = {k: np.zeros(b.n_bodies, np.float32) for k in ("pos_x", "pos_y", "pos_z")}
    for i, sc in enumerate(scs):
        sl = b.scene_slice(i)
        nd = sc.n - 1
        for k, name in enumerate(("pos_x", "pos_y", "pos_z")):
            adj[name][sl.start:sl.start + nd] = 2.0 * (st[name][sl][:nd].astype(np.float64) - sc.meta["target"][:, k])
    g = d.backward(adj)
    rng = np.random.default_rng(7)
    errs = []
    for i in sample:
        sc = scs[i]
        sl = b.scene_slice(i)
        nd = sc.n - 1
        u_v = rng.normal(size=(nd, 3))
        u_im = np.concatenate([rng.normal(size=nd) * 0.2, [0.0]])
        gv = np.stack([g["vel_x"][sl][:nd], g["vel_y"][sl][:nd], g["vel_z"][sl][:nd]], 1).astype(np.float64)
        gpu_d = float(np.sum(gv * u_v) + np.sum(g["inv_mass"][sl].astype(np.float64) * u_im))
        host_sc = _substep_scene(sc) if sub > 1 else sc
        want, f32_st = _host_directional(hm, host_sc, frames * sub, np.vstack([u_v, np.zeros((1, 3))]), u_im,
                                         hm.ha_sim_dir_f)
        for k, name in enumerate(("pos_x", "pos_y", "pos_z", "rot_w", "rot_x", "rot_y", "rot_z")):
            assert np.array_equal(bits(f32_st[k].astype(np.float32)), bits(st[name][sl])), (i, name)
        errs.append(abs(gpu_d - want) / max(abs(want), 1e-12))
        print(f"[diff at size] scene {i} ({sc.n} bodies): gpu {gpu_d:.6e} forward mode {want:.6e} "
              f"({errs[-1]:.1e})", flush=True)
    assert max(errs) <= 1e-3, errs


def test_thrown_stacks_256_scenes_120_frames(gpu, hm):  # noqa: F811
    scs = scenes.diff_batch(256)
    assert len(scs) == 256 and all(sc.n == 65 for sc in scs)
    rng = np.random.default_rng(20260004)
    sample = sorted(rng.choice(256, 4, replace=False).tolist())
    print(f"[c4] 256 stacks x 120 frames, checkpoint 10; sample {sample}", flush=True)
    _grads_and_checks(gpu, hm, scs, 120, 10, sample)


def test_mixed_1024_scenes_240_frames_4_substeps(gpu, hm):  # noqa: F811
    scs = scenes.diff_at_scale()
    assert len(scs) == 1024
    counts = np.array([sc.n for sc in scs])
    order = np.argsort(counts, kind="stable")
    sample = [int(order[0]), int(order[1]), int(order[len(order) // 2])]
    print(f"[c5] 1024 scenes ({counts.sum()} bodies, largest {counts.max()}) x 240 frames x 4 substeps; "
          f"sample {sample} of {counts[sample].tolist()} bodies", flush=True)
    _grads_and_checks(gpu, hm, scs, 240, 20, sample, sub=4, tape_contacts_per_body=4.0, tape_bytes=40 << 30)
