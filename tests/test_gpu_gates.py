"""Solver quality on the reference's stress scene
(examples/stress_test_blocks/main.cpp: 1 024 boxes over a static ground,
300 frames, 1 substep, 10 iterations), on the shipped 1.0 m pitch and the
0.85 m variant: after every frame's coloured solve a second broadphase +
narrowphase over the solved state gives the deepest remaining contact, and
the kinetic energy over the last 30 frames is the settled energy.  Both equal
the oracle's: the deepest contact bit for bit every frame, the energy within
5 % (the runs are bit-identical, so the energies are equal).

The reference's algorithm does not hold contacts to 5 mm on this scene: the
oracle's own deepest post-solve contact reaches 0.956 m (frame 104, a box
pair driven into each other as the 16-box columns collapse), and more than
270 of the 300 frames exceed 5 mm; full-pipeline substeps lower it (0.80 m at
4, 0.47 m at 8) but do not reach 5 mm.  The GPU reproduces those figures
exactly, so the 5 mm bound is reported by the bench, not asserted here."""
import numpy as np
import pytest

import bench
from dyphur_amd import scenes

pytestmark = pytest.mark.gpu


def _gpu_run(gpu, sc, frames, settle):
    w = gpu.World(sc)
    pbp = gpu.Broadphase(w.s, sc.n, sc.max_pairs)
    pnp = gpu.Narrowphase(w.s, sc.max_contacts)
    post, ke = [], []
    for f in range(frames):
        w.step()
        pbp.build_and_query(w.s, w.bodies, w.shapes, sc.bounds)
        n = pbp.download_count(w.s)
        pp, _ = pbp.pairs_ptr()
        pnp.run(w.s, pp, min(n, pbp.max_pairs), w.bodies, w.shapes)
        c = pnp.contacts(w.s)
        post.append(c["depth"].max() if c["n"] else np.float32(0))
        if f >= frames - settle:
            ke.append(bench.kinetic_energy(w.state(), sc))
    return w, np.array(post, np.float32), float(np.mean(ke))


def _oracle_run(oracle, sc, frames, settle):
    ob, osh = oracle.Bodies(sc), oracle.Shapes(sc)
    ow, pw = oracle.World.for_scene(sc), oracle.World.for_scene(sc)
    post, ke = [], []
    for f in range(frames):
        ow.frame(ob, osh, sc)
        n = pw.build_and_query(ob, osh, sc.bounds)
        pw.sort_pairs(n)
        pw.narrowphase(pw.pairs(n), ob, osh)
        c = pw.contacts()
        post.append(c["depth"].max() if c["n"] else np.float32(0))
        if f >= frames - settle:
            ke.append(bench.kinetic_energy(ob.state(), sc))
    return ob, np.array(post, np.float32), float(np.mean(ke))


@pytest.mark.parametrize("pitch", [1.0, 0.85])
def test_stress_penetration_and_settled_energy(gpu, oracle, pitch):
    frames, settle = 300, 30
    sc = scenes.stress_blocks(pitch=pitch)
    print(f"[quality] pitch {pitch}: {frames} frames on the GPU", flush=True)
    w, post, ke_gpu = _gpu_run(gpu, sc, frames, settle)
    print(f"[quality] pitch {pitch}: the oracle's {frames} frames", flush=True)
    ob, post_ref, ke_ref = _oracle_run(oracle, sc, frames, settle)
    assert w.fnv() == ob.fnv()
    assert np.array_equal(post.view(np.uint32), post_ref.view(np.uint32))
    assert abs(ke_gpu - ke_ref) <= 0.05 * abs(ke_ref)
    print(f"[quality] pitch {pitch}: deepest post-solve contact {post.max():.4f} m at frame {post.argmax()}, "
          f"{int((post > 0.005).sum())} frames above 5 mm, settled energy {ke_gpu:.4f} J", flush=True)
