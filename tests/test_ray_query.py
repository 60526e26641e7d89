"""Ray queries over the LBVH (core/src/ray_query.cpp, core/include/core/ray_query.hpp).

CPU: the oracle against the reference's known answers
(core/tests/ray_query_test.cpp:44-181) and against a brute force over all
bodies (the closest exact hit among bodies whose binary16 leaf box the ray
enters); GPU: dy_ray_query bit for bit against the oracle.
"""
import numpy as np
import pytest

from dyphur_amd import scenes

F32 = np.float32
NO_HIT = 0xFFFFFFFF


def _ray(o, d, t_max=1e9):
    return np.array(list(o) + list(d) + [t_max], F32)


def _two(shape, half, pa, pb, bounds):
    sc = scenes.two_body(shape, shape, half, half, pa, pb)
    sc.bounds = np.array(bounds, F32)
    return sc


KATS = [
    # (scene, rays, expected (body, t) per ray; t None for a miss)
    ("sphere_x", lambda: _two(1, (1, 1, 1), (0, 0, 0), (0, 50, 0), (-6, -6, -6, 56, 56, 56)),
     [_ray((-5, 0, 0), (1, 0, 0)), _ray((-5, 3, 0), (1, 0, 0))], [(0, 4.0), (NO_HIT, None)]),
    ("box_x", lambda: _two(0, (1, 1, 1), (0, 0, 0), (0, 50, 0), (-6, -6, -6, 56, 56, 56)),
     [_ray((-5, 0, 0), (1, 0, 0)), _ray((-5, 2, 0), (1, 0, 0))], [(0, 4.0), (NO_HIT, None)]),
    ("closest_of_two", lambda: _two(1, (1, 1, 1), (0, 0, 0), (10, 0, 0), (-6, -6, -6, 16, 6, 6)),
     [_ray((-5, 0, 0), (1, 0, 0))], [(0, 4.0)]),
    ("t_max", lambda: _two(1, (1, 1, 1), (0, 0, 0), (50, 0, 0), (-6, -6, -6, 56, 6, 6)),
     [_ray((-5, 0, 0), (1, 0, 0), 2.0)], [(NO_HIT, None)]),
]


def _oracle_query(oracle, sc, rays):
    b, s = oracle.Bodies(sc), oracle.Shapes(sc)
    w = oracle.World.for_scene(sc)
    w.build_and_query(b, s, sc.bounds)
    return w.ray_query(b, s, sc.n, np.stack(rays)), w, b, s


@pytest.mark.parametrize("case", [k[0] for k in KATS])
def test_oracle_ray_query_reference_answers(oracle, case):
    name, mk, rays, want = next(k for k in KATS if k[0] == case)
    (t, body), *_ = _oracle_query(oracle, mk(), rays)
    for i, (wb, wt) in enumerate(want):
        assert body[i] == wb
        if wt is not None:
            assert abs(t[i] - wt) <= 1e-3


def _slab(o, inv, mn, mx):
    t1, t2 = (mn - o) * inv, (mx - o) * inv
    lo, hi = np.minimum(t1, t2), np.maximum(t1, t2)
    return lo.max(-1), hi.min(-1)


def _brute(sc, st, aabb16, sorted_idx, ray):
    """Closest exact hit among all bodies whose binary16 leaf box the ray
    enters before t_max, in float32 with the reference's operation order."""
    n = sc.n
    o, d, t_max = ray[:3], ray[3:6], ray[6]
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        inv = F32(1) / d
        box = aabb16[n - 1:, :6].copy().view(np.float16).astype(F32)
        lo, hi = _slab(o, inv, box[:, :3], box[:, 3:])
        enter = (hi >= 0) & (hi >= lo) & (lo < t_max)
        best_t, best = t_max, NO_HIT
        for leaf in np.nonzero(enter)[0]:
            i = int(sorted_idx[leaf])
            shp = int(sc.shape[i])
            p = np.array([st["pos_x"][i], st["pos_y"][i], st["pos_z"][i]], F32)
            h = sc.shape_half[shp].astype(F32)
            if sc.shape_type[shp] == scenes.SHAPE_SPHERE:
                oc = o - p
                a = F32(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
                b = F32(d[0] * oc[0] + d[1] * oc[1] + d[2] * oc[2])
                c = F32(F32(oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2]) - h[0] * h[0])
                disc = F32(b * b - a * c)
                if disc < 0:
                    continue
                sq = np.sqrt(disc)
                t = F32((-b - sq) / a)
                if t < 0:
                    t = F32((-b + sq) / a)
            else:
                qw, qx, qy, qz = (F32(st[k][i]) for k in ("rot_w", "rot_x", "rot_y", "rot_z"))
                u = np.array([-qx, -qy, -qz], F32)

                def rot(v):
                    tv = np.array([u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2],
                                   u[0] * v[1] - u[1] * v[0]], F32) * F32(2)
                    return (v + tv * qw) + np.array([u[1] * tv[2] - u[2] * tv[1], u[2] * tv[0] - u[0] * tv[2],
                                                     u[0] * tv[1] - u[1] * tv[0]], F32)

                lo_, ld = rot((o - p).astype(F32)), rot(d.astype(F32))
                il = F32(1) / ld
                t1, t2 = (-h - lo_) * il, (h - lo_) * il
                tmin, tmax = np.minimum(t1, t2).max(), np.maximum(t1, t2).min()
                if tmax < 0 or tmin > tmax:
                    continue
                t = tmin if tmin >= 0 else tmax
            if 0 <= t < best_t:
                best_t, best = F32(t), i
    return best_t, best


def test_oracle_ray_query_equals_brute_force(oracle):
    sc = scenes.random_mixed(800, seed=17, extent=(10, 6, 10))
    rng = np.random.default_rng(3)
    rays = []
    for _ in range(300):
        o = rng.uniform([-8, 0, -8], [8, 12, 8])
        tgt = rng.uniform([-5, 1, -5], [5, 7, 5])
        rays.append(_ray(o, (tgt - o) / np.linalg.norm(tgt - o), rng.uniform(3, 30)))
    (t, body), w, b, s = _oracle_query(oracle, sc, rays)
    bvh = w.bvh(sc.n)
    st = b.state()
    hits = 0
    for k, r in enumerate(rays):
        bt, bb = _brute(sc, st, bvh["aabb"], bvh["sorted_idx"], r)
        assert body[k] == bb, k
        assert t[k].view(np.uint32) == F32(bt).view(np.uint32), k
        hits += bb != NO_HIT
    assert hits > 100


@pytest.mark.gpu
@pytest.mark.parametrize("case", [k[0] for k in KATS] + ["mixed_3000"])
def test_gpu_ray_query_bitexact(gpu, oracle, case):
    if case == "mixed_3000":
        sc = scenes.random_mixed(3000, seed=11, extent=(12, 10, 12))
        rng = np.random.default_rng(8)
        rays = []
        for _ in range(20000):
            o = rng.uniform([-10, 0, -10], [10, 14, 10])
            tgt = rng.uniform([-6, 1, -6], [6, 10, 6])
            rays.append(_ray(o, tgt - o, rng.uniform(0.5, 3.0)))
    else:
        _, mk, rays, _ = next(k for k in KATS if k[0] == case)
        sc = mk()
    print(f"[ray] {case} n={sc.n} rays={len(rays)}", flush=True)
    (t_ref, b_ref), *_ = _oracle_query(oracle, sc, rays)
    s = gpu.Stream()
    db = gpu.DeviceBodies(sc, s)
    ds = gpu.DeviceShapes(sc, s)
    bp = gpu.Broadphase(s, sc.n, sc.max_pairs)
    bp.build_and_query(s, db, ds, sc.bounds)
    t, body = gpu.ray_query(s, bp, sc.n, db, ds, np.stack(rays))
    assert np.array_equal(body, b_ref)
    assert np.array_equal(t.view(np.uint32), t_ref.view(np.uint32))
