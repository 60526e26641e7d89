"""Pins of the CPU oracle by independent means (brute force, closed forms,
invariants, and a direct simulation of the reference's round rule), chosen so
that a dropped term, a swapped index or a wrong comparison in the oracle's
restatement fails here.  None of these re-types the oracle's own code path:
each check computes the expected value another way.
"""
import numpy as np
import pytest

from dyphur_amd import scenes

F32 = np.float32


def _scene_at(oracle, sc, frames):
    b, s = oracle.Bodies(sc), oracle.Shapes(sc)
    w = oracle.World.for_scene(sc)
    for _ in range(frames):
        w.frame(b, s, sc)
    return b, s


@pytest.fixture(scope="module")
def mixed(oracle):
    sc = scenes.random_mixed(1500, seed=21, extent=(10, 8, 10))
    b, s = _scene_at(oracle, sc, 25)
    w = oracle.World.for_scene(sc)
    n_pairs = w.build_and_query(b, s, sc.bounds)
    return sc, b, s, w, n_pairs


def _spread(v):
    out = 0
    for bit in range(10):
        out |= ((v >> bit) & 1) << (3 * bit)
    return out


# core/src/broadphase.cpp:21-33,117-150: key = morton30(q(p)) << 32 | index,
# q = (u32)(clamp01((p - min) / (max - min)) * 1023.f), x at bit 0; sorted
# ascending.  Recomputed here in numpy float32 and a bit loop, sorted by numpy.
def test_morton_keys_and_order_brute_force(oracle, mixed):
    sc, b, _, w, _ = mixed
    n = sc.n
    bvh = w.bvh(n)
    lo = sc.bounds[:3].astype(F32)
    ext = (sc.bounds[3:] - sc.bounds[:3]).astype(F32)
    keys = []
    for i in range(n):
        p = np.array([b.arr["pos_x"][i], b.arr["pos_y"][i], b.arr["pos_z"][i]], F32)
        t = np.clip((p - lo) / ext, F32(0), F32(1)).astype(F32)
        q = (t * F32(1023.0)).astype(F32).astype(np.uint32)
        m = _spread(int(q[0])) | (_spread(int(q[1])) << 1) | (_spread(int(q[2])) << 2)
        keys.append((m << 32) | i)
    ks = np.sort(np.array(keys, np.uint64))
    assert np.array_equal(bvh["morton"], ks)
    assert np.array_equal(bvh["sorted_idx"], (ks & np.uint64(0xFFFFFFFF)).astype(np.uint32))


def _clz64(x):
    return 64 - int(x).bit_length()


# core/src/broadphase.cpp:35-76,152-192: Karras numbering of the binary radix
# tree over distinct sorted keys.  Built here top-down by recursion: a node
# over [lo, hi] splits after the last key that shares a longer prefix with
# keys[lo] than keys[hi] does.
def test_karras_tree_recursive_brute_force(oracle, mixed):
    sc, _, _, w, _ = mixed
    n = sc.n
    bvh = w.bvh(n)
    k = [int(v) for v in bvh["morton"]]
    left = np.full(n - 1, -2, np.int64)
    right = np.full(n - 1, -2, np.int64)
    parent = np.full(2 * n - 1, -2, np.int64)
    parent[0] = -1
    stack = [(0, 0, n - 1)]
    while stack:
        node, lo, hi = stack.pop()
        common = _clz64(k[lo] ^ k[hi])
        g = lo
        for i in range(lo, hi):
            if _clz64(k[lo] ^ k[i]) > common:
                g = i
        lch = (n - 1 + g) if g == lo else g
        rch = (n - 1 + g + 1) if g + 1 == hi else g + 1
        left[node], right[node] = lch, rch
        parent[lch] = node
        parent[rch] = node
        if g != lo:
            stack.append((g, lo, g))
        if g + 1 != hi:
            stack.append((g + 1, g + 1, hi))
    assert np.array_equal(bvh["left"], left)
    assert np.array_equal(bvh["right"], right)
    assert np.array_equal(bvh["parent"], parent)


def _boxes_f32(aabb16):
    return aabb16[:, :6].copy().view(np.float16).astype(np.float64)


# core/src/broadphase.cpp:215-288 + core/include/core/broadphase.hpp:21-28:
# leaf boxes are the fp16 rounding of p +- r (spheres) and p +- |R| h (boxes);
# internal boxes are the exact union of their children.  Leaves are checked
# against the closed form evaluated in float64 (within one binary16 step);
# unions exactly.
def test_leaf_boxes_closed_form_and_refit_union(oracle, mixed):
    sc, b, _, w, _ = mixed
    n = sc.n
    bvh = w.bvh(n)
    box = _boxes_f32(bvh["aabb"])
    for slot in range(n):
        i = int(bvh["sorted_idx"][slot])
        p = np.array([b.arr[f][i] for f in ("pos_x", "pos_y", "pos_z")], np.float64)
        shp = int(sc.shape[i])
        h = sc.shape_half[shp].astype(np.float64)
        if sc.shape_type[shp] == scenes.SHAPE_SPHERE:
            e = np.array([h[0]] * 3)
        else:
            qw, qx, qy, qz = (float(b.arr[f][i]) for f in ("rot_w", "rot_x", "rot_y", "rot_z"))
            R = np.array([[1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qw * qz), 2 * (qx * qz + qw * qy)],
                          [2 * (qx * qy + qw * qz), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qw * qx)],
                          [2 * (qx * qz - qw * qy), 2 * (qy * qz + qw * qx), 1 - 2 * (qx * qx + qy * qy)]])
            e = np.abs(R) @ h
        got = box[n - 1 + slot]
        want = np.concatenate([p - e, p + e])
        tol = np.maximum(np.abs(want) * 2.0 ** -10, 2.0 ** -24) + 1e-5 * (1 + np.abs(want))
        assert (np.abs(got - want) <= tol).all(), (slot, got, want)
    for node in range(n - 1):
        l, r = int(bvh["left"][node]), int(bvh["right"][node])
        u = np.concatenate([np.minimum(box[l, :3], box[r, :3]), np.maximum(box[l, 3:], box[r, 3:])])
        assert np.array_equal(box[node], u), node


# core/src/broadphase.cpp:290-365: the pair set is every pair of leaves whose
# fp16 boxes overlap (touching counts), as (min, max) body index, ascending.
def test_pairs_equal_all_overlapping_leaf_boxes(oracle, mixed):
    sc, _, _, w, n_pairs = mixed
    n = sc.n
    bvh = w.bvh(n)
    box = _boxes_f32(bvh["aabb"])[n - 1:]
    body = bvh["sorted_idx"].astype(np.int64)
    mn, mx = box[:, :3], box[:, 3:]
    want = []
    for i in range(n):
        sep = (mn[i + 1:] > mx[i]).any(1) | (mx[i + 1:] < mn[i]).any(1)
        for j in np.nonzero(~sep)[0] + i + 1:
            a, c = body[i], body[j]
            want.append((min(a, c), max(a, c)))
    want.sort()
    w.sort_pairs(n_pairs)
    got = w.pairs(n_pairs)
    assert n_pairs == len(want)
    assert [tuple(x) for x in got.tolist()] == want


# core/src/narrowphase.cpp:1077-1127: contacts come out in pair order (one
# per touching pair), i.e. the order of a sequential pass over the pairs.
def test_contact_order_is_pair_order(oracle, mixed):
    sc, b, s, w, n_pairs = mixed
    w.sort_pairs(n_pairs)
    pairs = w.pairs(n_pairs)
    nc = w.narrowphase(pairs, b, s)
    c = w.contacts()
    assert nc == c["n"] and nc > 0
    pos = {tuple(p): k for k, p in enumerate(pairs.tolist())}
    idx = [pos[(int(a), int(bb))] for a, bb in zip(c["body_a"], c["body_b"])]
    assert idx == sorted(idx) and len(set(idx)) == len(idx)


def _fmix32(x):
    x = np.asarray(x, np.uint64) & np.uint64(0xFFFFFFFF)
    x ^= x >> np.uint64(16)
    x = (x * np.uint64(0x85EBCA6B)) & np.uint64(0xFFFFFFFF)
    x ^= x >> np.uint64(13)
    x = (x * np.uint64(0xC2B2AE35)) & np.uint64(0xFFFFFFFF)
    x ^= x >> np.uint64(16)
    return x


# core/src/xpbd_solver.cpp:31-34,625-695, simulated round by round as the
# reference states it: every uncoloured k posts its priority (hash(k) << 32 |
# k, hash = Murmur3 fmix32) to its dynamic endpoints with atomic max; a k that
# holds the max at all its dynamic endpoints takes the lowest bit clear in
# mask[a] | mask[b] and sets it in both masks.
def test_jp_colours_match_round_simulation(oracle, mixed):
    sc, b, s, w, n_pairs = mixed
    w.sort_pairs(n_pairs)
    nc = w.narrowphase(w.pairs(n_pairs), b, s)
    assert nc > 256
    c = w.contacts()
    im = b.arr["inv_mass"]
    A, B = c["body_a"].astype(np.int64), c["body_b"].astype(np.int64)
    pri = (_fmix32(np.arange(nc)) << np.uint64(32)) | np.arange(nc, dtype=np.uint64)
    ends = [[x for x in (A[k], B[k]) if im[x] != 0] for k in range(nc)]
    colour = np.full(nc, -1, np.int64)
    mask = np.zeros(sc.n, np.uint64)
    left = set(range(nc))
    rounds = 0
    while left and rounds < 256:
        rounds += 1
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
    assert not left and colour.max() < 64
    bb = b.copy()
    st = w.solve(bb, sc.dt)
    assert st["used_serial"] == 0
    assert np.array_equal(w.colors(nc).astype(np.int64), colour)
    # proper colouring: no dynamic body sees one colour twice
    for x in range(sc.n):
        ks = [k for k in range(nc) if x in ends[k]]
        assert len({int(colour[k]) for k in ks}) == len(ks)


# core/src/xpbd_solver.cpp:413-521 with mu = 0: a head-on contact between two
# equal spheres changes their velocities by equal and opposite impulses, so
# linear momentum is conserved and the pair stops approaching.
def test_solver_conserves_momentum_without_friction(oracle):
    sc = scenes.two_body(1, 1, (0.5, 0.5, 0.5), (0.5, 0.5, 0.5), (0, 0, 0), (0.9, 0, 0),
                         vel_a=(1.0, 0.2, 0), vel_b=(-1.0, 0.2, 0), inertia=(0.1, 0.1))
    sc.mu = 0.0
    b, s = oracle.Bodies(sc), oracle.Shapes(sc)
    w = oracle.World(2, 32, 16, n_iters=4, mu=0.0)
    assert w.narrowphase(np.array([[0, 1]], np.uint32), b, s) == 1
    p0 = np.array([b.arr["vel_x"].sum(), b.arr["vel_y"].sum(), b.arr["vel_z"].sum()], np.float64)
    w.solve(b, 1 / 60)
    p1 = np.array([b.arr["vel_x"].sum(), b.arr["vel_y"].sum(), b.arr["vel_z"].sum()], np.float64)
    assert np.allclose(p0, p1, atol=1e-6)
    assert b.arr["vel_x"][0] - b.arr["vel_x"][1] <= 1e-6  # no longer approaching along x
