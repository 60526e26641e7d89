"""The kernels' per-element arithmetic (dy_math.hpp, compiled for the host)
against the oracle, bit for bit.  This runs without a GPU; the GPU parity
tests then check that the device build of the same code agrees."""
import math

import numpy as np
import pytest

from tests import _hostmath as hmod

F32 = np.float32
RNG = np.random.default_rng(1234)


@pytest.fixture(scope="module")
def hm():
    return hmod.lib()


def _rand_quat(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return q.astype(F32)


def test_half_conversion_matches_oracle_and_numpy(hm, oracle):
    x = np.concatenate([
        RNG.normal(scale=30, size=200000), RNG.uniform(-70000, 70000, 50000),
        RNG.uniform(-1e-4, 1e-4, 50000), np.array([0.0, -0.0, 65504, 65519.99, 65520, 1e-8, 6e-8, 3e-8]),
    ]).astype(F32)
    # Exact ties between adjacent halves.
    h = np.arange(0, 0x7bff, 7, dtype=np.uint16).view(np.float16).astype(np.float32)
    nxt = (np.arange(0, 0x7bff, 7, dtype=np.uint16) + 1).view(np.float16).astype(np.float32)
    x = np.concatenate([x, ((h.astype(np.float64) + nxt) / 2).astype(F32)])
    a = np.zeros(len(x), np.uint16)
    b = np.zeros(len(x), np.uint16)
    hm.hm_f32_to_f16_array(x.ctypes.data, a.ctypes.data, len(x))
    oracle.lib().rc_half_from_float_array(x.ctypes.data, b.ctypes.data, len(x))
    assert np.array_equal(a, b)
    assert np.array_equal(a, x.astype(np.float16).view(np.uint16))


def test_morton_and_quantize(hm, oracle):
    L = oracle.lib()
    for v in range(1024):
        for (x, y, z) in ((v, 0, 0), (0, v, 0), (0, 0, v), (v, 1023 - v, v ^ 0x2aa)):
            assert hm.hm_morton3(x, y, z) == L.rc_morton3(x, y, z)
    p = RNG.uniform(-30, 30, 20000).astype(F32)
    for val in p[:2000]:
        q = hm.hm_quantize10(val, F32(-25), F32(25))
        t = (F32(val) - F32(-25)) / (F32(25) - F32(-25))
        t = min(max(t, F32(0)), F32(1))
        assert q == int(np.float32(t) * F32(1023))


def _keys(n, rng):
    codes = rng.integers(0, 1 << 30, n, dtype=np.uint64)
    keys = (codes << np.uint64(32)) | np.arange(n, dtype=np.uint64)
    return np.sort(keys)


def test_karras_nodes(hm, oracle):
    L = oracle.lib()
    for n in (2, 3, 5, 17, 100, 1000):
        keys = _keys(n, RNG)
        a = np.zeros(5, np.int32)
        b = np.zeros(5, np.int32)
        for i in range(n - 1):
            hm.hm_karras_node(keys.ctypes.data, n, i, a.ctypes.data)
            L.rc_karras_node(keys.ctypes.data, n, i, b.ctypes.data)
            assert a[:3].tolist() == b[:3].tolist()


def test_leaf_extent(hm, oracle):
    L = oracle.lib()
    q = _rand_quat(RNG, 3000)
    he = RNG.uniform(0.05, 2.0, (3000, 3)).astype(F32)
    a, b = np.zeros(3, F32), np.zeros(3, F32)
    for i in range(3000):
        for t in (0, 1):
            hq, hh = np.ascontiguousarray(q[i]), np.ascontiguousarray(he[i])
            hm.hm_leaf_extent(t, hh.ctypes.data, hq.ctypes.data, a.ctypes.data)
            L.rc_leaf_extent(t, hh.ctypes.data, hq.ctypes.data, b.ctypes.data)
            assert a.view(np.uint32).tolist() == b.view(np.uint32).tolist()


def _pair_cases(rng, n):
    """Random and structured pair poses: axis-aligned stacks, random rotations,
    near-touching and deep overlaps, tiny and large boxes."""
    cases = []
    for i in range(n):
        kind = i % 4
        ta, tb = [(0, 0), (1, 1), (0, 1), (1, 0)][i % 4] if i % 8 < 4 else (0, 0)
        ha = rng.uniform(0.1, 0.6, 3).astype(F32)
        hb = rng.uniform(0.1, 0.6, 3).astype(F32)
        if ta == 1:
            ha[:] = ha[0]
        if tb == 1:
            hb[:] = hb[0]
        pa = rng.uniform(-2, 2, 3).astype(F32)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        dist = rng.uniform(0.0, 1.4) * (float(np.max(ha)) + float(np.max(hb)))
        pb = (pa + d * dist).astype(F32)
        if kind == 0:  # stacked, axis aligned
            qa = np.array([1, 0, 0, 0], F32)
            qb = np.array([1, 0, 0, 0], F32)
            pb = (pa + np.array([rng.uniform(-0.1, 0.1), ha[1] + hb[1] - rng.uniform(-0.05, 0.1),
                                 rng.uniform(-0.1, 0.1)])).astype(F32)
        elif kind == 1:  # slightly tilted
            ang = rng.normal(scale=0.05, size=(2, 3))
            qa = np.array([1, *ang[0] / 2], F32); qa /= np.linalg.norm(qa)
            qb = np.array([1, *ang[1] / 2], F32); qb /= np.linalg.norm(qb)
            qa, qb = qa.astype(F32), qb.astype(F32)
        else:
            qa, qb = _rand_quat(rng, 2)
        cases.append((ta, tb, ha, hb, pa, pb, qa, qb))
    return cases


def test_pair_contacts(hm, oracle):
    """Contact values and feature words (kind, axes, incident-vertex mask,
    SAT axis) of the kernels' arithmetic equal the oracle's bit for bit."""
    L = oracle.lib()
    hits = 0
    kinds = set()
    a, b = np.zeros(7, F32), np.zeros(7, F32)
    fa, fb = np.zeros(1, np.uint32), np.zeros(1, np.uint32)
    for (ta, tb, ha, hb, pa, pb, qa, qb) in _pair_cases(RNG, 20000):
        args = [np.ascontiguousarray(v).ctypes.data for v in (ha, hb, pa, pb, qa, qb)]
        a[:] = 0
        b[:] = 0
        ra = hm.hm_pair_contact_feat(ta, tb, *args, a.ctypes.data, fa.ctypes.data)
        rb = L.rc_pair_contact_feat(ta, tb, *args, b.ctypes.data, fb.ctypes.data)
        assert ra == rb
        if ra:
            hits += 1
            assert a.view(np.uint32).tolist() == b.view(np.uint32).tolist()
            assert fa[0] == fb[0]
            kinds.add(int(fa[0]) & 7)
    assert hits > 5000
    assert kinds == {0, 1, 2, 3, 4, 5}


def _rand_body(rng, dynamic=True):
    q = _rand_quat(rng, 1)[0]
    q = (q * rng.uniform(0.98, 1.02)).astype(F32)  # un-normalised, as mid-solve
    body = np.zeros(20, F32)
    body[0:3] = rng.uniform(-1, 1, 3)
    body[3:7] = q
    body[7:10] = rng.normal(size=3)
    body[10:13] = rng.normal(size=3)
    if dynamic:
        body[13] = rng.uniform(0.2, 3)
        body[14:17] = rng.uniform(0.5, 20, 3)
        body[17:20] = rng.normal(scale=0.3, size=3)
    return body


def test_project_contact(hm, oracle):
    L = oracle.lib()
    for i in range(20000):
        n = RNG.normal(size=3)
        n /= np.linalg.norm(n)
        c = np.concatenate([RNG.uniform(-1, 1, 3), n, [RNG.uniform(-0.01, 0.2)]]).astype(F32)
        A = _rand_body(RNG, dynamic=(i % 5 != 0))
        B = _rand_body(RNG, dynamic=(i % 7 != 0))
        lam0 = F32(RNG.uniform(0, 0.1)) if i % 3 == 0 else F32(0)
        mu = F32(0.5) if i % 4 else F32(0)
        Ah, Bh, Ao, Bo = A.copy(), B.copy(), A.copy(), B.copy()
        lh = np.array([lam0], F32)
        lo = np.array([lam0], F32)
        hm.hm_project_contact(c.ctypes.data, Ah.ctypes.data, Bh.ctypes.data, mu, lh.ctypes.data, F32(1 / 60))
        L.rc_project_contact(c.ctypes.data, Ao.ctypes.data, Bo.ctypes.data, mu, lo.ctypes.data, F32(1 / 60))
        assert Ah.view(np.uint32).tolist() == Ao.view(np.uint32).tolist(), i
        assert Bh.view(np.uint32).tolist() == Bo.view(np.uint32).tolist(), i
        assert lh.view(np.uint32)[0] == lo.view(np.uint32)[0]


def test_integrate_body(hm, oracle):
    from dyphur_amd import scenes
    n = 500
    sc = scenes.two_body(0, 0, (1, 1, 1), (1, 1, 1), (0, 0, 0), (0, 0, 0))
    b = scenes._alloc(n)
    b["pos"][:] = RNG.uniform(-5, 5, (n, 3))
    b["rot"][:] = _rand_quat(RNG, n)
    b["vel"][:] = RNG.normal(size=(n, 3))
    b["ang"][:] = RNG.normal(scale=3, size=(n, 3))
    for k, v in b.items():
        setattr(sc, k, v)
    for sub in (1, 4):
        ob = oracle.Bodies(sc)
        oracle.integrate(ob, (0, -9.81, 0), 1 / 60, sub)
        for i in range(n):
            kin = np.concatenate([sc.pos[i], sc.rot[i], sc.vel[i], sc.ang[i]]).astype(F32)
            hm.hm_integrate_body(kin.ctypes.data, F32(0), F32(-9.81), F32(0), F32(1 / 60) / F32(sub), sub)
            ref = np.array([ob.arr[k][i] for k in ("pos_x", "pos_y", "pos_z", "rot_w", "rot_x", "rot_y",
                                                   "rot_z", "vel_x", "vel_y", "vel_z")], F32)
            assert kin[:10].view(np.uint32).tolist() == ref.view(np.uint32).tolist()


def test_pair_projection_matches_project_contact(hm):
    """The sweep's two-lane projection (project_contact_pair: each lane one
    body, the shared sums formed in the same order from exchanged terms)
    gives project_contact's lambda and both bodies bit for bit, on 3 000
    random contacts (static second bodies and accumulated lambdas among
    them); the two lanes run as two host threads meeting at every exchange."""
    assert hm.hm_pair_projection_mismatches(3000, 20260007) == 0
