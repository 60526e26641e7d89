"""Derivatives of the per-element arithmetic (dyphur_amd/csrc/adjoint.hpp),
compiled for the host: the templated replays instantiated at float give the
forward's exact bits, and the hand-written adjoints and the dual-number
contact Jacobian agree with central finite differences in float64.

Pins: the forward maps are the ones the kernels run (bit equality with
dy_math.hpp's float functions, which the oracle pins); the derivatives are
checked against differences of those same maps evaluated in double, on
random inputs away from branch boundaries."""
import ctypes as C

import numpy as np
import pytest

from tests import _hostmath

F32, F64 = np.float32, np.float64
vp = C.c_void_p
BOX, SPHERE = 0, 1


@pytest.fixture(scope="module")
def hm():
    L = _hostmath.lib()
    d, f, i = C.c_double, C.c_float, C.c_int
    L.ha_pair_contact_f.restype = i
    L.ha_pair_contact_f.argtypes = [C.c_uint32, C.c_uint32] + [vp] * 8
    L.ha_pair_contact_d.restype = i
    L.ha_pair_contact_d.argtypes = [C.c_uint32, C.c_uint32] + [vp] * 8
    L.ha_contact_adj_d.restype = i
    L.ha_contact_adj_d.argtypes = [C.c_uint32, C.c_uint32] + [vp] * 9 + [i]
    L.ha_pair_feat_f.restype = i
    L.ha_pair_feat_f.argtypes = [C.c_uint32, C.c_uint32] + [vp] * 7
    L.ha_pair_feat_d.restype = i
    L.ha_pair_feat_d.argtypes = [C.c_uint32, C.c_uint32] + [vp] * 7
    L.ha_geom_adj_d.restype = None
    L.ha_geom_adj_d.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32] + [vp] * 9
    L.ha_geom_adj_f.restype = None
    L.ha_geom_adj_f.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32] + [vp] * 9
    L.hm_pair_contact_feat.restype = i
    L.hm_pair_contact_feat.argtypes = [C.c_uint32, C.c_uint32] + [vp] * 8
    L.ha_contact_adj_f.restype = i
    L.ha_contact_adj_f.argtypes = [C.c_uint32, C.c_uint32] + [vp] * 9
    L.ha_project_f.restype = i
    L.ha_project_f.argtypes = [vp, vp, vp, f, vp, f]
    L.ha_project_d.restype = i
    L.ha_project_d.argtypes = [vp, vp, vp, d, vp, d]
    L.ha_project_adj_d.argtypes = [vp, vp, vp, d, d, d, vp, vp, vp, vp]
    L.ha_integrate_f.argtypes = [vp, vp, f]
    L.ha_integrate_d.argtypes = [vp, vp, d]
    L.ha_integrate_adj_d.argtypes = [vp, vp, d, vp]
    L.ha_renorm_d.argtypes = [vp]
    L.ha_renorm_adj_d.argtypes = [vp, vp]
    return L


def P(a):
    return a.ctypes.data


def _quat(rng):
    q = rng.normal(size=4)
    return q / np.linalg.norm(q)


def _pair(rng, ta, tb):
    ha = rng.uniform(0.1, 0.5, 3)
    hb = rng.uniform(0.1, 0.5, 3)
    if ta == SPHERE:
        ha[:] = ha[0]
    if tb == SPHERE:
        hb[:] = hb[0]
    pa = rng.uniform(-0.3, 0.3, 3)
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    pb = pa + d * (ha.max() + hb.max()) * rng.uniform(0.5, 1.1)
    return ha, hb, pa, pb, _quat(rng), _quat(rng)


TYPES = [(BOX, BOX), (SPHERE, BOX), (BOX, SPHERE), (SPHERE, SPHERE)]


def test_contact_replay_float_equals_forward_bits(hm):
    rng = np.random.default_rng(1)
    hits = 0
    for it in range(3000):
        ta, tb = TYPES[it % 4]
        ha, hb, pa, pb, qa, qb = (x.astype(F32) for x in _pair(rng, ta, tb))
        o1, o2 = np.zeros(7, F32), np.zeros(7, F32)
        w = C.c_int(-1)
        r1 = hm.hm_pair_contact(ta, tb, P(ha), P(hb), P(pa), P(pb), P(qa), P(qb), P(o1))
        r2 = hm.ha_pair_contact_f(ta, tb, P(ha), P(hb), P(pa), P(pb), P(qa), P(qb), P(o2), C.byref(w))
        assert r1 == r2
        if r1:
            hits += 1
            assert np.array_equal(o1.view(np.uint32), o2.view(np.uint32)), (ta, tb, o1, o2)
    assert hits > 1000


def _body(rng, static=False):
    b = np.zeros(20)
    b[0:3] = rng.uniform(-0.5, 0.5, 3)
    b[3:7] = _quat(rng) * rng.uniform(0.98, 1.02)
    b[7:10] = rng.normal(0, 1.0, 3)
    b[10:13] = rng.normal(0, 1.0, 3)
    if not static:
        b[13] = rng.uniform(0.5, 2.0)
        d = rng.uniform(0.5, 3.0, 3)
        b[14:17] = d
        b[17:20] = rng.uniform(-0.1, 0.1, 3)
    return b


def _contact(rng, a, b):
    p = 0.5 * (a[0:3] + b[0:3]) + rng.normal(0, 0.05, 3)
    n = a[0:3] - b[0:3]
    n /= np.linalg.norm(n)
    return np.concatenate([p, n, [rng.uniform(0.005, 0.05)]])


def test_project_t_float_equals_project_contact_bits(hm):
    rng = np.random.default_rng(2)
    for it in range(2000):
        a = _body(rng).astype(F32)
        b = _body(rng, static=(it % 5 == 0)).astype(F32)
        c = _contact(rng, a, b).astype(F32)
        lam = F32(rng.uniform(0, 0.01)) if it % 3 else F32(0)
        a1, b1, a2, b2 = a.copy(), b.copy(), a.copy(), b.copy()
        l1, l2 = np.array([lam], F32), np.array([lam], F32)
        mu = F32(0.5 if it % 4 else 0.0)
        r1 = hm.hm_project_contact(P(c), P(a1), P(b1), mu, P(l1), F32(1 / 60))
        r2 = hm.ha_project_f(P(c), P(a2), P(b2), mu, P(l2), F32(1 / 60))
        assert r1 == r2
        for x, y in ((a1, a2), (b1, b2), (l1, l2)):
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), it


def test_integrate_t_float_equals_integrate_body_bits(hm):
    rng = np.random.default_rng(3)
    g = np.array([0, -9.81, 0], F32)
    for _ in range(1000):
        k = np.concatenate([rng.uniform(-1, 1, 3), _quat(rng), rng.normal(0, 1, 3), rng.normal(0, 2, 3)]).astype(F32)
        k1, k2 = k.copy(), k.copy()
        hm.hm_integrate_body(P(k1), g[0], g[1], g[2], F32(1 / 60), 1)
        hm.ha_integrate_f(P(k2), P(g), F32(1 / 60))
        assert np.array_equal(k1.view(np.uint32), k2.view(np.uint32))


def _fd(f, x, eps):
    """Central differences of scalar f at x, one coordinate at a time."""
    g = np.zeros_like(x)
    for i in range(len(x)):
        xp, xm = x.copy(), x.copy()
        xp[i] += eps
        xm[i] -= eps
        g[i] = (f(xp) - f(xm)) / (2 * eps)
    return g


def _close(got, want, tol):
    scale = max(1.0, np.abs(want).max())
    assert np.abs(got - want).max() <= tol * scale, (got, want)


def _project_case(hm, rng, static_b, mu, cap):
    """One projection with output weights; returns (loss(x), adjoint, x)."""
    dt = 1 / 60
    for _ in range(200):
        a = _body(rng)
        b = _body(rng, static=static_b)
        c = _contact(rng, a, b)
        lam = rng.uniform(0, 0.004)
        if cap:
            a[7:13] *= 0.05  # small tangential speed: friction impulse capped by mu * lambda
        x = np.concatenate([a[:14], b[:14], c, [lam]])
        wa, wb, wl = rng.normal(size=13), rng.normal(size=13), rng.normal()

        def loss(x, a=a, b=b):
            A, B = a.copy(), b.copy()
            A[:14], B[:14] = x[:14], x[14:28]
            cc = x[28:35].copy()
            lm = np.array([x[35]])
            hm.ha_project_d(P(cc), P(A), P(B), mu, P(lm), dt)
            return wa @ A[:13] + wb @ B[:13] + wl * lm[0]

        ab = np.concatenate([wa, [0.0]])
        bb = np.concatenate([wb, [0.0]])
        lb = np.array([wl])
        cb = np.zeros(7)
        hm.ha_project_adj_d(P(c), P(a), P(b), mu, lam, dt, P(ab), P(bb), P(lb), P(cb))
        adj = np.concatenate([ab, bb, cb, lb])
        if static_b:  # a static body's state gets no adjoint
            adj[14:28] = np.nan
        return loss, adj, x
    raise AssertionError("no case")


@pytest.mark.parametrize("static_b,mu,cap", [(False, 0.5, False), (False, 0.5, True), (False, 0.0, False),
                                             (True, 0.5, False), (True, 0.5, True)])
def test_project_adjoint_matches_central_differences(hm, static_b, mu, cap):
    rng = np.random.default_rng(10 + 7 * static_b + 3 * cap + int(10 * mu))
    for _ in range(25):
        loss, adj, x = _project_case(hm, rng, static_b, mu, cap)
        fd = _fd(loss, x, 1e-7)
        keep = ~np.isnan(adj)
        if static_b:
            keep[14:28] = False
        _close(adj[keep], fd[keep], 2e-5)


def test_integrate_adjoint_matches_central_differences(hm):
    rng = np.random.default_rng(4)
    g = np.array([0, -9.81, 0])
    for _ in range(50):
        k = np.concatenate([rng.uniform(-1, 1, 3), _quat(rng) * 1.01, rng.normal(0, 1, 3), rng.normal(0, 3, 3)])
        w = rng.normal(size=13)

        def loss(x):
            y = x.copy()
            hm.ha_integrate_d(P(y), P(g), 1 / 60)
            return w @ y

        bar = w.copy()
        hm.ha_integrate_adj_d(P(k), P(g), 1 / 60, P(bar))
        _close(bar, _fd(loss, k, 1e-7), 1e-6)


def test_renorm_adjoint_matches_central_differences(hm):
    rng = np.random.default_rng(5)
    for _ in range(50):
        q = _quat(rng) * rng.uniform(0.9, 1.1)
        w = rng.normal(size=4)

        def loss(x):
            y = x.copy()
            hm.ha_renorm_d(P(y))
            return w @ y

        qb = w.copy()
        hm.ha_renorm_adj_d(P(q), P(qb))
        _close(qb, _fd(loss, q, 1e-7), 1e-7)


def test_contact_geometry_adjoint_matches_central_differences(hm):
    rng = np.random.default_rng(6)
    seen = {}
    for it in range(4000):
        ta, tb = TYPES[it % 4]
        ha, hb, pa, pb, qa, qb = _pair(rng, ta, tb)
        out = np.zeros(7)
        w0 = C.c_int(-1)
        if not hm.ha_pair_contact_d(ta, tb, P(ha), P(hb), P(pa), P(pb), P(qa), P(qb), P(out), C.byref(w0)):
            continue
        kind = (ta, tb, w0.value >> 4 if (ta, tb) == (BOX, BOX) else -1)
        if seen.get(kind, 0) >= 15:
            continue
        hbar = rng.normal(size=7)
        x = np.concatenate([pa, qa, pb, qb])

        def loss(x):
            o = np.zeros(7)
            w = C.c_int(-1)
            xa, xqa, xb, xqb = x[0:3].copy(), x[3:7].copy(), x[7:10].copy(), x[10:14].copy()
            r = hm.ha_pair_contact_d(ta, tb, P(ha), P(hb), P(xa), P(xb), P(xqa), P(xqb), P(o), C.byref(w))
            assert r and w.value == w0.value
            return hbar @ o

        try:
            fd = _fd(loss, x, 1e-7)
        except AssertionError:
            continue  # the perturbation crossed a branch: not an active-set-stable point
        got = {}
        for nt in (1, 7, 14):
            ga, gb = np.zeros(7), np.zeros(7)
            assert hm.ha_contact_adj_d(ta, tb, P(ha), P(hb), P(pa), P(pb), P(qa), P(qb), P(hbar), P(ga), P(gb), nt)
            got[nt] = np.concatenate([ga, gb])
        assert np.allclose(got[1], got[14], rtol=1e-12, atol=1e-12) and np.allclose(got[7], got[14], rtol=1e-12,
                                                                                    atol=1e-12)
        _close(got[14], fd, 2e-6)
        feat = C.c_uint32(0)
        assert hm.ha_pair_feat_d(ta, tb, P(ha), P(hb), P(pa), P(pb), P(qa), P(qb), C.byref(feat))
        ga, gb = np.zeros(7), np.zeros(7)
        hm.ha_geom_adj_d(ta, tb, feat.value, P(ha), P(hb), P(pa), P(pb), P(qa), P(qb), P(hbar), P(ga), P(gb))
        _close(np.concatenate([ga, gb]), fd, 2e-6)
        seen[kind] = seen.get(kind, 0) + 1
    # every dispatch branch and both box-box kinds (face of A / of B, edge) exercised
    assert {(0, 0, 0), (0, 0, 1), (1, 0, -1), (0, 1, -1), (1, 1, -1)} <= set(seen), seen


# ---- the float64 / forward-mode frame reference used by the GPU gradient test
def _sim_args(hm, sc):
    L = hm
    vp_, d, i = C.c_void_p, C.c_double, C.c_int
    L.ha_sim_d.argtypes = [i] + [vp_] * 7 + [d, i, i, d, i, C.c_uint32, C.c_uint32]
    L.ha_sim_dir_d.argtypes = [i] + [vp_] * 9 + [d, i, i, d, i, C.c_uint32, C.c_uint32]
    from dyphur_amd import capi
    inv, iI, fl = capi.body_dynamics(sc.mass, sc.inertia, sc.flags)
    n = sc.n
    st = np.concatenate([sc.pos.T, sc.rot.T, sc.vel.T, sc.ang.T]).astype(np.float64).ravel()
    I6 = iI.T.astype(np.float64).ravel()
    stype = np.ascontiguousarray(sc.shape_type[sc.shape], np.uint32)
    hs = sc.shape_half[sc.shape].T.astype(np.float64).ravel()
    return dict(n=n, st=st, im=inv.astype(np.float64), I6=I6, stype=stype, hs=hs,
                flags=np.ascontiguousarray(fl, np.uint32), g=np.array(sc.gravity, np.float64))


def _sim(hm, a, frames, st=None, im=None):
    s = (a["st"] if st is None else st).copy()
    m = a["im"] if im is None else im
    hm.ha_sim_d(a["n"], P(s), P(m), P(a["I6"]), P(a["stype"]), P(a["hs"]), P(a["flags"]), P(a["g"]),
                1 / 60, frames, 10, 0.5, 0, 1 << 20, 1 << 20)
    return s


def test_host_reference_tracks_the_oracle(hm, oracle):
    from dyphur_amd import scenes
    sc = scenes.diff_stack(scenes.Pcg32(7))
    a = _sim_args(hm, sc)
    ob, osh = oracle.Bodies(sc), oracle.Shapes(sc)
    w = oracle.World.for_scene(sc)
    s = a["st"].copy()
    for f in range(30):
        w.frame(ob, osh, sc)
        s = _sim(hm, a, 1, st=s)
    ref = ob.state()
    n = sc.n
    got = s.reshape(13, n)
    for k, name in enumerate(("pos_x", "pos_y", "pos_z")):
        assert np.abs(got[k] - ref[name]).max() < 1e-3, name


def test_host_forward_mode_matches_central_differences(hm):
    from dyphur_amd import scenes
    sc = scenes.diff_stack(scenes.Pcg32(8))
    a = _sim_args(hm, sc)
    n = sc.n
    rng = np.random.default_rng(9)
    u_st = np.zeros(13 * n)
    u_st[7 * n:10 * n] = rng.normal(size=3 * n)  # velocity directions
    u_st[7 * n + n - 1::n] = 0.0                  # not the ground
    u_im = np.concatenate([rng.normal(size=n - 1) * 0.1, [0.0]])
    frames = 12
    st, sd = a["st"].copy(), u_st.copy()
    hm.ha_sim_dir_d(n, P(st), P(sd), P(a["im"]), P(u_im), P(a["I6"]), P(a["stype"]), P(a["hs"]),
                    P(a["flags"]), P(a["g"]), 1 / 60, frames, 10, 0.5, 0, 1 << 20, 1 << 20)
    eps = 1e-6
    sp = _sim(hm, a, frames, st=a["st"] + eps * u_st, im=a["im"] + eps * u_im)
    sm = _sim(hm, a, frames, st=a["st"] - eps * u_st, im=a["im"] - eps * u_im)
    fd = (sp - sm) / (2 * eps)
    assert np.abs(sd - fd).max() <= 1e-4 * max(1.0, np.abs(fd).max())


FEAT_KINDS = {0: "face of A", 1: "face of B", 2: "edge", 3: "sphere-sphere", 4: "sphere-box outside",
              5: "sphere-box inside"}


def _pair_any(rng, it):
    """Random touching candidates of every dispatch branch; every fifth
    sphere-box pair puts the sphere's centre inside the box."""
    ta, tb = TYPES[it % 4]
    ha, hb, pa, pb, qa, qb = _pair(rng, ta, tb)
    if (ta, tb) != (BOX, BOX) and ta != tb and (it // 4) % 5 == 0:
        box_h, box_c, box_q = (hb, pb, qb) if ta == SPHERE else (ha, pa, qa)
        w, x, y, z = box_q
        E = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y + z * w), 2 * (x * z - y * w)],
                      [2 * (x * y - z * w), 1 - 2 * (x * x + z * z), 2 * (y * z + x * w)],
                      [2 * (x * z + y * w), 2 * (y * z - x * w), 1 - 2 * (x * x + y * y)]])
        centre = box_c + E.T @ (rng.uniform(-0.9, 0.9, 3) * box_h)
        if ta == SPHERE:
            pa = centre
        else:
            pb = centre
    return ta, tb, ha, hb, pa, pb, qa, qb


def test_replay_feature_words_equal_the_forward(hm):
    """The templated replay at float reports the forward's feature word (kind,
    face or edge, passing vertices, winning axis) on every contact."""
    rng = np.random.default_rng(14)
    seen = set()
    for it in range(6000):
        ta, tb, *xs = _pair_any(rng, it)
        ha, hb, pa, pb, qa, qb = (np.asarray(x, F32) for x in xs)
        out = np.zeros(7, F32)
        f1, f2 = C.c_uint32(0), C.c_uint32(0)
        hit = hm.hm_pair_contact_feat(ta, tb, P(ha), P(hb), P(pa), P(pb), P(qa), P(qb), P(out), C.byref(f1))
        assert hm.ha_pair_feat_f(ta, tb, P(ha), P(hb), P(pa), P(pb), P(qa), P(qb), C.byref(f2)) == hit
        if hit:
            assert f1.value == f2.value, (ta, tb, hex(f1.value), hex(f2.value))
            seen.add(f1.value & 7)
    assert seen == set(FEAT_KINDS), seen


def test_geometry_adjoints_equal_dual_jacobian(hm):
    """The written-out adjoint of every contact kind, its branches taken from
    the feature word, equals the transposed dual-number Jacobian (float64)."""
    rng = np.random.default_rng(12)
    count = dict.fromkeys(FEAT_KINDS, 0)
    for it in range(8000):
        ta, tb, ha, hb, pa, pb, qa, qb = _pair_any(rng, it)
        feat = C.c_uint32(0)
        if not hm.ha_pair_feat_d(ta, tb, P(ha), P(hb), P(pa), P(pb), P(qa), P(qb), C.byref(feat)):
            continue
        hbar = rng.normal(size=7)
        g1a, g1b, g2a, g2b = np.zeros(7), np.zeros(7), np.zeros(7), np.zeros(7)
        assert hm.ha_contact_adj_d(ta, tb, P(ha), P(hb), P(pa), P(pb), P(qa), P(qb), P(hbar), P(g1a), P(g1b), 14)
        hm.ha_geom_adj_d(ta, tb, feat.value, P(ha), P(hb), P(pa), P(pb), P(qa), P(qb), P(hbar), P(g2a), P(g2b))
        want, got = np.concatenate([g1a, g1b]), np.concatenate([g2a, g2b])
        kind = feat.value & 7
        assert np.abs(got - want).max() <= 1e-10 * max(1.0, np.abs(want).max()), (FEAT_KINDS[kind], got, want)
        count[kind] += 1
    print({FEAT_KINDS[k]: v for k, v in count.items()})
    assert min(count.values()) >= 40, count


def test_geometry_adjoints_in_float_track_float64(hm):
    """At float (the kernels' precision) the written-out adjoints stay within
    float rounding of the float64 Jacobian, no further from it than the
    dual-number replay at float."""
    rng = np.random.default_rng(15)
    worst = {}
    for it in range(6000):
        ta, tb, *xs = _pair_any(rng, it)
        f32 = [np.asarray(x, F32) for x in xs]
        h32 = np.asarray(rng.normal(size=7), F32)
        feat = C.c_uint32(0)
        if not hm.ha_pair_feat_f(ta, tb, *(P(x) for x in f32), C.byref(feat)):
            continue
        fd = C.c_uint32(0)
        d64 = [x.astype(F64) for x in f32]
        if not hm.ha_pair_feat_d(ta, tb, *(P(x) for x in d64), C.byref(fd)) or fd.value != feat.value:
            continue  # float and double round to different branches here
        ga, gb = np.zeros(7), np.zeros(7)
        h64 = h32.astype(F64)
        hm.ha_contact_adj_d(ta, tb, *(P(x) for x in d64), P(h64), P(ga), P(gb), 14)
        want = np.concatenate([ga, gb])
        fa, fb = np.zeros(7, F32), np.zeros(7, F32)
        hm.ha_geom_adj_f(ta, tb, feat.value, *(P(x) for x in f32), P(h32), P(fa), P(fb))
        da, db = np.zeros(7, F32), np.zeros(7, F32)
        assert hm.ha_contact_adj_f(ta, tb, *(P(x) for x in f32), P(h32), P(da), P(db))
        scale = max(1.0, np.abs(want).max())
        e_new = np.abs(np.concatenate([fa, fb]) - want).max() / scale
        e_dual = np.abs(np.concatenate([da, db]) - want).max() / scale
        kind = feat.value & 7
        worst[kind] = max(worst.get(kind, 0.0), e_new)
        assert e_new <= max(1e-4, 4.0 * e_dual), (FEAT_KINDS[kind], e_new, e_dual)
    print({FEAT_KINDS[k]: f"{v:.1e}" for k, v in worst.items()})
    assert set(worst) == set(FEAT_KINDS), worst
