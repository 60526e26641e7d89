"""The allocation-free narrowphase entry point (dy_narrowphase_run_noalloc)
against dy_narrowphase_run_device_count through the C ABI: the same contacts,
counts and feature words bit for bit, on two calls in a row, the second with
a smaller input (its scratch keeps the first call's contents)."""
import ctypes as C

import numpy as np
import pytest

from dyphur_amd import scenes

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _noalloc(gpu, s, dp, dcount, bound, db, ds, cap):
    """Runs the entry point into a store of our own; returns (counts, contacts, features)."""
    arrs = {f: gpu.DeviceArray(cap, dt) for f, dt in gpu.CONTACT_FIELDS}
    counts = gpu.DeviceArray(5, np.uint32)
    feat = gpu.DeviceArray(cap, np.uint32)
    cv = gpu.ContactView()
    for f, _ in gpu.CONTACT_FIELDS:
        setattr(cv, f, arrs[f].ptr)
    cv.n = None
    cv.capacity = cap
    bv, sv = db.view(), ds.view()
    gpu.check(gpu.lib().dy_narrowphase_run_noalloc(s.h, gpu.vp(dp.ptr), gpu.vp(dcount.ptr), bound, C.byref(bv),
                                                   C.byref(sv), C.byref(cv), gpu.vp(counts.ptr), gpu.vp(feat.ptr)),
              "dy_narrowphase_run_noalloc")
    c = counts.download(s)
    n = int(c[0])
    out = {f: arrs[f].download(s, n) for f, _ in gpu.CONTACT_FIELDS}
    return c, out, feat.download(s, n)


def _reference(gpu, s, dp, dcount, bound, db, ds, cap):
    np_ = gpu.Narrowphase(s, cap)
    np_.set_features(True)
    np_.run_device_count(s, dp.ptr, dcount.ptr, bound, db, ds)
    return np_.download_count(s), np_.last_count(s), np_.contacts(s), np_.features(s)


@pytest.mark.parametrize("clamped", [False, True])
def test_noalloc_matches_device_count_run(gpu, oracle, clamped):
    s = gpu.Stream()
    # first call: a settled 4 000-body mixed scene; second: a smaller stack
    inputs = []
    for sc, frames in ((scenes.random_mixed(4000, seed=31, extent=(14, 10, 14)), 30),
                       (scenes.stress_blocks(4, 6, 4), 40)):
        ob, osh = oracle.Bodies(sc), oracle.Shapes(sc)
        ow = oracle.World.for_scene(sc)
        for _ in range(frames):
            ow.frame(ob, osh, sc)
        n_pairs = ow.build_and_query(ob, osh, sc.bounds)
        ow.sort_pairs(n_pairs)
        inputs.append((sc, ob, ow.pairs(n_pairs)))
    assert len(inputs[1][2]) < len(inputs[0][2])
    for k, (sc, ob, pairs) in enumerate(inputs):
        print(f"[no-alloc] call {k}: {len(pairs)} pairs, store clamped: {clamped}", flush=True)
        db = gpu.DeviceBodies(sc, s)
        db.upload_state(ob.state(), s)
        ds = gpu.DeviceShapes(sc, s)
        bound = len(pairs) + 777
        dp = gpu.DeviceArray.from_numpy(np.concatenate([pairs.ravel(), np.zeros(2 * 777, np.uint32)]), s)
        dcount = gpu.DeviceArray.from_numpy(np.array([len(pairs)], np.uint32), s)
        cap = max(16, len(pairs) // 4) if clamped else sc.max_contacts
        stored, true, ref, ref_feat = _reference(gpu, s, dp, dcount, bound, db, ds, cap)
        c, got, feat = _noalloc(gpu, s, dp, dcount, bound, db, ds, cap)
        assert (int(c[0]), int(c[1]), int(c[2]), int(c[3])) == (stored, true, 0, 0)
        if clamped:
            assert true > cap == stored  # capacity clamps the store, keeps the true count
        for f, _ in gpu.CONTACT_FIELDS:
            assert np.array_equal(_bits(got[f]), _bits(ref[f])), f
        assert np.array_equal(feat, ref_feat)


def test_noalloc_rejects_oversized_bound(gpu):
    s = gpu.Stream()
    sc = scenes.stress_blocks(2, 2, 2)
    db, ds = gpu.DeviceBodies(sc, s), gpu.DeviceShapes(sc, s)
    cap = 16
    arrs = {f: gpu.DeviceArray(cap, dt) for f, dt in gpu.CONTACT_FIELDS}
    counts = gpu.DeviceArray(5, np.uint32)
    cv = gpu.ContactView()
    for f, _ in gpu.CONTACT_FIELDS:
        setattr(cv, f, arrs[f].ptr)
    cv.capacity = cap
    dp = gpu.DeviceArray(2, np.uint32)
    bv, sv = db.view(), ds.view()
    st = gpu.lib().dy_narrowphase_run_noalloc(s.h, gpu.vp(dp.ptr), None, (1 << 24) + 1, C.byref(bv), C.byref(sv),
                                              C.byref(cv), gpu.vp(counts.ptr), None)
    assert st == gpu.DY_ERR_INVALID_ARG
