"""ctypes binding of the C ABI in include/dyphur_hip.h (libdyphur_hip.so).

This is the binding a Python caller of the reference's path would add on top of
the drop-in library; the GPU parity tests and bench.py call the frame step
through it.  ``load()`` raises when the library is missing: there is no CPU
fallback for the product path.

Device memory is allocated through the ABI (``dy_malloc``); ``DeviceBodies`` and
``DeviceShapes`` hold SoA arrays laid out exactly like the reference's
BodyView / ShapeView.
"""
from __future__ import annotations

import atexit
import ctypes as C
import itertools
import os
import weakref

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libdyphur_hip.so")
_lib = None

DY_OK = 0
DY_ERR_INVALID_ARG = -1
DY_ERR_UNSUPPORTED = -2
DY_ERR_NO_DEVICE = -3
DY_ERR_INTERNAL = -4

vp = C.c_void_p
u32 = C.c_uint32
f32 = C.c_float

BODY_FLOAT_FIELDS = ("pos_x", "pos_y", "pos_z", "rot_w", "rot_x", "rot_y", "rot_z",
                     "vel_x", "vel_y", "vel_z", "ang_x", "ang_y", "ang_z", "inv_mass",
                     "iI_xx", "iI_yy", "iI_zz", "iI_xy", "iI_xz", "iI_yz")
CONTACT_FIELDS = (("body_a", np.uint32), ("body_b", np.uint32), ("pos_x", np.float32),
                  ("pos_y", np.float32), ("pos_z", np.float32), ("norm_x", np.float32),
                  ("norm_y", np.float32), ("norm_z", np.float32), ("depth", np.float32))


class BodyView(C.Structure):
    _fields_ = [(f, vp) for f in BODY_FLOAT_FIELDS] + [("shape", vp), ("flags", vp), ("n", u32)]


class ShapeView(C.Structure):
    _fields_ = [("type", vp), ("half_x", vp), ("half_y", vp), ("half_z", vp), ("ext_id", vp), ("n", u32)]


class ContactView(C.Structure):
    _fields_ = [(f, vp) for f, _ in CONTACT_FIELDS] + [("n", vp), ("capacity", u32)]


JOINT_F32 = ("anchor_px", "anchor_py", "anchor_pz", "anchor_cx", "anchor_cy", "anchor_cz",
             "axis_px", "axis_py", "axis_pz", "limit_lo", "limit_hi", "target_pos", "target_vel",
             "stiffness", "damping", "compliance_pos", "compliance_ang")


class JointView(C.Structure):
    _fields_ = [("body_parent", vp), ("body_child", vp)] + [(f, vp) for f in JOINT_F32] + [
        ("type", vp), ("n", u32)]


class Aabb(C.Structure):
    _fields_ = [(n, f32) for n in ("min_x", "min_y", "min_z", "max_x", "max_y", "max_z")]


class BvhView(C.Structure):
    _fields_ = [("left", vp), ("right", vp), ("parent", vp), ("root", vp), ("sorted_idx", vp),
                ("aabb", vp), ("n", u32)]


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 256), ("is_gpu", C.c_int), ("arch_major", C.c_int),
                ("arch_minor", C.c_int), ("compute_units", C.c_size_t), ("global_mem_bytes", C.c_size_t)]


class SolveStats(C.Structure):
    _fields_ = [(n, u32) for n in ("n_constraints", "used_serial", "n_classes", "rounds", "overflow",
                                   "dyn_endpoints")]


class SolverTiming(C.Structure):
    _fields_ = [("solves", u32), ("assign_ms", C.c_double), ("sweep_ms", C.c_double),
                ("sweep_bytes", C.c_double)]


class BroadphaseTiming(C.Structure):
    _fields_ = [("builds", u32), ("sort_ms", C.c_double), ("karras_ms", C.c_double), ("refit_ms", C.c_double),
                ("traverse_ms", C.c_double), ("emit_ms", C.c_double)]


class BatchSceneStats(C.Structure):
    _fields_ = [(k, u32) for k in ("n_pairs", "n_contacts", "n_classes", "used_serial", "rounds",
                                   "overflow", "error", "substeps", "sum_pairs", "sum_contacts",
                                   "sum_dyn_endpoints", "solves", "grid_substeps")]


class BatchLayout(C.Structure):
    _fields_ = [("pairs", vp), ("contacts", ContactView), ("body_offset", C.POINTER(u32)),
                ("pair_offset", C.POINTER(u32)), ("contact_offset", C.POINTER(u32)), ("n_scenes", u32)]


class DyError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        msg = lib().dy_status_string(status).decode() if _lib else str(status)
        super().__init__(f"{what} failed: {msg} (status {status})")


_SIGS = {
    "dy_status_string": (C.c_char_p, [C.c_int]),
    "dy_version": (C.c_int, [C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "dy_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "dy_device_get_info": (C.c_int, [C.c_int, C.POINTER(DeviceInfo)]),
    "dy_set_device": (C.c_int, [C.c_int]),
    "dy_stream_create": (C.c_int, [C.POINTER(vp)]),
    "dy_stream_destroy": (C.c_int, [vp]),
    "dy_stream_synchronize": (C.c_int, [vp]),
    "dy_malloc": (C.c_int, [C.POINTER(vp), C.c_size_t]),
    "dy_free": (C.c_int, [vp]),
    "dy_memcpy_h2d_async": (C.c_int, [vp, vp, C.c_size_t, vp]),
    "dy_memcpy_d2h": (C.c_int, [vp, vp, C.c_size_t, vp]),
    "dy_memcpy_d2d_async": (C.c_int, [vp, vp, C.c_size_t, vp]),
    "dy_memcpy_async": (C.c_int, [vp, vp, C.c_size_t, vp]),
    "dy_memset_async": (C.c_int, [vp, C.c_int, C.c_size_t, vp]),
    "dy_event_create": (C.c_int, [C.POINTER(vp)]),
    "dy_event_destroy": (C.c_int, [vp]),
    "dy_event_record": (C.c_int, [vp, vp]),
    "dy_event_synchronize": (C.c_int, [vp]),
    "dy_event_elapsed_ms": (C.c_int, [vp, vp, C.POINTER(f32)]),
    "dy_sort_by_key_u64_u32": (C.c_int, [vp, vp, vp, C.c_size_t]),
    "dy_sort_by_key_u32_u32": (C.c_int, [vp, vp, vp, C.c_size_t]),
    "dy_sort_u32": (C.c_int, [vp, vp, C.c_size_t]),
    "dy_sort_u64": (C.c_int, [vp, vp, C.c_size_t]),
    "dy_reduce_sum_f32": (C.c_int, [vp, vp, C.c_size_t, C.POINTER(f32)]),
    "dy_scan_exclusive_u32": (C.c_int, [vp, vp, vp, C.c_size_t, vp]),
    "dy_body_dynamics_host": (C.c_int, [u32, vp, vp, vp, vp, vp, vp]),
    "dy_integrate": (C.c_int, [vp, C.POINTER(BodyView), f32, f32, f32, f32, C.c_int]),
    "dy_broadphase_create": (C.c_int, [vp, u32, u32, C.POINTER(vp)]),
    "dy_broadphase_destroy": (C.c_int, [vp]),
    "dy_broadphase_build_and_query": (C.c_int, [vp, vp, C.POINTER(BodyView), C.POINTER(ShapeView),
                                                 C.POINTER(Aabb)]),
    "dy_broadphase_pairs": (C.c_int, [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(u32)]),
    "dy_broadphase_sort_pairs": (C.c_int, [vp, vp, u32]),
    "dy_broadphase_download_count": (C.c_int, [vp, vp, C.POINTER(u32)]),
    "dy_broadphase_bvh_view": (C.c_int, [vp, u32, C.POINTER(BvhView)]),
    "dy_broadphase_set_timing": (C.c_int, [vp, C.c_int]),
    "dy_broadphase_get_timing": (C.c_int, [vp, C.POINTER(BroadphaseTiming)]),
    "dy_broadphase_morton_keys": (C.c_int, [vp, C.POINTER(vp)]),
    "dy_narrowphase_create": (C.c_int, [vp, u32, C.POINTER(vp)]),
    "dy_narrowphase_destroy": (C.c_int, [vp]),
    "dy_narrowphase_run": (C.c_int, [vp, vp, vp, u32, C.POINTER(BodyView), C.POINTER(ShapeView)]),
    "dy_narrowphase_run_device_count": (C.c_int, [vp, vp, vp, vp, u32, C.POINTER(BodyView),
                                                   C.POINTER(ShapeView)]),
    "dy_narrowphase_contacts": (C.c_int, [vp, C.POINTER(ContactView)]),
    "dy_narrowphase_download_count": (C.c_int, [vp, vp, C.POINTER(u32)]),
    "dy_narrowphase_last_count": (C.c_int, [vp, vp, C.POINTER(u32)]),
    "dy_narrowphase_set_features": (C.c_int, [vp, C.c_int]),
    "dy_narrowphase_features": (C.c_int, [vp, C.POINTER(vp)]),
    "dy_narrowphase_run_noalloc": (C.c_int, [vp, vp, vp, u32, C.POINTER(BodyView), C.POINTER(ShapeView),
                                              C.POINTER(ContactView), vp, vp]),
    "dy_xpbd_create": (C.c_int, [vp, C.c_int, u32, f32, C.POINTER(vp)]),
    "dy_xpbd_destroy": (C.c_int, [vp]),
    "dy_xpbd_set_force_serial": (C.c_int, [vp, C.c_int]),
    "dy_xpbd_solve": (C.c_int, [vp, vp, C.POINTER(ContactView), C.POINTER(JointView), C.POINTER(BodyView), f32]),
    "dy_xpbd_last_stats": (C.c_int, [vp, C.POINTER(SolveStats)]),
    "dy_xpbd_download_classes": (C.c_int, [vp, vp, vp, u32]),
    "dy_xpbd_set_timing": (C.c_int, [vp, C.c_int]),
    "dy_xpbd_get_timing": (C.c_int, [vp, vp, C.POINTER(SolverTiming)]),
    "dy_state_hash_fnv1a64": (C.c_int, [vp, C.POINTER(BodyView), C.POINTER(C.c_uint64)]),
    "dy_ray_query": (C.c_int, [vp, C.POINTER(BvhView), C.POINTER(BodyView), C.POINTER(ShapeView), vp, vp, u32]),
    "dy_graph_capture_begin": (C.c_int, [vp]),
    "dy_graph_capture_end": (C.c_int, [vp, C.POINTER(vp)]),
    "dy_graph_launch": (C.c_int, [vp, vp]),
    "dy_graph_destroy": (C.c_int, [vp]),
    "dy_batch_create": (C.c_int, [vp, u32, vp, vp, vp, C.c_int, f32, C.POINTER(vp)]),
    "dy_batch_destroy": (C.c_int, [vp]),
    "dy_batch_set_force_serial": (C.c_int, [vp, C.c_int]),
    "dy_batch_set_shape": (C.c_int, [vp, u32, u32]),
    "dy_batch_set_lds_limit": (C.c_int, [vp, u32]),
    "dy_batch_get_shape": (C.c_int, [vp, C.POINTER(u32), C.POINTER(u32)]),
    "dy_batch_step": (C.c_int, [vp, vp, C.POINTER(BodyView), C.POINTER(ShapeView), f32, f32, f32, f32,
                                C.c_int, C.c_int]),
    "dy_batch_autotune": (C.c_int, [vp, vp, C.POINTER(BodyView), C.POINTER(ShapeView), f32, f32, f32, f32,
                                    C.c_int, C.c_int, C.POINTER(u32), C.POINTER(u32), C.POINTER(f32)]),
    "dy_batch_stats": (C.c_int, [vp, vp, C.POINTER(BatchSceneStats)]),
    "dy_batch_layout_get": (C.c_int, [vp, C.POINTER(BatchLayout)]),
    "dy_batch_set_timing": (C.c_int, [vp, vp, C.c_int]),
    "dy_diff_create": (C.c_int, [vp, u32, u32, u32, f32, C.c_uint64, C.POINTER(vp)]),
    "dy_diff_tape_steps": (C.c_int, [vp, C.POINTER(u32), C.POINTER(C.c_uint64)]),
    "dy_diff_destroy": (C.c_int, [vp]),
    "dy_diff_forward": (C.c_int, [vp, vp, C.POINTER(BodyView), C.POINTER(ShapeView), f32, f32, f32, f32, C.c_int]),
    "dy_diff_backward": (C.c_int, [vp, vp, C.POINTER(BodyView), C.POINTER(ShapeView), f32, f32, f32, f32,
                                   C.POINTER(BodyView), C.POINTER(BodyView)]),
    "dy_batch_get_timing": (C.c_int, [vp, vp, vp, C.POINTER(C.c_double)]),
}

# Symbols declared in include/dyphur_hip.h (checked by the CPU test suite).
DECLARED = tuple(k for k in _SIGS if k != "dy_broadphase_morton_keys")


def load(path: str = LIB_PATH):
    """Load libdyphur_hip.so; raises RuntimeError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libdyphur_hip.so not built ({path}); run python -m dyphur_amd.build")
    L = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def lib():
    return load()


def check(status: int, what: str):
    if status != DY_OK:
        raise DyError(status, what)


# ---------------------------------------------------------------------------
# Lifetimes: every object that owns a device resource (stream, event, graph,
# stage object, batch, tape, allocation) is tracked, and release_all() closes
# the live ones newest first.  It runs at interpreter exit (atexit, while the
# library is still loaded), so the HIP runtime has nothing of ours left to
# tear down inside exit(), where a profiler that intercepts the runtime may
# already have finalised.
# ---------------------------------------------------------------------------
_live: "dict[int, weakref.ref]" = {}
_next_id = itertools.count()


def _track(obj) -> None:
    k = next(_next_id)
    _live[k] = weakref.ref(obj, lambda _r, k=k: _live.pop(k, None))


def release_all() -> None:
    """Closes every live stream, event, graph, stage object, batch, tape and
    device allocation made through this module, newest first."""
    for k in sorted(_live, reverse=True):
        r = _live.pop(k, None)
        o = r() if r is not None else None
        if o is not None:
            try:
                o.close()
            except Exception:
                pass


atexit.register(release_all)


class _Owned:
    """Base of the resource owners: close() releases once, __del__ closes."""

    def close(self):  # pragma: no cover - overridden
        pass

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_count() -> int:
    n = C.c_int(0)
    check(lib().dy_device_count(C.byref(n)), "dy_device_count")
    return n.value


# ---------------------------------------------------------------------------
# Streams, events, device memory
# ---------------------------------------------------------------------------
class Stream(_Owned):
    def __init__(self):
        h = vp()
        check(lib().dy_stream_create(C.byref(h)), "dy_stream_create")
        self.h = h
        _track(self)

    def wait(self):
        check(lib().dy_stream_synchronize(self.h), "dy_stream_synchronize")

    def close(self):
        if getattr(self, "h", None):
            lib().dy_stream_synchronize(self.h)
            lib().dy_stream_destroy(self.h)
            self.h = None


class Event(_Owned):
    def __init__(self):
        h = vp()
        check(lib().dy_event_create(C.byref(h)), "dy_event_create")
        self.h = h
        _track(self)

    def record(self, s: Stream):
        check(lib().dy_event_record(self.h, s.h), "dy_event_record")

    def sync(self):
        check(lib().dy_event_synchronize(self.h), "dy_event_synchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = f32(0)
        check(lib().dy_event_elapsed_ms(self.h, end.h, C.byref(ms)), "dy_event_elapsed_ms")
        return ms.value

    def close(self):
        if getattr(self, "h", None):
            lib().dy_event_destroy(self.h)
            self.h = None


class DeviceArray(_Owned):
    """A device allocation with a dtype and length."""

    def __init__(self, n: int, dtype):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = vp()
        check(lib().dy_malloc(C.byref(p), max(self.nbytes, 1)), "dy_malloc")
        self.ptr = p.value
        _track(self)

    @classmethod
    def from_numpy(cls, a: np.ndarray, s: Stream) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a, s)
        return d

    def upload(self, a: np.ndarray, s: Stream, count: int | None = None):
        a = np.ascontiguousarray(a, self.dtype)
        nb = (a.size if count is None else count) * self.dtype.itemsize
        check(lib().dy_memcpy_h2d_async(self.ptr, a.ctypes.data, nb, s.h), "dy_memcpy_h2d_async")
        s.wait()  # the numpy source may be freed after return

    def download(self, s: Stream, count: int | None = None) -> np.ndarray:
        n = self.n if count is None else count
        out = np.empty(n, self.dtype)
        if n:
            check(lib().dy_memcpy_d2h(out.ctypes.data, self.ptr, n * self.dtype.itemsize, s.h), "dy_memcpy_d2h")
        return out

    def free(self):
        if getattr(self, "ptr", None):
            lib().dy_free(vp(self.ptr))
            self.ptr = None

    def close(self):
        self.free()

    def copy_from_device(self, src_ptr: int, s: Stream, count: int | None = None):
        """Device-to-device copy of count elements (default all) from src_ptr."""
        nb = (self.n if count is None else count) * self.dtype.itemsize
        if nb:
            check(lib().dy_memcpy_d2d_async(vp(self.ptr), vp(src_ptr), nb, s.h), "dy_memcpy_d2d_async")


def download_ptr(ptr: int, n: int, dtype, s: Stream) -> np.ndarray:
    dtype = np.dtype(dtype)
    out = np.empty(n, dtype)
    if n:
        check(lib().dy_memcpy_d2h(out.ctypes.data, vp(ptr), n * dtype.itemsize, s.h), "dy_memcpy_d2h")
    return out


# ---------------------------------------------------------------------------
# Body / shape stores
# ---------------------------------------------------------------------------
def body_dynamics(mass: np.ndarray, inertia: np.ndarray, flags: np.ndarray):
    """BodyStore::add dynamics on the host: inv_mass, iI (n, 6), flags."""
    n = len(mass)
    mass = np.ascontiguousarray(mass, np.float32)
    inertia = np.ascontiguousarray(inertia, np.float32).reshape(n, 9)
    flags = np.ascontiguousarray(flags, np.uint32)
    inv = np.zeros(n, np.float32)
    iI = np.zeros((n, 6), np.float32)
    fo = np.zeros(n, np.uint32)
    check(lib().dy_body_dynamics_host(n, mass.ctypes.data, inertia.ctypes.data, flags.ctypes.data,
                                      inv.ctypes.data, iI.ctypes.data, fo.ctypes.data), "dy_body_dynamics_host")
    return inv, iI, fo


class DeviceBodies:
    """Device SoA body arrays (BodyView layout) for a Scene."""

    def __init__(self, scene, s: Stream, capacity: int | None = None):
        self.n = scene.n
        self.capacity = capacity or scene.n
        self.f = {k: DeviceArray(self.capacity, np.float32) for k in BODY_FLOAT_FIELDS}
        self.shape = DeviceArray(self.capacity, np.uint32)
        self.flags = DeviceArray(self.capacity, np.uint32)
        self.load(scene, s)

    def host_arrays(self, scene) -> dict:
        inv, iI, fo = body_dynamics(scene.mass, scene.inertia, scene.flags)
        h = {}
        for i, k in enumerate(("pos_x", "pos_y", "pos_z")):
            h[k] = scene.pos[:, i]
        for i, k in enumerate(("rot_w", "rot_x", "rot_y", "rot_z")):
            h[k] = scene.rot[:, i]
        for i, k in enumerate(("vel_x", "vel_y", "vel_z")):
            h[k] = scene.vel[:, i]
        for i, k in enumerate(("ang_x", "ang_y", "ang_z")):
            h[k] = scene.ang[:, i]
        h["inv_mass"] = inv
        for i, k in enumerate(("iI_xx", "iI_yy", "iI_zz", "iI_xy", "iI_xz", "iI_yz")):
            h[k] = iI[:, i]
        h["shape"] = scene.shape
        h["flags"] = fo
        return h

    def load(self, scene, s: Stream):
        h = self.host_arrays(scene)
        for k in BODY_FLOAT_FIELDS:
            self.f[k].upload(h[k], s)
        self.shape.upload(h["shape"], s)
        self.flags.upload(h["flags"], s)
        self.n = scene.n

    def view(self) -> BodyView:
        v = BodyView()
        for k in BODY_FLOAT_FIELDS:
            setattr(v, k, self.f[k].ptr)
        v.shape = self.shape.ptr
        v.flags = self.flags.ptr
        v.n = self.n
        return v

    def download(self, s: Stream) -> dict:
        d = {k: self.f[k].download(s, self.n) for k in BODY_FLOAT_FIELDS}
        d["shape"] = self.shape.download(s, self.n)
        d["flags"] = self.flags.download(s, self.n)
        return d

    def upload_state(self, state: dict, s: Stream):
        for k in BODY_FLOAT_FIELDS:
            if k in state:
                self.f[k].upload(state[k], s)

    def fnv(self, s: Stream) -> int:
        out = C.c_uint64(0)
        v = self.view()
        check(lib().dy_state_hash_fnv1a64(s.h, C.byref(v), C.byref(out)), "dy_state_hash_fnv1a64")
        return out.value


class DeviceJoints:
    """Device SoA joint arrays (the reference JointStore, JointView layout)
    from a dict of host arrays: body_parent, body_child (uint32), type
    (uint8: 0 fixed, 1 revolute, 2 prismatic, 3 ball) and the float32
    fields of JOINT_F32."""

    def __init__(self, joints: dict, s: Stream):
        self.n = len(joints["body_parent"])
        self.a = {"body_parent": DeviceArray.from_numpy(np.asarray(joints["body_parent"], np.uint32), s),
                  "body_child": DeviceArray.from_numpy(np.asarray(joints["body_child"], np.uint32), s),
                  "type": DeviceArray.from_numpy(np.asarray(joints["type"], np.uint8), s)}
        for f in JOINT_F32:
            a = np.asarray(joints[f], np.float32)
            if len(a) != self.n:
                raise ValueError(f"joint field {f} has {len(a)} entries, expected {self.n}")
            self.a[f] = DeviceArray.from_numpy(a, s)

    def set_targets(self, pos: np.ndarray, vel: np.ndarray, s: Stream):
        """JointStore::set_targets (core/include/core/joint_store.hpp:31)."""
        self.a["target_pos"].upload(np.asarray(pos, np.float32), s)
        self.a["target_vel"].upload(np.asarray(vel, np.float32), s)

    def view(self) -> JointView:
        v = JointView()
        for k, d in self.a.items():
            setattr(v, k, d.ptr)
        v.n = self.n
        return v


class DeviceShapes:
    def __init__(self, scene, s: Stream):
        self.n = scene.n_shapes
        self.type = DeviceArray.from_numpy(np.asarray(scene.shape_type, np.uint32), s)
        self.hx = DeviceArray.from_numpy(np.asarray(scene.shape_half[:, 0], np.float32), s)
        self.hy = DeviceArray.from_numpy(np.asarray(scene.shape_half[:, 1], np.float32), s)
        self.hz = DeviceArray.from_numpy(np.asarray(scene.shape_half[:, 2], np.float32), s)
        self.ext = DeviceArray.from_numpy(np.zeros(self.n, np.uint32), s)

    def view(self) -> ShapeView:
        return ShapeView(self.type.ptr, self.hx.ptr, self.hy.ptr, self.hz.ptr, self.ext.ptr, self.n)


def aabb(bounds) -> Aabb:
    return Aabb(*[float(x) for x in bounds])


# ---------------------------------------------------------------------------
# Stages
# ---------------------------------------------------------------------------
def integrate(s: Stream, bodies: DeviceBodies, gravity, dt: float, n_substeps: int):
    v = bodies.view()
    check(lib().dy_integrate(s.h, C.byref(v), gravity[0], gravity[1], gravity[2], dt, n_substeps), "dy_integrate")


class Broadphase(_Owned):
    def __init__(self, s: Stream, max_bodies: int, max_pairs: int):
        h = vp()
        check(lib().dy_broadphase_create(s.h, max_bodies, max_pairs, C.byref(h)), "dy_broadphase_create")
        self.h = h
        _track(self)
        self.max_bodies, self.max_pairs = max_bodies, max_pairs

    def build_and_query(self, s: Stream, bodies: DeviceBodies, shapes: DeviceShapes, bounds):
        bv, sv, bd = bodies.view(), shapes.view(), aabb(bounds)
        check(lib().dy_broadphase_build_and_query(self.h, s.h, C.byref(bv), C.byref(sv), C.byref(bd)),
              "dy_broadphase_build_and_query")

    def pairs_ptr(self):
        p, c, m = vp(), vp(), u32()
        check(lib().dy_broadphase_pairs(self.h, C.byref(p), C.byref(c), C.byref(m)), "dy_broadphase_pairs")
        return p.value, c.value

    def download_count(self, s: Stream) -> int:
        out = u32(0)
        check(lib().dy_broadphase_download_count(self.h, s.h, C.byref(out)), "dy_broadphase_download_count")
        return out.value

    def sort_pairs(self, s: Stream, n_pairs: int):
        check(lib().dy_broadphase_sort_pairs(self.h, s.h, n_pairs), "dy_broadphase_sort_pairs")

    def set_timing(self, on: bool):
        check(lib().dy_broadphase_set_timing(self.h, int(bool(on))), "dy_broadphase_set_timing")

    def timing(self) -> dict:
        t = BroadphaseTiming()
        check(lib().dy_broadphase_get_timing(self.h, C.byref(t)), "dy_broadphase_get_timing")
        return {k: getattr(t, k) for k, _ in BroadphaseTiming._fields_}

    def pairs(self, s: Stream, n: int) -> np.ndarray:
        n = min(n, self.max_pairs)
        p, _ = self.pairs_ptr()
        return download_ptr(p, 2 * n, np.uint32, s).reshape(n, 2)

    def bvh(self, s: Stream, n: int) -> dict:
        v = BvhView()
        check(lib().dy_broadphase_bvh_view(self.h, n, C.byref(v)), "dy_broadphase_bvh_view")
        k = vp()
        check(lib().dy_broadphase_morton_keys(self.h, C.byref(k)), "dy_broadphase_morton_keys")
        out = dict(morton=download_ptr(k.value, n, np.uint64, s),
                   sorted_idx=download_ptr(v.sorted_idx, n, np.uint32, s),
                   root=download_ptr(v.root, 1, np.int32, s))
        if n >= 2:
            out.update(left=download_ptr(v.left, n - 1, np.int32, s),
                       right=download_ptr(v.right, n - 1, np.int32, s),
                       parent=download_ptr(v.parent, 2 * n - 1, np.int32, s),
                       aabb=download_ptr(v.aabb, 8 * (2 * n - 1), np.uint16, s).reshape(2 * n - 1, 8))
        return out

    def close(self):
        if getattr(self, "h", None):
            lib().dy_broadphase_destroy(self.h)
            self.h = None


class Narrowphase(_Owned):
    def __init__(self, s: Stream, max_contacts: int):
        h = vp()
        check(lib().dy_narrowphase_create(s.h, max_contacts, C.byref(h)), "dy_narrowphase_create")
        self.h = h
        _track(self)
        self.max_contacts = max_contacts

    def run(self, s: Stream, pairs_ptr: int, n_pairs: int, bodies: DeviceBodies, shapes: DeviceShapes):
        bv, sv = bodies.view(), shapes.view()
        check(lib().dy_narrowphase_run(self.h, s.h, vp(pairs_ptr), n_pairs, C.byref(bv), C.byref(sv)),
              "dy_narrowphase_run")

    def run_device_count(self, s: Stream, pairs_ptr: int, count_ptr: int, max_pairs: int,
                         bodies: DeviceBodies, shapes: DeviceShapes):
        bv, sv = bodies.view(), shapes.view()
        check(lib().dy_narrowphase_run_device_count(self.h, s.h, vp(pairs_ptr), vp(count_ptr), max_pairs,
                                                    C.byref(bv), C.byref(sv)), "dy_narrowphase_run_device_count")

    def contacts_view(self) -> ContactView:
        v = ContactView()
        check(lib().dy_narrowphase_contacts(self.h, C.byref(v)), "dy_narrowphase_contacts")
        return v

    def download_count(self, s: Stream) -> int:
        out = u32(0)
        check(lib().dy_narrowphase_download_count(self.h, s.h, C.byref(out)), "dy_narrowphase_download_count")
        return out.value

    def last_count(self, s: Stream) -> int:
        out = u32(0)
        check(lib().dy_narrowphase_last_count(self.h, s.h, C.byref(out)), "dy_narrowphase_last_count")
        return out.value

    def set_features(self, on: bool):
        check(lib().dy_narrowphase_set_features(self.h, int(bool(on))), "dy_narrowphase_set_features")

    def features(self, s: Stream) -> np.ndarray:
        """Feature words of the stored contacts of the last run (set_features(True) first)."""
        n = self.download_count(s)
        p = vp()
        check(lib().dy_narrowphase_features(self.h, C.byref(p)), "dy_narrowphase_features")
        return download_ptr(p.value, n, np.uint32, s)

    def contacts(self, s: Stream) -> dict:
        n = self.download_count(s)
        v = self.contacts_view()
        out = {"n": n}
        for f, dt in CONTACT_FIELDS:
            out[f] = download_ptr(getattr(v, f), n, dt, s)
        return out

    def close(self):
        if getattr(self, "h", None):
            lib().dy_narrowphase_destroy(self.h)
            self.h = None


class XpbdSolver(_Owned):
    def __init__(self, s: Stream, n_iters: int = 10, max_contacts: int = 16384, friction: float = 0.5):
        h = vp()
        check(lib().dy_xpbd_create(s.h, n_iters, max_contacts, friction, C.byref(h)), "dy_xpbd_create")
        self.h = h
        _track(self)
        self.max_contacts = max_contacts

    def set_force_serial(self, v: bool):
        check(lib().dy_xpbd_set_force_serial(self.h, int(bool(v))), "dy_xpbd_set_force_serial")

    def solve(self, s: Stream, contacts: ContactView, bodies: DeviceBodies, dt: float,
              joints: "DeviceJoints | None" = None):
        bv = bodies.view()
        jv = C.byref(joints.view()) if joints is not None and joints.n else None
        check(lib().dy_xpbd_solve(self.h, s.h, C.byref(contacts), jv, C.byref(bv), dt), "dy_xpbd_solve")

    def stats(self) -> dict:
        st = SolveStats()
        check(lib().dy_xpbd_last_stats(self.h, C.byref(st)), "dy_xpbd_last_stats")
        return {k: getattr(st, k) for k, _ in SolveStats._fields_}

    def set_timing(self, on: bool):
        check(lib().dy_xpbd_set_timing(self.h, int(bool(on))), "dy_xpbd_set_timing")

    def timing(self, s: Stream) -> dict:
        t = SolverTiming()
        check(lib().dy_xpbd_get_timing(self.h, s.h, C.byref(t)), "dy_xpbd_get_timing")
        return {k: getattr(t, k) for k, _ in SolverTiming._fields_}

    def classes(self, s: Stream, n: int) -> np.ndarray:
        out = np.zeros(n, np.uint32)
        check(lib().dy_xpbd_download_classes(self.h, s.h, out.ctypes.data, n), "dy_xpbd_download_classes")
        return out

    def close(self):
        if getattr(self, "h", None):
            lib().dy_xpbd_destroy(self.h)
            self.h = None


class World:
    """The stress driver's per-frame sequence (examples/stress_test_blocks/main.cpp:264-273)
    over the C ABI: integrate -> build_and_query -> download_count -> sort_pairs ->
    narrowphase -> solve."""

    def __init__(self, scene, s: Stream | None = None):
        self.scene = scene
        self.s = s or Stream()
        self.bodies = DeviceBodies(scene, self.s)
        self.shapes = DeviceShapes(scene, self.s)
        self.bp = Broadphase(self.s, scene.n, scene.max_pairs)
        self.np = Narrowphase(self.s, scene.max_contacts)
        self.solver = XpbdSolver(self.s, scene.n_iters, scene.max_contacts, scene.mu)
        joints = getattr(scene, "joints", None)
        self.joints = DeviceJoints(joints, self.s) if joints is not None else None
        self.n_joints = self.joints.n if self.joints is not None else 0
        self.last = {}

    def step_async(self, force_serial: bool = False):
        """One frame with no host synchronisation: the pair count stays on the
        device (the reference's async-capable Narrowphase::run overload,
        core/include/core/narrowphase.hpp:39-46) and the solve is enqueued.
        Errors surface at the next blocking call (step(), fnv(), check())."""
        sc, s = self.scene, self.s
        integrate(s, self.bodies, sc.gravity, sc.dt, sc.n_substeps)
        self.bp.build_and_query(s, self.bodies, self.shapes, sc.bounds)
        pairs, count = self.bp.pairs_ptr()
        self.np.run_device_count(s, pairs, count, self.bp.max_pairs, self.bodies, self.shapes)
        self.solver.set_force_serial(force_serial)
        self.solver.solve(s, self.np.contacts_view(), self.bodies, sc.dt, self.joints)

    def check(self) -> dict:
        """Blocks until the enqueued frames are done and raises on any device
        error flag; returns the last frame's counts."""
        n_pairs = self.bp.download_count(self.s)
        n_contacts = self.np.download_count(self.s)
        st = self.solver.stats()
        return dict(n_pairs=n_pairs, n_contacts=n_contacts, n_colors=st["n_classes"],
                    used_serial=st["used_serial"])

    def step(self, force_serial: bool = False) -> dict:
        sc, s = self.scene, self.s
        integrate(s, self.bodies, sc.gravity, sc.dt, sc.n_substeps)
        self.bp.build_and_query(s, self.bodies, self.shapes, sc.bounds)
        n_pairs = self.bp.download_count(s)
        self.bp.sort_pairs(s, n_pairs)
        pairs, _ = self.bp.pairs_ptr()
        self.np.run(s, pairs, min(n_pairs, self.bp.max_pairs), self.bodies, self.shapes)
        self.solver.set_force_serial(force_serial)
        self.solver.solve(s, self.np.contacts_view(), self.bodies, sc.dt, self.joints)
        st = self.solver.stats()
        self.last = dict(n_pairs=n_pairs, n_contacts=st["n_constraints"] - self.n_joints, n_colors=st["n_classes"],
                         used_serial=st["used_serial"], dyn_endpoints=st["dyn_endpoints"], rounds=st["rounds"])
        return self.last

    def state(self) -> dict:
        return self.bodies.download(self.s)

    def fnv(self) -> int:
        return self.bodies.fnv(self.s)


def ray_query(s: Stream, bp: "Broadphase", n: int, bodies: "DeviceBodies", shapes: "DeviceShapes",
              rays: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Closest hits of rays (m, 7: origin xyz, direction xyz, t_max) over the
    BVH of bp's last build_and_query over n bodies: (t[m], body[m])."""
    r = np.ascontiguousarray(rays, np.float32)
    m = len(r)
    d_rays = DeviceArray.from_numpy(r.ravel(), s)
    d_hits = DeviceArray(2 * max(m, 1), np.uint32)
    bvh = BvhView()
    check(lib().dy_broadphase_bvh_view(bp.h, n, C.byref(bvh)), "dy_broadphase_bvh_view")
    bv, sv = bodies.view(), shapes.view()
    check(lib().dy_ray_query(s.h, C.byref(bvh), C.byref(bv), C.byref(sv), d_rays.ptr, d_hits.ptr, m),
          "dy_ray_query")
    out = d_hits.download(s, 2 * m).reshape(m, 2)
    return out[:, 0].view(np.float32).copy(), out[:, 1].copy()


class Graph(_Owned):
    """Calls issued on stream s inside ``with Graph.capture(s) as g:`` are
    recorded into a HIP graph; ``g.launch(s)`` replays them."""

    def __init__(self):
        self.h = vp()
        _track(self)

    @classmethod
    def capture(cls, s: Stream) -> "_Capture":
        return _Capture(cls(), s)

    def launch(self, s: Stream):
        check(lib().dy_graph_launch(self.h, s.h), "dy_graph_launch")

    def close(self):
        if getattr(self, "h", None):
            lib().dy_graph_destroy(self.h)
            self.h = vp()


class _Capture:
    def __init__(self, g: Graph, s: Stream):
        self.g, self.s = g, s

    def __enter__(self) -> Graph:
        check(lib().dy_graph_capture_begin(self.s.h), "dy_graph_capture_begin")
        return self.g

    def __exit__(self, et, ev, tb):
        st = lib().dy_graph_capture_end(self.s.h, C.byref(self.g.h))
        if et is None:
            check(st, "dy_graph_capture_end")
        return False


BATCH_STAT_FIELDS = ("n_pairs", "n_contacts", "n_classes", "used_serial", "rounds", "overflow", "error",
                     "substeps", "sum_pairs", "sum_contacts", "sum_dyn_endpoints", "solves", "grid_substeps")


class Batch(_Owned):
    """Many independent scenes stepped together (dy_batch_*): one launch runs
    every scene's frames, one workgroup per scene.  Body indices in pairs and
    contacts are scene-local, as each scene's own World would report them."""

    def __init__(self, scene_list, s: Stream | None = None, n_iters: int | None = None,
                 friction: float | None = None):
        from . import scenes as _scenes
        self.scenes = list(scene_list)
        if any(getattr(sc, "joints", None) is not None for sc in self.scenes):
            raise ValueError("the batched form solves contacts only; step scenes with joints with World")
        overridden = {"n_iters"} if n_iters is not None else set()
        if friction is not None:
            overridden.add("mu")
        self.merged = _scenes.concat(self.scenes, uniform=[k for k in _scenes.RUN_PARAMS if k not in overridden])
        self.s = s or Stream()
        m = self.merged.meta
        self.body_offset = np.ascontiguousarray(m["body_offset"], np.uint32)
        self.max_pairs = np.ascontiguousarray(m["max_pairs"], np.uint32)
        self.max_contacts = np.ascontiguousarray(m["max_contacts"], np.uint32)
        first = self.scenes[0]
        self.n_iters = first.n_iters if n_iters is None else n_iters
        self.mu = first.mu if friction is None else friction
        self.bodies = DeviceBodies(self.merged, self.s)
        self.shapes = DeviceShapes(self.merged, self.s)
        self.h = vp()
        check(lib().dy_batch_create(self.s.h, len(self.scenes), self.body_offset.ctypes.data,
                                    self.max_pairs.ctypes.data, self.max_contacts.ctypes.data,
                                    int(self.n_iters), float(self.mu), C.byref(self.h)), "dy_batch_create")
        _track(self)
        lay = BatchLayout()
        check(lib().dy_batch_layout_get(self.h, C.byref(lay)), "dy_batch_layout_get")
        self.layout = lay
        ns = len(self.scenes)
        self.pair_offset = np.ctypeslib.as_array(lay.pair_offset, (ns + 1,)).copy()
        self.contact_offset = np.ctypeslib.as_array(lay.contact_offset, (ns + 1,)).copy()

    @property
    def n_scenes(self) -> int:
        return len(self.scenes)

    @property
    def n_bodies(self) -> int:
        return int(self.body_offset[-1])

    def set_force_serial(self, v: bool):
        check(lib().dy_batch_set_force_serial(self.h, 1 if v else 0), "dy_batch_set_force_serial")

    def set_shape(self, threads: int, wave_scene_max: int = 128):
        check(lib().dy_batch_set_shape(self.h, threads, wave_scene_max), "dy_batch_set_shape")

    def set_lds_limit(self, nbytes: int):
        """LDS one workgroup may take (0: the device's maximum); beyond it a
        scene's state stays in global memory."""
        check(lib().dy_batch_set_lds_limit(self.h, int(nbytes)), "dy_batch_set_lds_limit")

    def shape(self) -> tuple[int, int]:
        t, w = u32(0), u32(0)
        check(lib().dy_batch_get_shape(self.h, C.byref(t), C.byref(w)), "dy_batch_get_shape")
        return t.value, w.value

    def step(self, n_frames: int = 1, dt: float | None = None, n_substeps: int | None = None, gravity=None):
        sc = self.scenes[0]
        g = sc.gravity if gravity is None else gravity
        bv, sv = self.bodies.view(), self.shapes.view()
        check(lib().dy_batch_step(self.h, self.s.h, C.byref(bv), C.byref(sv), g[0], g[1], g[2],
                                  sc.dt if dt is None else dt, sc.n_substeps if n_substeps is None else n_substeps,
                                  n_frames), "dy_batch_step")

    TUNE_SHAPES = ((64, 0), (128, 64), (128, 128), (256, 64), (256, 128), (256, 256))

    def autotune(self, trial_frames: int = 5) -> tuple[tuple[int, int], list[float]]:
        """Times the six launch shapes from the current state (which it
        restores) and keeps the fastest; returns it and the six times in ms."""
        sc = self.scenes[0]
        g = sc.gravity
        bv, sv = self.bodies.view(), self.shapes.view()
        t, w = u32(0), u32(0)
        ms = (f32 * 6)()
        check(lib().dy_batch_autotune(self.h, self.s.h, C.byref(bv), C.byref(sv), g[0], g[1], g[2], sc.dt,
                                      sc.n_substeps, trial_frames, C.byref(t), C.byref(w), ms),
              "dy_batch_autotune")
        return (t.value, w.value), list(ms)

    def stats(self, raise_on_error: bool = True) -> dict:
        """Per-scene statistics; raises DyError when a scene's device error
        flag is set (hull/mesh shape met, internal check) unless told not to."""
        arr = (BatchSceneStats * self.n_scenes)()
        st = lib().dy_batch_stats(self.h, self.s.h, arr)
        if raise_on_error:
            check(st, "dy_batch_stats")
        raw = np.ctypeslib.as_array(C.cast(arr, C.POINTER(u32)), (self.n_scenes, len(BATCH_STAT_FIELDS))).copy()
        return {k: raw[:, i] for i, k in enumerate(BATCH_STAT_FIELDS)}

    BATCH_STAGES = ("integrate_boxes", "pairs", "narrowphase", "assign", "sweep")

    def set_timing(self, on: bool):
        check(lib().dy_batch_set_timing(self.h, self.s.h, 1 if on else 0), "dy_batch_set_timing")

    def timing(self) -> dict:
        """Seconds per stage per scene accumulated while timing was on."""
        ticks = np.zeros((self.n_scenes, 5), np.uint64)
        hz = C.c_double(0)
        check(lib().dy_batch_get_timing(self.h, self.s.h, ticks.ctypes.data, C.byref(hz)), "dy_batch_get_timing")
        sec = ticks.astype(np.float64) / hz.value
        return {k: sec[:, i] for i, k in enumerate(self.BATCH_STAGES)}

    def pairs_all(self) -> np.ndarray:
        n = int(self.pair_offset[-1])
        return download_ptr(self.layout.pairs, 2 * n, np.uint32, self.s).reshape(n, 2)

    def contacts_all(self) -> dict:
        n = int(self.contact_offset[-1])
        cv = self.layout.contacts
        return {k: download_ptr(getattr(cv, k), n, dt, self.s) for k, dt in CONTACT_FIELDS}

    def state(self) -> dict:
        return self.bodies.download(self.s)

    def scene_slice(self, i: int) -> slice:
        return slice(int(self.body_offset[i]), int(self.body_offset[i + 1]))

    def close(self):
        if getattr(self, "h", None):
            lib().dy_batch_destroy(self.h)
            self.h = vp()


MUTABLE_FIELDS = ("pos_x", "pos_y", "pos_z", "rot_w", "rot_x", "rot_y", "rot_z",
                  "vel_x", "vel_y", "vel_z", "ang_x", "ang_y", "ang_z")


class Diff(_Owned):
    """Reverse-mode derivatives of a Batch's frames (dy_diff_*): forward()
    steps the batch keeping checkpoints; backward() turns the adjoint of the
    final state into the adjoint of the initial state and of inv_mass."""

    def __init__(self, batch: Batch, max_frames: int, checkpoint_every: int = 10, n_substeps: int = 1,
                 tape_contacts_per_body: float = 0.0, tape_bytes: int = 0):
        self.batch = batch
        self.h = vp()
        self.n_substeps = n_substeps
        check(lib().dy_diff_create(batch.h, max_frames, checkpoint_every, n_substeps, tape_contacts_per_body,
                                   tape_bytes, C.byref(self.h)), "dy_diff_create")
        _track(self)
        n = batch.n_bodies
        self.adj = {k: DeviceArray(n, np.float32) for k in MUTABLE_FIELDS + ("inv_mass",)}
        self.frames = 0

    def forward(self, n_frames: int):
        b = self.batch
        sc = b.scenes[0]
        g = sc.gravity
        bv, sv = b.bodies.view(), b.shapes.view()
        check(lib().dy_diff_forward(self.h, b.s.h, C.byref(bv), C.byref(sv), g[0], g[1], g[2], sc.dt, n_frames),
              "dy_diff_forward")
        self.frames = n_frames

    def tape_steps(self) -> int:
        return self.tape_layout()[0]

    def tape_layout(self) -> tuple[int, int]:
        """(substeps one tape holds, bytes of one recorded substep)."""
        steps, nbytes = u32(0), C.c_uint64(0)
        check(lib().dy_diff_tape_steps(self.h, C.byref(steps), C.byref(nbytes)), "dy_diff_tape_steps")
        return steps.value, nbytes.value

    def _adj_view(self) -> BodyView:
        v = BodyView()
        for k in BODY_FLOAT_FIELDS:
            setattr(v, k, self.adj[k].ptr if k in self.adj else None)
        v.n = self.batch.n_bodies
        return v

    def backward(self, adj_final: dict, download: bool = True, device_ptrs: bool = False) -> dict | None:
        """adj_final: host arrays (n,) for any of the 13 mutable fields (others
        zero), or with device_ptrs device addresses of float32 arrays (n,).
        Returns the adjoint of the initial state's 13 fields and of inv_mass
        (host arrays) unless download is False; self.adj holds them on the
        device either way."""
        b = self.batch
        sc = b.scenes[0]
        g = sc.gravity
        n = b.n_bodies
        for k in MUTABLE_FIELDS:
            if device_ptrs and k in adj_final:
                self.adj[k].copy_from_device(adj_final[k], b.s)
            else:
                self.adj[k].upload(np.ascontiguousarray(adj_final.get(k, np.zeros(n)), np.float32), b.s)
        v = self._adj_view()
        bv, sv = b.bodies.view(), b.shapes.view()
        check(lib().dy_diff_backward(self.h, b.s.h, C.byref(bv), C.byref(sv), g[0], g[1], g[2], sc.dt,
                                     C.byref(v), C.byref(v)), "dy_diff_backward")
        if not download:
            return None
        return {k: self.adj[k].download(b.s) for k in self.adj}

    def close(self):
        if getattr(self, "h", None):
            lib().dy_diff_destroy(self.h)
            self.h = vp()
