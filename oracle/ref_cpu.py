"""ctypes wrapper of the CPU oracle (libref_cpu.so).  TEST INFRASTRUCTURE:
imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, always as the checker / baseline, never as the measured product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libref_cpu.so")
_lib = None

u32p = C.POINTER(C.c_uint32)
f32p = C.POINTER(C.c_float)


class RcBodies(C.Structure):
    _fields_ = [(n, f32p) for n in (
        "pos_x", "pos_y", "pos_z", "rot_w", "rot_x", "rot_y", "rot_z",
        "vel_x", "vel_y", "vel_z", "ang_x", "ang_y", "ang_z", "inv_mass",
        "iI_xx", "iI_yy", "iI_zz", "iI_xy", "iI_xz", "iI_yz")] + [
        ("shape", u32p), ("flags", u32p), ("n", C.c_uint32)]


class RcShapes(C.Structure):
    _fields_ = [("type", u32p), ("half_x", f32p), ("half_y", f32p), ("half_z", f32p),
                ("ext_id", u32p), ("n", C.c_uint32)]


class RcAabb(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("min_x", "min_y", "min_z", "max_x", "max_y", "max_z")]


class RcContacts(C.Structure):
    _fields_ = [("body_a", u32p), ("body_b", u32p)] + [(n, f32p) for n in (
        "pos_x", "pos_y", "pos_z", "norm_x", "norm_y", "norm_z", "depth")] + [
        ("n", C.c_uint32), ("last_count", C.c_uint32)]


class RcSolveStats(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("n_constraints", "used_serial", "n_colors",
                                          "rounds", "overflow")]


JOINT_F32 = ("anchor_px", "anchor_py", "anchor_pz", "anchor_cx", "anchor_cy", "anchor_cz",
             "axis_px", "axis_py", "axis_pz", "limit_lo", "limit_hi", "target_pos", "target_vel",
             "stiffness", "damping", "compliance_pos", "compliance_ang")


class RcJoints(C.Structure):
    _fields_ = [("body_parent", u32p), ("body_child", u32p)] + [(n, f32p) for n in JOINT_F32] + [
        ("type", C.POINTER(C.c_uint8)), ("n", C.c_uint32)]


class RcFrameStats(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("n_pairs", "n_contacts", "n_colors", "used_serial")]


BODY_FIELDS = ("pos_x", "pos_y", "pos_z", "rot_w", "rot_x", "rot_y", "rot_z",
               "vel_x", "vel_y", "vel_z", "ang_x", "ang_y", "ang_z", "inv_mass",
               "iI_xx", "iI_yy", "iI_zz", "iI_xy", "iI_xz", "iI_yz")


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "libref_cpu.so"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        L = C.CDLL(_LIB_PATH)
        vp = C.c_void_p
        L.rc_half_from_float.restype = C.c_uint16
        L.rc_half_from_float.argtypes = [C.c_float]
        L.rc_half_to_float.restype = C.c_float
        L.rc_half_to_float.argtypes = [C.c_uint16]
        L.rc_half_from_float_array.argtypes = [vp, vp, C.c_size_t]
        L.rc_hash_u32.restype = C.c_uint32
        L.rc_hash_u32.argtypes = [C.c_uint32]
        L.rc_morton3.restype = C.c_uint32
        L.rc_morton3.argtypes = [C.c_uint32] * 3
        L.rc_fnv1a64.restype = C.c_uint64
        L.rc_fnv1a64.argtypes = [C.c_uint64, vp, C.c_size_t]
        L.rc_fnv1a64_state.restype = C.c_uint64
        L.rc_fnv1a64_state.argtypes = [C.POINTER(RcBodies)]
        L.rc_body_setup.argtypes = [C.c_uint32, vp, vp, vp, C.POINTER(RcBodies)]
        L.rc_integrate.argtypes = [C.POINTER(RcBodies), C.c_float, C.c_float, C.c_float,
                                   C.c_float, C.c_int]
        L.rc_bitonic_sort_u64_u32.argtypes = [vp, vp, C.c_size_t]
        L.rc_bitonic_sort_u32_u32.argtypes = [vp, vp, C.c_size_t]
        L.rc_bitonic_sort_u32.argtypes = [vp, C.c_size_t]
        L.rc_world_create.restype = vp
        L.rc_world_create.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_float]
        L.rc_world_destroy.argtypes = [vp]
        L.rc_bp_build_and_query.restype = C.c_uint32
        L.rc_bp_build_and_query.argtypes = [vp, C.POINTER(RcBodies), C.POINTER(RcShapes),
                                            C.POINTER(RcAabb)]
        L.rc_bp_copy_out.argtypes = [vp, C.c_uint32] + [vp] * 8
        L.rc_bp_sort_pairs.argtypes = [vp, C.c_uint32]
        L.rc_bp_copy_pairs.argtypes = [vp, vp, C.c_uint32]
        L.rc_bp_pairs.restype = vp
        L.rc_bp_pairs.argtypes = [vp]
        L.rc_np_run.restype = C.c_uint32
        L.rc_np_run.argtypes = [vp, vp, C.c_uint32, C.POINTER(RcBodies), C.POINTER(RcShapes)]
        L.rc_np_contacts.argtypes = [vp, C.POINTER(RcContacts)]
        L.rc_np_set_contacts.argtypes = [vp, C.c_uint32] + [vp] * 9
        L.rc_solve.argtypes = [vp, C.POINTER(RcBodies), C.c_float, C.c_int, C.POINTER(RcSolveStats)]
        L.rc_solve_joints.argtypes = [vp, C.POINTER(RcBodies), C.POINTER(RcJoints), C.c_float, C.c_int,
                                      C.POINTER(RcSolveStats)]
        L.rc_set_libc_atan2.argtypes = [C.c_int]
        L.rc_joint_atan2.restype = C.c_float
        L.rc_joint_atan2.argtypes = [C.c_float, C.c_float]
        L.rc_solve_colors.argtypes = [vp, vp, C.c_uint32]
        L.rc_frame.argtypes = [vp, C.POINTER(RcBodies), C.POINTER(RcShapes), C.POINTER(RcAabb),
                               C.c_float, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int,
                               C.POINTER(RcFrameStats)]
        L.rc_frame_joints.argtypes = [vp, C.POINTER(RcBodies), C.POINTER(RcShapes), C.POINTER(RcAabb),
                                      C.POINTER(RcJoints), C.c_float, C.c_float, C.c_float, C.c_float, C.c_int,
                                      C.c_int, C.POINTER(RcFrameStats)]
        L.rc_pair_contact.restype = C.c_int
        L.rc_pair_contact.argtypes = [C.c_uint32, C.c_uint32] + [vp] * 7
        L.rc_pair_contact_feat.restype = C.c_int
        L.rc_pair_contact_feat.argtypes = [C.c_uint32, C.c_uint32] + [vp] * 8
        L.rc_np_features.argtypes = [vp, vp, C.c_uint32]
        L.rc_project_contact.argtypes = [vp, vp, vp, C.c_float, vp, C.c_float]
        L.rc_leaf_extent.argtypes = [C.c_uint32, vp, vp, vp]
        L.rc_karras_node.argtypes = [vp, C.c_int, C.c_int, vp]
        L.rc_ray_query.argtypes = [vp, C.c_uint32, C.POINTER(RcBodies), C.POINTER(RcShapes), vp, vp, C.c_uint32]
        L.rc_max_threads.restype = C.c_int
        L.rc_set_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def _ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(C.POINTER(ct))


class Bodies:
    """Host SoA body state owned by numpy arrays (the oracle's BodyStore)."""

    def __init__(self, scene=None, n: int = 0):
        if scene is not None:
            n = scene.n
        self.n = n
        self.arr = {f: np.zeros(n, np.float32) for f in BODY_FIELDS}
        self.shape = np.zeros(n, np.uint32)
        self.flags = np.zeros(n, np.uint32)
        if scene is not None:
            self.load(scene)

    def load(self, sc):
        a = self.arr
        for i, f in enumerate(("pos_x", "pos_y", "pos_z")):
            a[f][:] = sc.pos[:, i]
        for i, f in enumerate(("rot_w", "rot_x", "rot_y", "rot_z")):
            a[f][:] = sc.rot[:, i]
        for i, f in enumerate(("vel_x", "vel_y", "vel_z")):
            a[f][:] = sc.vel[:, i]
        for i, f in enumerate(("ang_x", "ang_y", "ang_z")):
            a[f][:] = sc.ang[:, i]
        self.shape[:] = sc.shape
        mass = np.ascontiguousarray(sc.mass, np.float32)
        inertia = np.ascontiguousarray(sc.inertia, np.float32)
        flags_in = np.ascontiguousarray(sc.flags, np.uint32)
        lib().rc_body_setup(self.n, mass.ctypes.data, inertia.ctypes.data, flags_in.ctypes.data,
                            C.byref(self.view()))

    def view(self) -> RcBodies:
        v = RcBodies()
        for f in BODY_FIELDS:
            setattr(v, f, _ptr(self.arr[f], C.c_float))
        v.shape = _ptr(self.shape, C.c_uint32)
        v.flags = _ptr(self.flags, C.c_uint32)
        v.n = self.n
        self._keep = v
        return v

    def copy(self) -> "Bodies":
        o = Bodies(n=self.n)
        for f in BODY_FIELDS:
            o.arr[f][:] = self.arr[f]
        o.shape[:] = self.shape
        o.flags[:] = self.flags
        return o

    def state(self) -> dict:
        d = {f: self.arr[f].copy() for f in BODY_FIELDS}
        d["shape"] = self.shape.copy()
        d["flags"] = self.flags.copy()
        return d

    def fnv(self) -> int:
        return int(lib().rc_fnv1a64_state(C.byref(self.view())))


class Shapes:
    def __init__(self, scene):
        self.type = np.ascontiguousarray(scene.shape_type, np.uint32)
        self.hx = np.ascontiguousarray(scene.shape_half[:, 0], np.float32)
        self.hy = np.ascontiguousarray(scene.shape_half[:, 1], np.float32)
        self.hz = np.ascontiguousarray(scene.shape_half[:, 2], np.float32)
        self.ext = np.zeros(len(self.type), np.uint32)

    def view(self) -> RcShapes:
        v = RcShapes(_ptr(self.type, C.c_uint32), _ptr(self.hx, C.c_float),
                     _ptr(self.hy, C.c_float), _ptr(self.hz, C.c_float),
                     _ptr(self.ext, C.c_uint32), len(self.type))
        self._keep = v
        return v


class Joints:
    """Host joint SoA (the oracle's JointStore): a dict of arrays with the
    reference JointView's fields (body_parent, body_child: uint32; type:
    uint8; the rest float32)."""

    def __init__(self, arrays: dict):
        n = len(arrays["body_parent"])
        self.n = n
        self.arr = {"body_parent": np.ascontiguousarray(arrays["body_parent"], np.uint32),
                    "body_child": np.ascontiguousarray(arrays["body_child"], np.uint32),
                    "type": np.ascontiguousarray(arrays["type"], np.uint8)}
        for f in JOINT_F32:
            self.arr[f] = np.ascontiguousarray(arrays[f], np.float32)
            assert len(self.arr[f]) == n, f

    def view(self) -> RcJoints:
        v = RcJoints()
        v.body_parent = _ptr(self.arr["body_parent"], C.c_uint32)
        v.body_child = _ptr(self.arr["body_child"], C.c_uint32)
        for f in JOINT_F32:
            setattr(v, f, _ptr(self.arr[f], C.c_float))
        v.type = _ptr(self.arr["type"], C.c_uint8)
        v.n = self.n
        self._keep = v
        return v


def aabb(bounds) -> RcAabb:
    return RcAabb(*[float(x) for x in bounds])


class World:
    """Oracle Broadphase + Narrowphase + XpbdSolver buffers for one scene."""

    def __init__(self, max_bodies: int, max_pairs: int, max_contacts: int,
                 n_iters: int = 10, mu: float = 0.5):
        self.L = lib()
        self.max_bodies, self.max_pairs, self.max_contacts = max_bodies, max_pairs, max_contacts
        self.h = self.L.rc_world_create(max_bodies, max_pairs, max_contacts, n_iters, mu)

    @classmethod
    def for_scene(cls, sc, max_contacts=None):
        return cls(sc.n, sc.max_pairs, max_contacts or sc.max_contacts, sc.n_iters, sc.mu)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.rc_world_destroy(self.h)
            self.h = None

    def build_and_query(self, bodies: Bodies, shapes: Shapes, bounds) -> int:
        bd = aabb(bounds)
        return int(self.L.rc_bp_build_and_query(self.h, C.byref(bodies.view()),
                                                C.byref(shapes.view()), C.byref(bd)))

    def bvh(self, n: int) -> dict:
        m = max(n, 1)
        out = dict(morton=np.zeros(m, np.uint64), sorted_idx=np.zeros(m, np.uint32),
                   left=np.zeros(max(n - 1, 1), np.int32), right=np.zeros(max(n - 1, 1), np.int32),
                   parent=np.zeros(max(2 * n - 1, 1), np.int32), root=np.zeros(1, np.int32),
                   aabb=np.zeros((max(2 * n - 1, 1), 8), np.uint16),
                   refit_rounds=np.zeros(1, np.uint32))
        self.L.rc_bp_copy_out(self.h, n, *[out[k].ctypes.data for k in (
            "morton", "sorted_idx", "left", "right", "parent", "root", "aabb", "refit_rounds")])
        return out

    def sort_pairs(self, n_pairs: int):
        self.L.rc_bp_sort_pairs(self.h, n_pairs)

    def pairs(self, n: int) -> np.ndarray:
        n = min(n, self.max_pairs)
        out = np.zeros((n, 2), np.uint32)
        if n:
            self.L.rc_bp_copy_pairs(self.h, out.ctypes.data, n)
        return out

    def narrowphase(self, pairs: np.ndarray, bodies: Bodies, shapes: Shapes) -> int:
        p = np.ascontiguousarray(pairs, np.uint32)
        return int(self.L.rc_np_run(self.h, p.ctypes.data, len(p), C.byref(bodies.view()),
                                    C.byref(shapes.view())))

    def contacts(self) -> dict:
        c = RcContacts()
        self.L.rc_np_contacts(self.h, C.byref(c))
        n = c.n
        out = {"n": n, "last_count": c.last_count}
        for f, dt in (("body_a", np.uint32), ("body_b", np.uint32), ("pos_x", np.float32),
                      ("pos_y", np.float32), ("pos_z", np.float32), ("norm_x", np.float32),
                      ("norm_y", np.float32), ("norm_z", np.float32), ("depth", np.float32)):
            ptr = getattr(c, f)
            out[f] = np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt).copy() if n else np.zeros(0, dt)
        return out

    def features(self) -> np.ndarray:
        """Feature words of the stored contacts (see ref_cpu.h)."""
        c = RcContacts()
        self.L.rc_np_contacts(self.h, C.byref(c))
        out = np.zeros(c.n, np.uint32)
        if c.n:
            self.L.rc_np_features(self.h, out.ctypes.data, c.n)
        return out

    def set_contacts(self, c: dict):
        n = len(c["body_a"])
        arrs = [np.ascontiguousarray(c[f], dt) for f, dt in (
            ("body_a", np.uint32), ("body_b", np.uint32), ("pos_x", np.float32),
            ("pos_y", np.float32), ("pos_z", np.float32), ("norm_x", np.float32),
            ("norm_y", np.float32), ("norm_z", np.float32), ("depth", np.float32))]
        self.L.rc_np_set_contacts(self.h, n, *[a.ctypes.data for a in arrs])

    def solve(self, bodies: Bodies, dt: float, force_serial: bool = False, joints: "Joints | None" = None) -> dict:
        st = RcSolveStats()
        if joints is None:
            self.L.rc_solve(self.h, C.byref(bodies.view()), dt, int(force_serial), C.byref(st))
        else:
            self.L.rc_solve_joints(self.h, C.byref(bodies.view()), C.byref(joints.view()), dt, int(force_serial),
                                   C.byref(st))
        return {k: getattr(st, k) for k, _ in RcSolveStats._fields_}

    def colors(self, n: int) -> np.ndarray:
        out = np.zeros(n, np.uint32)
        self.L.rc_solve_colors(self.h, out.ctypes.data, n)
        return out

    def ray_query(self, bodies: Bodies, shapes: Shapes, n: int, rays: np.ndarray):
        """Closest hits of rays (m, 7: origin, direction, t_max) against the
        BVH of the last build_and_query over n bodies: (t[m], body[m])."""
        r = np.ascontiguousarray(rays, np.float32)
        out = np.zeros((len(r), 2), np.uint32)
        self.L.rc_ray_query(self.h, n, C.byref(bodies.view()), C.byref(shapes.view()), r.ctypes.data,
                            out.ctypes.data, len(r))
        return out[:, 0].view(np.float32).copy(), out[:, 1].copy()

    def frame(self, bodies: Bodies, shapes: Shapes, sc, force_serial: bool = False,
              joints: "Joints | None" = None) -> dict:
        st = RcFrameStats()
        bd = aabb(sc.bounds)
        g = sc.gravity
        if joints is None and getattr(sc, "joints", None) is not None:
            joints = Joints(sc.joints)
        if joints is None:
            self.L.rc_frame(self.h, C.byref(bodies.view()), C.byref(shapes.view()), C.byref(bd),
                            g[0], g[1], g[2], sc.dt, sc.n_substeps, int(force_serial), C.byref(st))
        else:
            self.L.rc_frame_joints(self.h, C.byref(bodies.view()), C.byref(shapes.view()), C.byref(bd),
                                   C.byref(joints.view()), g[0], g[1], g[2], sc.dt, sc.n_substeps,
                                   int(force_serial), C.byref(st))
        return {k: getattr(st, k) for k, _ in RcFrameStats._fields_}


def integrate(bodies: Bodies, gravity, dt: float, n_substeps: int):
    lib().rc_integrate(C.byref(bodies.view()), gravity[0], gravity[1], gravity[2], dt, n_substeps)


def pair_contact(ta: int, tb: int, ha, hb, pa, pb, qa, qb):
    """One pair's contact: (hit, [px py pz nx ny nz depth], feature word)."""
    arrs = [np.ascontiguousarray(x, np.float32) for x in (ha, hb, pa, pb, qa, qb)]
    out = np.zeros(7, np.float32)
    feat = np.zeros(1, np.uint32)
    hit = lib().rc_pair_contact_feat(ta, tb, *[a.ctypes.data for a in arrs], out.ctypes.data, feat.ctypes.data)
    return bool(hit), out, int(feat[0])


def set_threads(n: int):
    lib().rc_set_threads(int(n))


def max_threads() -> int:
    return int(lib().rc_max_threads())
