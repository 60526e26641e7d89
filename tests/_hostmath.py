"""ctypes access to libdyphur_hostmath.so: the kernels' per-element arithmetic
(dyphur_amd/csrc/dy_math.hpp) compiled for the host."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PATH = os.path.join(ROOT, "dyphur_amd", "lib", "libdyphur_hostmath.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(PATH):
            raise RuntimeError(f"{PATH} not built (python -m dyphur_amd.build)")
        L = C.CDLL(PATH)
        vp = C.c_void_p
        L.hm_f32_to_f16.restype = C.c_uint16
        L.hm_f32_to_f16.argtypes = [C.c_float]
        L.hm_f16_to_f32.restype = C.c_float
        L.hm_f16_to_f32.argtypes = [C.c_uint16]
        L.hm_f32_to_f16_array.argtypes = [vp, vp, C.c_size_t]
        L.hm_morton3.restype = C.c_uint32
        L.hm_morton3.argtypes = [C.c_uint32] * 3
        L.hm_quantize10.restype = C.c_uint32
        L.hm_quantize10.argtypes = [C.c_float] * 3
        L.hm_hash_u32.restype = C.c_uint32
        L.hm_hash_u32.argtypes = [C.c_uint32]
        L.hm_karras_node.argtypes = [vp, C.c_int, C.c_int, vp]
        L.hm_leaf_extent.argtypes = [C.c_uint32, vp, vp, vp]
        L.hm_pair_contact.restype = C.c_int
        L.hm_pair_contact.argtypes = [C.c_uint32, C.c_uint32] + [vp] * 7
        L.hm_pair_contact_feat.restype = C.c_int
        L.hm_pair_contact_feat.argtypes = [C.c_uint32, C.c_uint32] + [vp] * 8
        L.hm_pair_projection_mismatches.restype = C.c_int
        L.hm_pair_projection_mismatches.argtypes = [C.c_int, C.c_uint32]
        L.hm_project_contact.restype = C.c_int
        L.hm_project_contact.argtypes = [vp, vp, vp, C.c_float, vp, C.c_float]
        L.hm_integrate_body.argtypes = [vp, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int]
        L.hm_renormalise.argtypes = [vp]
        L.hm_atan2_cr.restype = C.c_float
        L.hm_atan2_cr.argtypes = [C.c_float, C.c_float]
        L.hm_atan2_cr_array.argtypes = [vp, vp, vp, C.c_size_t]
        L.hm_joint_pair.argtypes = [vp, vp, vp, C.c_uint32, C.c_float, C.c_int]
        _lib = L
    return _lib


def f32(*a):
    return np.ascontiguousarray(np.array(a, np.float32).ravel())
