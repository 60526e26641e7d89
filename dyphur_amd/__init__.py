"""dyphur_amd — the dyphur frame step (integrate -> LBVH broadphase -> pair
sort -> SAT narrowphase -> XPBD solve) as hand-written gfx950 HIP kernels
behind a C ABI (include/dyphur_hip.h, dyphur_amd/lib/libdyphur_hip.so).

Python entry points:
  dyphur_amd.capi    ctypes binding of the C ABI (streams, stores, stages, World)
  dyphur_amd.scenes  synthetic scenes (stress stack, random mixed, batched piles)
  dyphur_amd.build   in-tree build of the native libraries
"""
__all__ = ["capi", "scenes", "build"]
__version__ = "0.1.0"
