"""Build the native libraries in-tree.

  dyphur_amd/lib/libdyphur_hip.so       gfx950 kernels + the C ABI (include/dyphur_hip.h)
  dyphur_amd/lib/libdyphur_hostmath.so  the same per-element arithmetic (dy_math.hpp)
                                        compiled for the host, for CPU parity checks
  dyphur_amd/dyphur_py*.so              the Python module over the C++ surface
  tests/cpp/_build/dyphur_cpp_tests     the C++ API tests (dyphur_amd/cpp/include)

Run ``python -m dyphur_amd.build`` (or __graft_entry__.build()).  Every call
recompiles.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
BUILD = os.path.join(PKG, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DYPHUR_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["runtime.hip", "primitives.hip", "broadphase.hip", "narrowphase.hip", "narrowphase_noalloc.hip",
               "solver.hip", "batch.hip", "ray_query.hip"]

# -ffp-contract=off: parity with the reference requires every float op to round
# on its own.  Packed-FP32 instructions are disabled: some of their forms have
# given wrong results on these GPUs.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++20", "-fPIC", "-ffp-contract=off",
    "-fno-gpu-rdc", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", f"-I{os.path.join(ROOT, 'include')}",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build_hip(verbose: bool = False) -> str:
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BUILD, exist_ok=True)

    def compile_one(src):
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=min(8, len(HIP_SOURCES))) as ex:
        objs = list(ex.map(compile_one, HIP_SOURCES))
    out = os.path.join(LIB, "libdyphur_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build_hostmath(verbose: bool = False) -> str:
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libdyphur_hostmath.so")
    tmp = out + ".tmp"
    _run([HIPCC, "-x", "hip", "--cuda-host-only", "-O2", "-std=c++20", "-fPIC", "-shared",
          "-ffp-contract=off", f"-I{os.path.join(ROOT, 'include')}",
          os.path.join(CSRC, "hostmath.cpp"), "-o", tmp])
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build_python_module(verbose: bool = False) -> str | None:
    src = os.path.join(CSRC, "bindings", "dyphur_py.cpp")
    if not os.path.exists(src):
        return None
    import pybind11
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = os.path.join(PKG, "dyphur_py" + suffix)
    tmp = out + ".tmp"
    inc = [f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}",
           f"-I{os.path.join(ROOT, 'include')}", f"-I{os.path.join(PKG, 'cpp', 'include')}"]
    _run(["g++", "-O2", "-std=c++20", "-fPIC", "-shared", "-fvisibility=hidden", *inc, src,
          f"-L{LIB}", "-ldyphur_hip", "-Wl,-rpath,$ORIGIN/lib", "-o", tmp])
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


CPP_TEST_SOURCES = ["main.cpp", "test_stores_integrator.cpp", "test_pipeline.cpp", "test_compute.cpp", "test_ray_query.cpp",
                    "test_joints.cpp"]


def build_cpp_tests(verbose: bool = False) -> str:
    """The C++ API tests (tests/cpp) against the header layer in
    dyphur_amd/cpp/include and libdyphur_hip.so; hipcc, because parallel_for
    callers compile device lambdas."""
    src_dir = os.path.join(ROOT, "tests", "cpp")
    out_dir = os.path.join(src_dir, "_build")
    os.makedirs(out_dir, exist_ok=True)
    flags = [f"--offload-arch={ARCH}", "-O2", "-std=c++20", "-fno-gpu-rdc", "-Xclang", "-target-feature",
             "-Xclang", "-packed-fp32-ops", "-Wall", f"-I{os.path.join(ROOT, 'include')}",
             f"-I{os.path.join(PKG, 'cpp', 'include')}"]

    def compile_one(src):
        obj = os.path.join(out_dir, src.replace(".cpp", ".o"))
        _run([HIPCC, *flags, "-x", "hip", "-c", os.path.join(src_dir, src), "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=len(CPP_TEST_SOURCES)) as ex:
        objs = list(ex.map(compile_one, CPP_TEST_SOURCES))
    out = os.path.join(out_dir, "dyphur_cpp_tests")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-o", tmp, *objs, f"-L{LIB}", "-ldyphur_hip",
          "-Wl,-rpath,$ORIGIN/../../../dyphur_amd/lib"])
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build_examples(verbose: bool = False) -> list:
    """Drivers over the C++ layer (dyphur_amd/cpp/examples) -> dyphur_amd/bin."""
    out_dir = os.path.join(PKG, "bin")
    os.makedirs(out_dir, exist_ok=True)
    src_dir = os.path.join(PKG, "cpp", "examples")
    outs = []
    for src in sorted(os.listdir(src_dir)):
        if not src.endswith(".cpp"):
            continue
        out = os.path.join(out_dir, src[:-4])
        _run(["g++", "-O2", "-std=c++20", "-Wall", f"-I{os.path.join(ROOT, 'include')}",
              f"-I{os.path.join(PKG, 'cpp', 'include')}", os.path.join(src_dir, src), "-o", out + ".tmp",
              f"-L{LIB}", "-ldyphur_hip", "-Wl,-rpath,$ORIGIN/../lib"])
        os.replace(out + ".tmp", out)
        outs.append(out)
        if verbose:
            print("built", out)
    return outs


def build_all(verbose: bool = False):
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    paths = [build_hip(verbose), build_hostmath(verbose), build_cpp_tests(verbose), *build_examples(verbose)]
    py = build_python_module(verbose)
    if py:
        paths.append(py)
    return paths


if __name__ == "__main__":
    for p in build_all(verbose=True):
        pass
    sys.exit(0)
