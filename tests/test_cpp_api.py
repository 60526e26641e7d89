"""The C++ layer (dyphur_amd/cpp/include: the reference's compute/ and core/
class surface over the C ABI), exercised by tests/cpp, which follow the
reference's Catch2 suites case by case."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "dyphur_cpp_tests")

EXPECTED = [
    "BodyStore: add and upload round-trip",
    "integrate: free fall matches the closed form",
    "Broadphase: four mutually overlapping spheres give all six pairs",
    "Narrowphase: box-box face contact",
    "XpbdSolver: coloured equals serial on independent pairs",
    "Determinism: golden free-fall hash, run to run",
    "sort_by_key: reverse order, u64 keys with u32 values",
    "reduce_sum: ones, and the same bits every run",
]


def _binary():
    if not os.path.exists(BIN):
        from dyphur_amd import build
        build.build_cpp_tests()
    return BIN


def test_cpp_tests_build_and_list():
    out = subprocess.run([_binary(), "--list"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    names = out.stdout.splitlines()
    assert len(names) >= 31
    for n in EXPECTED:
        assert n in names


def test_cpp_host_only_case():
    out = subprocess.run([_binary(), "--only", "AABB"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "1 of 1 test cases passed" in out.stdout, out.stdout


@pytest.mark.gpu
def test_cpp_api_suite_on_gpu(gpu):
    print("[cpp] running tests/cpp suite", flush=True)
    out = subprocess.run([_binary()], capture_output=True, text=True, timeout=110)
    print(out.stdout[-4000:], flush=True)
    assert out.returncode == 0, out.stdout[-4000:] + out.stderr[-2000:]
    assert "FAIL" not in out.stdout
