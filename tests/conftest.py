import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs the product kernels")


@pytest.fixture(scope="session")
def oracle():
    from oracle import ref_cpu
    ref_cpu.build()
    ref_cpu.lib()
    return ref_cpu


@pytest.fixture(scope="session")
def capi():
    from dyphur_amd import capi as c
    c.load()
    return c


@pytest.fixture(scope="session")
def gpu(capi):
    if capi.device_count() < 1:
        pytest.fail("GPU test without a HIP device")
    return capi
