import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__ as ge  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def pkg():
    return ge.load_package()


@pytest.fixture(scope="session")
def oracle():
    o = ge.load_oracle()
    o.lib()
    return o


@pytest.fixture(scope="session")
def hipmod(pkg):
    from cpra_amd import hip

    hip.lib()
    return hip


@pytest.fixture(scope="session")
def gpu(hipmod):
    """GPU tests allocate through the library (hip.DeviceArray) and never
    import torch, so the process holds a single HIP runtime."""
    if hipmod.device_count() < 1:
        pytest.fail("GPU test collected on a machine without a GPU")
    return True
