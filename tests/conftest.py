import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built libsoliton.so")


@pytest.fixture(scope="session")
def gpu_available():
    """True when HIP sees a device.  A missing library is an error, not a skip."""
    from soliton_solver_amd import _lib
    return _lib.device_count() > 0


@pytest.fixture(autouse=True)
def _require_gpu(request, gpu_available):
    if request.node.get_closest_marker("gpu") and not gpu_available:
        pytest.skip("no GPU visible to HIP")
