import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
os.environ.setdefault("MKL_THREADING_LAYER", "GNU")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long GPU case at full workload size")


def _gpu_available():
    try:
        from hmat_lib_amd import device
        return device.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Initialise the device; fail loudly (not skip) when a gpu test runs
    without one, so a missing native path never passes silently."""
    from hmat_lib_amd import device
    if device.device_count() <= 0:
        pytest.fail("gpu test requested but no HIP device is visible")
    device.init(0)
    return device
