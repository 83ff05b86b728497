import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libmdn_mech.so on the device)")
    config.addinivalue_line("markers", "slow: long-running CPU oracle case")


def _gpu_available():
    try:
        import torch  # noqa: F401  (device discovery only)
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """A ForceFieldManager factory on device 0; GPU tests fail (not skip) without a device."""
    from picgo_amd import mech
    mech.lib()
    return mech.ForceFieldManager
