import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: CPU test that takes more than a few seconds")


def gpu_available():
    """True when a ROCm device node exists (the library itself decides whether it can run)."""
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def oracle():
    from oracle import atmo
    atmo.lib()
    return atmo
