import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (REPO, os.path.join(REPO, "tart-gnss_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (runs on the MI355X box)")


def gpu_available():
    try:
        import tga
        return tga.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skip-free guard: GPU tests must fail loudly, not silently fall back, when the device
    or the library is missing."""
    import tga
    n = tga.device_count()
    assert n > 0, "gpu-marked test run without a visible HIP device"
    return n
