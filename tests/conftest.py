import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "lklight-gpu_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def gpu():
    """Skip unless the HIP library loads and a gfx950 device is present."""
    import lklight
    if not lklight.device_available():
        pytest.fail("gpu-marked test but no gfx950 device is available")
    return lklight
