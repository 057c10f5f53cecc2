import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the product solve path on it")


@pytest.fixture(scope="session")
def gpu():
    """Fails loudly when the native library or the device is missing.  Device
    memory comes from the library's own HIP runtime (no torch in the process)."""
    import pargemslr_amd as pg
    pg.lib()
    assert pg.device_count() > 0, "GPU test without a visible HIP device"
    return 0
