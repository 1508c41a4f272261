import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) HIP device")


@pytest.fixture(scope="session")
def hip():
    """The HIP library, required to have a usable device (gpu tests only)."""
    from spatialrust_amd import native
    lib = native.load()
    if not lib.sr_device_available():
        pytest.fail("marked gpu but no usable gfx950 device: " + native.last_error())
    return lib


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle
