import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (MI355X / gfx950) and libqsx.so")


@pytest.fixture(scope="session")
def gpu():
    import qspace_amd._capi as capi
    import ctypes
    err = ctypes.create_string_buffer(256)
    rc = capi.lib().qsx_device_init(0, err, 256)
    if rc:
        pytest.fail(f"GPU test on a box without a usable device: {err.value.decode()}")
    return True
