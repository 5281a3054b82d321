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


@pytest.fixture(scope="session")
def heis4000():
    """the headline operator (SU(2) Heisenberg two-site H*psi, D* = 4000),
    a seeded vector x and the oracle's H*psi of x (both terms), computed once
    per session: (HL, psi, HR, x, ref)"""
    import numpy as np
    from qspace_amd import gen
    from test_gpu_parity import _hpsi_oracle
    HL, psi, HR = gen.heis_hpsi(seed=1003, Dstar=4000)
    x = np.random.RandomState(4).standard_normal(psi.numel())
    ref = _hpsi_oracle(HL, psi.with_packed(x), HR, 1).packed()
    return HL, psi, HR, x, ref
