import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def gpu():
    """The HIP library, initialised on device 0."""
    from cc_mpi_amd import _native
    lib = _native.gpu_lib()
    rc = lib.cc_hip_init(0)
    assert rc == 0, lib.cc_hip_last_error().decode()
    return lib
