import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on the device)")


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib

    return oracle_lib


@pytest.fixture(scope="session")
def vx():
    """The product package. GPU tests fail (never skip) when no device is visible."""
    import vortexstark_amd

    assert vortexstark_amd.gpu_available(), "no HIP device visible to libvortex_hip"
    return vortexstark_amd
