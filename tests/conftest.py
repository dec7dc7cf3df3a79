import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libhyburn_hip.so on a GPU)")


@pytest.fixture(scope="session")
def orc():
    from oracle import hb_oracle
    hb_oracle.lib()
    return hb_oracle


@pytest.fixture(scope="session")
def gpu():
    """The HIP library with at least one device; GPU tests fail loudly without it."""
    from hyburn_amd import _lib
    L = _lib.lib()
    n = L.hb_device_count()
    assert n > 0, "no HIP device visible"
    return L
