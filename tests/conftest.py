import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def orc():
    """The CPU oracle (built from oracle/ if its library is missing)."""
    import oracle
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    return oracle


@pytest.fixture(scope="session")
def dev():
    """The GPU device through the C-ABI; fails loudly without the library."""
    import barney_amd as bx
    d = bx.Device(0)
    yield d
    d.close()
