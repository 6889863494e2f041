import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def bk():
    """The product library; on a GPU run it must load or the test fails."""
    import bricklib_amd
    bricklib_amd.load()
    return bricklib_amd


@pytest.fixture(scope="session")
def gpu(bk):
    n = bk.device_count()
    if n == 0:
        pytest.fail("no HIP device visible to a gpu-marked test")
    return bk


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle
