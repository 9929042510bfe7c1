import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "afivo-streamer_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def A():
    import afgpu
    return afgpu


@pytest.fixture(scope="session")
def S():
    import setups
    return setups


@pytest.fixture(scope="session")
def CH():
    import chemistry
    return chemistry
