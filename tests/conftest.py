import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP engine")


@pytest.fixture(scope="session")
def engine():
    """The HIP engine on device 0.  On a GPU run a missing or unloadable
    library is an error, never a skip."""
    from spinwalk_amd import engine as E
    eng = E.Engine(0)
    yield eng
    eng.close()
