import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on cuda:0)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def orc():
    from oracle import build, orc as _orc
    if not os.path.exists(os.path.join(ROOT, "oracle", "liborc.so")):
        build.build_oracle()
    return _orc


@pytest.fixture(scope="session")
def tri():
    """A device context; GPU tests fail loudly when the HIP library is missing."""
    import gstar4d_amd as g
    t = g.Triangulator(grid_size=64)
    yield t
    t.close()
