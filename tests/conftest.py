import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libmyrt_amd.so on the GPU)")


@pytest.fixture(scope="session")
def ensure_built():
    """The libraries are built by __graft_entry__.build(); build here only if absent."""
    import myrt_amd
    import oracle
    if not os.path.exists(myrt_amd.LIB_PATH):
        from myrt_amd import build
        build.build()
    if not os.path.exists(oracle.LIB_PATH):
        from oracle import build as ob
        ob.build()
    return True
