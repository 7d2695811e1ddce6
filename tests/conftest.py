import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    oracle_lib.build()
    return oracle_lib


@pytest.fixture(scope="session")
def vdi():
    """The product package with its HIP library loaded (GPU tests only)."""
    import vdropimpact_amd
    if not os.path.exists(vdropimpact_amd.LIB_PATH):
        from vdropimpact_amd import build
        build.build()
    vdropimpact_amd.load_library()
    return vdropimpact_amd
