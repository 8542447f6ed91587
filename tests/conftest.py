import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on the device)")


@pytest.fixture(scope="session")
def oracle_lib():
    """The CPU oracle, built in-tree when this container has a compiler and it is missing."""
    from oracle import oracle as orc
    if not orc.LIB_PATH.exists():
        from oracle import build as ob
        ob.build()
    orc.load()
    return orc


@pytest.fixture(scope="session")
def hip():
    """The HIP backend on device 0; every test using it is marked gpu."""
    from aleph_amd import _lib
    if _lib.device_count() == 0:
        pytest.skip("no HIP device")
    from aleph_amd.backend import HipSvBackend
    return HipSvBackend(seed=1234)


@pytest.fixture(scope="session")
def hip32():
    from aleph_amd import _lib
    if _lib.device_count() == 0:
        pytest.skip("no HIP device")
    from aleph_amd.backend import HipSvBackend
    return HipSvBackend(seed=4321, precision=32)
