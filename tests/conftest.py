import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
TESTS = Path(__file__).resolve().parent
if str(TESTS) not in sys.path:
    sys.path.insert(0, str(TESTS))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and libflowx_hip.so")


@pytest.fixture(scope="session")
def fx():
    import __graft_entry__

    return __graft_entry__.load_package()


@pytest.fixture(scope="session")
def oracle():
    import oracle as orc

    orc.lib()
    return orc
