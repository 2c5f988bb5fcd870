import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU and libc2gpu.so")
    config.addinivalue_line("markers", "slow: large parity case (tens of seconds of oracle time)")


@pytest.fixture(scope="session")
def ctx():
    from critic2_amd._lib import Context
    c = Context(0)
    yield c
    c.close()
