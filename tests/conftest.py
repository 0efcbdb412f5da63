import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the device)")


@pytest.fixture(scope="session")
def gi():
    import easy_graphit_amd as gi
    if gi.device_count() < 1:
        pytest.skip("no GPU visible")
    return gi


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle
