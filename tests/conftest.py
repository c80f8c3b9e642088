import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def gpu():
    from rootpwa_amd import capi
    if capi.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    return capi
