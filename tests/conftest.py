import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def gpu_available():
    from cms2d_amd import _lib
    n = _lib.device_count()
    if n < 1:
        pytest.fail("no HIP device visible to a test marked gpu")
    return n
