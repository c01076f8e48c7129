import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running CPU case")


@pytest.fixture(scope="session")
def dev():
    import gpu_bitruss_amd as gb

    if gb.device_count() == 0:
        pytest.fail("no HIP device visible to a test marked gpu")
    d = gb.Device(0)
    yield d
    d.close()
