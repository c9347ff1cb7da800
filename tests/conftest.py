import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("oracle", "tetsuhiro-wos_amd"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libwos_hip's C ABI)")


@pytest.fixture(scope="session")
def wos():
    """libwos_hip on a GPU; a missing library or device is an error, not a skip."""
    import wos_hip
    wos_hip.load()
    if wos_hip.device_count() <= 0:
        pytest.fail("no HIP device visible to libwos_hip")
    wos_hip.init()
    return wos_hip


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O
