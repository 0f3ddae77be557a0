import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
PKG = os.path.join(REPO, "helmholtz-3d-torrone_amd")
for p in (HERE, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


def gpu_available():
    try:
        import ibim
        return ibim.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    import ibim
    if ibim.device_count() <= 0:
        pytest.fail("no GPU visible to the HIP library")
    return ibim
