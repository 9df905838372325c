import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "murmuration-kernels_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def fh():
    import flock_hip

    flock_hip.lib()  # raises when libflock_hip.so is not built
    return flock_hip


@pytest.fixture(scope="session")
def fr():
    from oracle import flockref

    flockref.lib()
    return flockref
