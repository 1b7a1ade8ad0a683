import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built libgcam_denoise.so")


@pytest.fixture(scope="session")
def oracle_threads():
    import oracle
    n = max(1, min(16, os.cpu_count() or 1))
    oracle.set_threads(n)
    return n
