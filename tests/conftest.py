import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libatsat_gpu.so)")
    config.addinivalue_line("markers", "slow: native-grid multi-step cases")


@pytest.fixture(scope="session")
def defaults():
    import atsat_amd
    return atsat_amd.defaults("resident")
