import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (runs on the MI355X box)")


@pytest.fixture(scope="session")
def gpu():
    """The HIP library with a usable gfx950 device; fails (never skips) when absent,
    so a GPU run cannot pass on a fallback."""
    import tentaco_amd as tt

    tt.device_check()
    return tt
