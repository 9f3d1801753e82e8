import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libpfdtd.so")


def _have_gpu() -> bool:
    try:
        from parallelfdtd_amd import native
        return native.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skip-free guard: a GPU test on a machine without a GPU is an error, not a skip."""
    from parallelfdtd_amd import native
    n = native.device_count()
    assert n > 0, "no GPU visible"
    return 0
