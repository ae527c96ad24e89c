import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libdoon.so")
    config.addinivalue_line("markers", "slow: long-running")


def _have_gpu() -> bool:
    try:
        from doonengine_amd._hip import device_count
        return device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skips GPU tests where no HIP device is visible; fails loudly if the library is missing."""
    if not _have_gpu():
        pytest.skip("no HIP device")
    from doonengine_amd import _lib
    lib = _lib.load()
    _lib.check_layout(lib)
    return lib
