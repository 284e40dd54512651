import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def _gpu_available() -> bool:
    try:
        import ctypes

        from isai_kernel_amd import lib

        n = ctypes.c_int(0)
        return lib.isai_device_count(ctypes.byref(n)) == 0 and n.value > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skips nothing: a gpu-marked test on a machine without a GPU must fail loudly."""
    if not _gpu_available():
        pytest.fail("gpu test selected but no HIP device is visible")
    return True
