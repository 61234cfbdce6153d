import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device and the built HIP library")


def gpu_available():
    try:
        import torch  # noqa: F401  (device enumeration only)
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def hip_lib():
    """The product library; GPU tests fail loudly when it is missing."""
    from acoustics_amd import capi
    return capi.load_library()
