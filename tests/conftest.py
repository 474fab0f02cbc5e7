import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "ssg.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libssg_hip.so")


def _gpu_present() -> bool:
    try:
        import torch  # noqa: F401  (device count does not initialise HIP)
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not _gpu_present():
        pytest.fail("GPU test collected but no HIP device is visible")
    import ssgamd
    ssgamd.load_library()
    return ssgamd
