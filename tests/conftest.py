import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "vapoursynth-lgcr_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity against the oracle")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def device():
    """One HIP device context for the whole GPU session."""
    if not gpu_available():
        pytest.skip("no HIP device")
    os.environ.setdefault("LGCR_CAPTURE", "1")
    from lgcr_amd import abi
    dev = abi.Device(0)
    yield dev
    dev.close()
