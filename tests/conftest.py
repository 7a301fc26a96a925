import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (runs the HIP library)")


@pytest.fixture(scope="session")
def lib():
    from pedalsky_amd import abi
    return abi.load()


@pytest.fixture(scope="session")
def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def require_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected without a GPU: the HIP path has no CPU fallback")
