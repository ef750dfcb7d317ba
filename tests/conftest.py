import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def gpu_available():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(params=["oracle", pytest.param("gpu", marks=pytest.mark.gpu)])
def TAT(request):
    """The TAT-shaped module under test: the CPU oracle facade, or the GPU
    package (which must load its native library; no fallback)."""
    if request.param == "oracle":
        from oracle import pytat

        return pytat
    if not gpu_available():
        pytest.fail("gpu test selected but no HIP device is visible")
    from tnsp_amd import TAT as gpu_tat

    return gpu_tat
