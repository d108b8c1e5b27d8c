import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through libvgpu.so)")


@pytest.fixture(scope="session")
def dev():
    """Initializes the device through the C ABI (the HIP path must load: no fallback)."""
    import torch
    from vectisdb_amd import capi
    assert torch.cuda.is_available(), "GPU test without a GPU"
    capi.check(capi.load().vgpu_device_init(0))
    return torch.device("cuda:0")


@pytest.fixture(autouse=True)
def _device_before_library(request):
    """GPU tests bring the device up through torch first: the HIP runtime torch ships must initialize before
    the library's own HIP calls, or torch no longer sees the device in that process."""
    if request.node.get_closest_marker("gpu") is not None:
        request.getfixturevalue("dev")
