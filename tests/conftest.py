import os
import sys

import pytest

# PyTorch-ROCm ships its own HIP runtime. Loading it before the engine lets the engine bind to
# that same runtime (one HIP runtime per process), which the GPU tests that hand torch device
# buffers to the engine rely on.
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP engine)")


def _gpu_visible():
    import vecsearch_amd as vs  # raises ImportError if the engine is not built

    return vs.hip_available()


def pytest_collection_modifyitems(config, items):
    if not any("gpu" in item.keywords for item in items):
        return
    try:
        visible = _gpu_visible()
    except ImportError:
        visible = False
    if visible:
        return
    skip = pytest.mark.skip(reason="no GPU visible to the HIP runtime")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)
