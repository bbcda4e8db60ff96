import os
import sys

import pytest

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
