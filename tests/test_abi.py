"""The C ABI library and the Python module load without a GPU, export what the headers
declare, and report errors with the reference's exception types (no compute calls here)."""
import ctypes
import os
import re

import numpy as np
import pytest

import vecsearch_amd as vs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(vsg_\w+)\s*\(", text, re.M))
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(vs.LIBRARY_PATH)
    names = declared_symbols()
    assert {"vsg_create", "vsg_add", "vsg_search", "vsg_search_lists", "vsg_last_error"} <= names
    for name in sorted(names):
        assert hasattr(lib, name), name


def test_device_count_and_errors_without_compute():
    lib = ctypes.CDLL(vs.LIBRARY_PATH)
    lib.vsg_last_error.restype = ctypes.c_char_p
    count = ctypes.c_int(-1)
    assert lib.vsg_device_count(ctypes.byref(count)) == 0 and count.value >= 0
    h = ctypes.c_void_p()
    assert lib.vsg_create(ctypes.c_size_t(0), 0, 0, ctypes.byref(h)) == 1  # dim == 0
    assert b"dim" in lib.vsg_last_error()
    assert lib.vsg_create(ctypes.c_size_t(8), 7, 0, ctypes.byref(h)) == 1  # bad metric
    if count.value == 0:
        assert lib.vsg_create(ctypes.c_size_t(8), 0, 0, ctypes.byref(h)) == 5  # no device
    assert lib.vsg_search(None, None, 0, 0, 0, None, None, None) == 1  # null index


def test_python_surface():
    assert callable(vs.cuda_available) and vs.cuda_available() == vs.hip_available()
    for name in ("search", "search_batch", "add", "dim", "metric", "__len__"):
        assert hasattr(vs.GpuBruteForceIndex, name)
    with pytest.raises(ValueError):
        vs.GpuBruteForceIndex(0)
    with pytest.raises(ValueError):
        vs.GpuBruteForceIndex(8, metric="hamming")
    if not vs.hip_available():
        with pytest.raises(RuntimeError):
            vs.GpuBruteForceIndex(8)


def test_helpers():
    x = np.array([[3.0, 4.0], [0.0, 0.0]], np.float64)
    n = vs.normalize(x)
    assert n.dtype == np.float32 and np.allclose(n[0], [0.6, 0.8]) and np.all(n[1] == 0)
    assert vs.recall_at_k(np.array([[1, 2, 3]]), np.array([[3, 2, 9]]), 3) == pytest.approx(2 / 3)
