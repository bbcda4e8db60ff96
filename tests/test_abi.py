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
    assert {"vsg_create", "vsg_add", "vsg_search", "vsg_search_lists", "vsg_last_error", "vsg_get_param",
            "vsg_graph_create", "vsg_graph_create_lists", "vsg_graph_launch", "vsg_server_create", "vsg_server_submit",
            "vsg_server_wait"} <= names
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
    # serving loop, captured searches and parameters reject null handles without touching a GPU
    srv = ctypes.c_void_p()
    assert lib.vsg_server_create(None, None, ctypes.byref(srv)) == 1
    ticket = ctypes.c_uint64()
    assert lib.vsg_server_submit(None, None, 0, 0, None, 0, None, None, ctypes.byref(ticket)) == 1
    assert lib.vsg_server_wait(None, ctypes.c_uint64(1), None) == 1
    g = ctypes.c_void_p()
    assert lib.vsg_graph_create(None, None, 1, 1, 0, None, None, ctypes.byref(g)) == 1
    assert lib.vsg_graph_create_lists(None, None, 1, None, 1, None, None, 0, None, None, ctypes.byref(g)) == 1
    assert lib.vsg_graph_launch(None, None) == 1
    v = ctypes.c_int64()
    assert lib.vsg_get_param(None, b"lists_qb", ctypes.byref(v)) == 1
    assert lib.vsg_set_param(None, b"lists_qb", ctypes.c_int64(4)) == 1
    lib.vsg_server_destroy(None)
    lib.vsg_graph_destroy(None)


def test_python_surface():
    assert callable(vs.cuda_available) and vs.cuda_available() == vs.hip_available()
    for name in ("search", "search_batch", "search_device", "add", "dim", "metric", "__len__",
                 "set_param", "get_param", "kernel_timing", "kernel_times", "short_stats"):
        assert hasattr(vs.GpuBruteForceIndex, name)
    for name in ("submit", "wait", "stats", "close"):
        assert hasattr(vs.Server, name)
    assert hasattr(vs.SearchGraph, "launch") and vs.ALL_ROWS == 2**64 - 1
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
