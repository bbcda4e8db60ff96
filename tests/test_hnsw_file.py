"""Saved HNSW indexes (hnsw.cpp:369-420) as a source of base rows, and HNSWIndex::search's
exact_below dispatch (hnsw.cpp:322-327) on the GPU engine.

The fixture tests/golden/hnsw_small.vshnsw was written by the reference's own
HNSWIndex::save (tests/golden/make_golden.py); hnsw_small.npz holds the rows it saved."""
import os
import struct

import numpy as np
import pytest

import oracle
import vecsearch_amd as vs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FILE = os.path.join(GOLDEN, "hnsw_small.vshnsw")


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def rows():
    return np.load(os.path.join(GOLDEN, "hnsw_small.npz"))["x"]


def test_header_of_the_reference_file():
    assert vs.hnsw_file_info(FILE) == (16, 300, "ip")
    # the vector block is the saved rows, in id order, right after the 52-byte header
    raw = open(FILE, "rb").read()
    assert raw[:8] == b"VSHNSW01" and struct.unpack("<QQ", raw[8:24]) == (16, 300)
    np.testing.assert_array_equal(np.frombuffer(raw[52:52 + 300 * 16 * 4], np.float32).reshape(300, 16), rows())


def test_a_file_the_reference_writes_now(tmp_path):
    ref = oracle.ref()
    if ref is None:
        pytest.skip("the compiled reference is not available here")
    x = np.random.default_rng(5).standard_normal((123, 24)).astype(np.float32)
    path = tmp_path / "x.vshnsw"
    ref.hnsw_save(x, path, "l2")
    assert vs.hnsw_file_info(str(path)) == (24, 123, "l2")


def test_malformed_files(tmp_path):
    raw = open(FILE, "rb").read()
    bad = tmp_path / "bad.vshnsw"
    bad.write_bytes(b"NOTHNSW1" + raw[8:])
    with pytest.raises(ValueError):
        vs.hnsw_file_info(str(bad))
    short = tmp_path / "short.vshnsw"
    short.write_bytes(raw[:40])
    with pytest.raises(ValueError):
        vs.hnsw_file_info(str(short))
    with pytest.raises(ValueError):
        vs.hnsw_file_info(str(tmp_path / "missing.vshnsw"))


@pytest.mark.gpu
def test_gpu_index_from_the_file(tmp_path):
    x = rows()
    q = np.random.default_rng(6).standard_normal((9, 16)).astype(np.float32)
    gpu = vs.GpuBruteForceIndex.from_hnsw_file(FILE)
    assert len(gpu) == 300 and gpu.dim == 16 and gpu.metric == "ip"
    ids, dists = gpu.search(q, 7)
    ei, ed = oracle.search(x, q, 7, "ip")
    np.testing.assert_array_equal(ids, ei.astype(np.int64))
    np.testing.assert_array_equal(bits(dists), bits(ed))
    # appending the file again continues the ids
    assert gpu.add_hnsw_file(FILE) == 300 and len(gpu) == 600
    with pytest.raises(ValueError):  # dim differs
        vs.GpuBruteForceIndex(8, "ip").add_hnsw_file(FILE)
    with pytest.raises(ValueError):  # metric differs
        vs.GpuBruteForceIndex(16, "l2").add_hnsw_file(FILE)
    truncated = tmp_path / "t.vshnsw"
    truncated.write_bytes(open(FILE, "rb").read()[:52 + 100 * 16 * 4])
    with pytest.raises(ValueError):
        vs.GpuBruteForceIndex(16, "ip").add_hnsw_file(str(truncated))


@pytest.mark.gpu
def test_exact_below_dispatch():
    x = rows()
    gpu = vs.GpuBruteForceIndex.from_hnsw_file(FILE)
    q = np.random.default_rng(7).standard_normal((5, 16)).astype(np.float32)
    mask = np.zeros(300, bool)
    mask[[3, 50, 51, 120, 299]] = True
    got = gpu.search_exact_below(q, 4, mask, exact_below=5)  # 5 allowed <= 5: exact scan
    assert got is not None
    ei, ed = oracle.search(x, q, 4, "ip", mask=mask)
    np.testing.assert_array_equal(got[0], ei.astype(np.int64))
    np.testing.assert_array_equal(bits(got[1]), bits(ed))
    if oracle.ref() is not None:  # the reference's own dispatch gives the same answer
        ri, rd = oracle.ref().hnsw_exact(x, q, 4, mask, "ip")
        np.testing.assert_array_equal(got[0], ri.astype(np.int64))
    assert gpu.search_exact_below(q, 4, mask, exact_below=4) is None  # graph search instead
    assert gpu.search_exact_below(q, 4, None, exact_below=1000) is None  # no filter
    ids, dists = gpu.search_exact_below(q, 4, np.zeros(300, bool), exact_below=0)  # hnsw.cpp:326
    assert (ids == -1).all() and np.isinf(dists).all()
