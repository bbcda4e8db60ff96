"""vsg_search_device_noalloc through the C ABI: the same bits as vsg_search_device on two calls
in a row (the second smaller), no device memory taken, and the scratch limit enforced."""
import ctypes

import numpy as np
import pytest

import oracle
import vecsearch_amd as vs
from vecsearch_amd import datagen as dg

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

c_sz, c_vp = ctypes.c_size_t, ctypes.c_void_p
METHOD = {"auto": 0, "naive": 1, "skinny": 2, "tiled": 3}


def lib():
    L = ctypes.CDLL(vs.LIBRARY_PATH)
    L.vsg_last_error.restype = ctypes.c_char_p
    L.vsg_create.argtypes = [c_sz, ctypes.c_int, ctypes.c_int, ctypes.POINTER(c_vp)]
    L.vsg_destroy.argtypes = [c_vp]
    L.vsg_destroy.restype = None
    L.vsg_add.argtypes = [c_vp, c_vp, c_sz]
    sig = [c_vp, c_vp, c_sz, c_sz, ctypes.c_int, c_vp, c_vp, c_vp, c_vp]
    L.vsg_search_device.argtypes = sig
    L.vsg_search_device_noalloc.argtypes = sig
    L.vsg_set_param.argtypes = [c_vp, ctypes.c_char_p, ctypes.c_int64]
    return L


def bits(t):
    return t.cpu().numpy().view(np.uint32)


def outputs(nq, k):
    return (torch.full((nq, k), -7, dtype=torch.int32, device="cuda"),
            torch.full((nq, k), -7.0, dtype=torch.float32, device="cuda"))


def run(L, fn, h, dq, nq, k, method, dmask, out=None, check_free=False):
    ids, dd = out if out is not None else outputs(nq, k)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0] if check_free else None
    rc = getattr(L, fn)(h, dq.data_ptr(), nq, k, METHOD[method], dmask, ids.data_ptr(), dd.data_ptr(), None)
    assert rc == 0, L.vsg_last_error()
    torch.cuda.synchronize()
    if check_free:
        assert torch.cuda.mem_get_info()[0] == free0  # no device memory taken by the call
    return ids, dd


def test_noalloc_matches_search_device():
    L = lib()
    n, d = 300_000, 64
    x = dg.mixture(n, d, seed=50)
    h = c_vp()
    assert L.vsg_create(d, 1, 0, ctypes.byref(h)) == 0, L.vsg_last_error()
    try:
        assert L.vsg_add(h, x.ctypes.data, n) == 0, L.vsg_last_error()
        q = torch.from_numpy(dg.mixture(96, d, seed=51)).cuda()
        mask = torch.from_numpy((np.random.default_rng(52).random(n) < 0.5).astype(np.uint8)).cuda()
        # one call first, so the runtime has loaded the scan's code object before memory is read
        run(L, "vsg_search_device_noalloc", h, q, 1, 10, "auto", None)
        # two calls in a row, the second with a smaller input (fewer queries, smaller k)
        for nq, k, method, dmask in ((96, 100, "auto", None), (5, 10, "naive", mask.data_ptr())):
            a = run(L, "vsg_search_device", h, q, nq, k, method, dmask)
            b = run(L, "vsg_search_device_noalloc", h, q, nq, k, method, dmask, out=outputs(nq, k),
                    check_free=True)
            np.testing.assert_array_equal(bits(a[0]), bits(b[0]))
            np.testing.assert_array_equal(bits(a[1]), bits(b[1]))
            m = None if dmask is None else mask.cpu().numpy()
            ei, ed = oracle.search(x, q[:nq].cpu().numpy(), k, "ip", threads=8, mask=m)
            np.testing.assert_array_equal(bits(b[0]), ei)
            np.testing.assert_array_equal(bits(b[1]), ed.view(np.uint32))
        # one pass of 8 queries x 1,024 lists x k = 128 needs 8 MiB: beyond the 4 MiB scratch
        assert L.vsg_set_param(h, b"scan_waves_per_cu", 32) == 0
        ids = torch.empty((8, 128), dtype=torch.int32, device="cuda")
        dd = torch.empty((8, 128), dtype=torch.float32, device="cuda")
        rc = L.vsg_search_device_noalloc(h, q.data_ptr(), 8, 128, 0, None, ids.data_ptr(), dd.data_ptr(), None)
        assert rc == 1 and b"bytes of list space" in L.vsg_last_error()
    finally:
        L.vsg_destroy(h)


def test_noalloc_empty_index():
    L = lib()
    h = c_vp()
    assert L.vsg_create(16, 0, 0, ctypes.byref(h)) == 0
    try:
        q = torch.zeros((3, 16), dtype=torch.float32, device="cuda")
        ids, dd = run(L, "vsg_search_device_noalloc", h, q, 3, 4, "auto", None)
        assert (ids.cpu().numpy() == -1).all() and np.isinf(dd.cpu().numpy()).all()
    finally:
        L.vsg_destroy(h)
