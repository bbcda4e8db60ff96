"""CPU oracle for the exact k-NN path (ctypes over oracle/libvs_oracle.so).

TEST INFRASTRUCTURE ONLY. The tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` use it as the checker; the product package ``vecsearch_amd`` never imports it.

It restates the reference's CPU arithmetic and result semantics:
  include/vecsearch/distance.h:18-73   hsum256, l2_sq, dot, distance (bit-exact op order)
  src/brute_force.cpp:116-169          BruteForceIndex::search_batch (top-k, (dist, id) merge,
                                       kNone / +inf padding)
  src/hnsw.cpp:322-354                 HNSWIndex::search / search_exact (filtered exact scan)

``ref()`` returns the compiled reference CPU index (oracle/_ref/libvsref.so, built from the
reference's own sources by oracle/Makefile) when that library exists, else None.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libvs_oracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libvsref.so")
NONE = 0xFFFFFFFF
METRICS = {"l2": 0, "ip": 1, "cosine": 1}

_f32p = ctypes.POINTER(ctypes.c_float)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_sz = ctypes.c_size_t

_lib = None
_ref = None


def build():
    """Compile the C oracle (and oracle/_ref when the reference tree is present)."""
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
    if os.path.isdir(os.environ.get("VS_REFERENCE", "/root/reference")):
        subprocess.run(["make", "-s", "-C", HERE, "ref", "dropin",
                        "REF=" + os.environ.get("VS_REFERENCE", "/root/reference")], check=True)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = ctypes.CDLL(LIB)
        lib.vso_dot.restype = ctypes.c_float
        lib.vso_dot.argtypes = [_f32p, _f32p, _sz]
        lib.vso_l2sq.restype = ctypes.c_float
        lib.vso_l2sq.argtypes = [_f32p, _f32p, _sz]
        lib.vso_distance.restype = ctypes.c_float
        lib.vso_distance.argtypes = [ctypes.c_int, _f32p, _f32p, _sz]
        lib.vso_dot_f64.restype = ctypes.c_double
        lib.vso_dot_f64.argtypes = [_f32p, _f32p, _sz]
        lib.vso_l2sq_f64.restype = ctypes.c_double
        lib.vso_l2sq_f64.argtypes = [_f32p, _f32p, _sz]
        lib.vso_search_batch.restype = None
        lib.vso_search_batch.argtypes = [_f32p, _sz, _sz, ctypes.c_int, _f32p, _sz, _sz, _u32p,
                                         _f32p, ctypes.c_int]
        lib.vso_search_masked.restype = None
        lib.vso_search_masked.argtypes = [_f32p, _sz, _sz, ctypes.c_int, _f32p, _sz, _sz, _u8p,
                                          _u32p, _f32p, ctypes.c_int]
        lib.vso_search_lists.restype = None
        lib.vso_search_lists.argtypes = [_f32p, _sz, _sz, ctypes.c_int, _f32p, _sz, _u32p, _sz,
                                         _u64p, _u32p, _u32p, _f32p, ctypes.c_int]
        _lib = lib
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _metric(m):
    return METRICS[m] if isinstance(m, str) else int(m)


def dot(a, b):
    a, b = _f32(a), _f32(b)
    return _load().vso_dot(_p(a, _f32p), _p(b, _f32p), a.size)


def l2sq(a, b):
    a, b = _f32(a), _f32(b)
    return _load().vso_l2sq(_p(a, _f32p), _p(b, _f32p), a.size)


def distance(metric, a, b):
    a, b = _f32(a), _f32(b)
    return _load().vso_distance(_metric(metric), _p(a, _f32p), _p(b, _f32p), a.size)


def distance_f64(metric, a, b):
    """FP64 value of the same distance (1 - dot or squared L2), for error reporting."""
    a, b = _f32(a), _f32(b)
    lib = _load()
    if _metric(metric) == 0:
        return lib.vso_l2sq_f64(_p(a, _f32p), _p(b, _f32p), a.size)
    return 1.0 - lib.vso_dot_f64(_p(a, _f32p), _p(b, _f32p), a.size)


def search(x, q, k, metric="l2", threads=1, mask=None):
    """Exact top-k: (ids uint32[nq,k], dists float32[nq,k]); missing = 0xFFFFFFFF / +inf."""
    x, q = _f32(x), _f32(q)
    if q.ndim == 1:
        q = q[None]
    n, d = x.shape if x.size else (0, q.shape[1])
    nq = q.shape[0]
    ids = np.full((nq, k), NONE, np.uint32)
    dists = np.full((nq, k), np.inf, np.float32)
    if nq == 0 or k == 0:
        return ids, dists
    lib = _load()
    if mask is None:
        lib.vso_search_batch(_p(x, _f32p), n, d, _metric(metric), _p(q, _f32p), nq, k,
                             _p(ids, _u32p), _p(dists, _f32p), threads)
    else:
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        assert m.shape == (n,)
        lib.vso_search_masked(_p(x, _f32p), n, d, _metric(metric), _p(q, _f32p), nq, k,
                              _p(m, _u8p), _p(ids, _u32p), _p(dists, _f32p), threads)
    return ids, dists


ALL_ROWS = 2**64 - 1  # a range beginning with this value is unfiltered (every row)


def search_lists(x, q, ks, ranges, allow_ids, metric="l2", threads=1, kmax=None):
    """Per-query k and allowed-id lists (query i scans allow_ids[ranges[i,0]:ranges[i,1]], or
    every row when ranges[i,0] == ALL_ROWS)."""
    x, q = _f32(x), _f32(q)
    ks = np.ascontiguousarray(ks, dtype=np.uint32)
    off = np.ascontiguousarray(ranges, dtype=np.uint64).reshape(-1)
    al = np.ascontiguousarray(allow_ids, dtype=np.uint32)
    nq, d = q.shape
    kmax = int(ks.max()) if kmax is None and nq else (kmax or 0)
    ids = np.full((nq, kmax), NONE, np.uint32)
    dists = np.full((nq, kmax), np.inf, np.float32)
    if nq == 0 or kmax == 0:
        return ids, dists
    if al.size == 0:
        al = np.zeros(1, np.uint32)
    _load().vso_search_lists(_p(x, _f32p), x.shape[0], d, _metric(metric), _p(q, _f32p), nq,
                             _p(ks, _u32p), kmax, _p(off, _u64p), _p(al, _u32p), _p(ids, _u32p),
                             _p(dists, _f32p), threads)
    return ids, dists


# ---- the compiled reference (optional) -------------------------------------------------------

class _RefIndex:
    """BruteForceIndex from the reference sources (brute_force.h:8-32)."""

    def __init__(self, lib, dim, metric):
        self.lib, self.dim = lib, dim
        self.h = lib.vsref_bf_create(dim, _metric(metric))

    def add(self, x):
        x = _f32(x)
        self.lib.vsref_bf_add(self.h, _p(x, _f32p), x.shape[0])

    def search(self, q, k, threads=1):
        q = _f32(q)
        if q.ndim == 1:
            q = q[None]
        nq = q.shape[0]
        ids = np.empty((nq, k), np.uint32)
        dists = np.empty((nq, k), np.float32)
        self.lib.vsref_bf_search_batch(self.h, _p(q, _f32p), nq, k, _p(ids, _u32p),
                                       _p(dists, _f32p), threads)
        return ids, dists

    def search_one(self, q, k):
        q = _f32(q)
        ids = np.full(k, NONE, np.uint32)
        dists = np.full(k, np.inf, np.float32)
        m = self.lib.vsref_bf_search_one(self.h, _p(q, _f32p), k, _p(ids, _u32p),
                                         _p(dists, _f32p))
        return ids[:m], dists[:m]

    def __del__(self):
        try:
            self.lib.vsref_bf_destroy(self.h)
        except Exception:
            pass


class _Ref:
    def __init__(self, lib):
        self.lib = lib
        vp = ctypes.c_void_p
        for name in ("vsref_dot", "vsref_l2sq"):
            getattr(lib, name).restype = ctypes.c_float
            getattr(lib, name).argtypes = [_f32p, _f32p, _sz]
        lib.vsref_distance.restype = ctypes.c_float
        lib.vsref_distance.argtypes = [ctypes.c_int, _f32p, _f32p, _sz]
        lib.vsref_bf_create.restype = vp
        lib.vsref_bf_create.argtypes = [_sz, ctypes.c_int]
        lib.vsref_bf_destroy.restype = None
        lib.vsref_bf_destroy.argtypes = [vp]
        lib.vsref_bf_add.restype = None
        lib.vsref_bf_add.argtypes = [vp, _f32p, _sz]
        lib.vsref_bf_search_batch.restype = None
        lib.vsref_bf_search_batch.argtypes = [vp, _f32p, _sz, _sz, _u32p, _f32p, ctypes.c_int]
        lib.vsref_bf_search_one.restype = _sz
        lib.vsref_bf_search_one.argtypes = [vp, _f32p, _sz, _u32p, _f32p]
        lib.vsref_hnsw_filtered.restype = None
        lib.vsref_hnsw_filtered.argtypes = [_f32p, _sz, _sz, ctypes.c_int, _f32p, _sz, _sz, _u8p,
                                            _sz, _u32p, _f32p, ctypes.c_int]

    def dot(self, a, b):
        a, b = _f32(a), _f32(b)
        return self.lib.vsref_dot(_p(a, _f32p), _p(b, _f32p), a.size)

    def l2sq(self, a, b):
        a, b = _f32(a), _f32(b)
        return self.lib.vsref_l2sq(_p(a, _f32p), _p(b, _f32p), a.size)

    def distance(self, metric, a, b):
        a, b = _f32(a), _f32(b)
        return self.lib.vsref_distance(_metric(metric), _p(a, _f32p), _p(b, _f32p), a.size)

    def index(self, dim, metric="l2"):
        return _RefIndex(self.lib, dim, metric)

    def search(self, x, q, k, metric="l2", threads=1):
        idx = self.index(x.shape[1], metric)
        idx.add(x)
        return idx.search(q, k, threads)

    def hnsw_save(self, x, path, metric="l2"):
        """HNSWIndex(x).save(path) from the reference sources (hnsw.cpp:375-389)."""
        x = _f32(x)
        self.lib.vsref_hnsw_save.restype = ctypes.c_int
        self.lib.vsref_hnsw_save.argtypes = [_f32p, _sz, _sz, ctypes.c_int, ctypes.c_char_p]
        rc = self.lib.vsref_hnsw_save(_p(x, _f32p), x.shape[0], x.shape[1], _metric(metric),
                                      str(path).encode())
        if rc:
            raise RuntimeError("reference HNSWIndex::save failed")

    def hnsw_exact(self, x, q, k, mask, metric="l2", threads=1):
        """HNSWIndex.search(filter=mask, exact_below=len(x)): the exact filtered scan."""
        x, q = _f32(x), _f32(q)
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        nq = q.shape[0]
        ids = np.empty((nq, k), np.uint32)
        dists = np.empty((nq, k), np.float32)
        self.lib.vsref_hnsw_filtered(_p(x, _f32p), x.shape[0], x.shape[1], _metric(metric),
                                     _p(q, _f32p), nq, k, _p(m, _u8p), x.shape[0], _p(ids, _u32p),
                                     _p(dists, _f32p), threads)
        return ids, dists


def ref():
    """The compiled reference CPU index, or None when oracle/_ref/libvsref.so is absent."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        _ref = _Ref(ctypes.CDLL(REF_LIB))
    return _ref
