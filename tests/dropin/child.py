"""Runs in its own process: the reference's Python module, built unchanged from its sources
(python/bindings.cpp, src/brute_force.cpp, src/hnsw.cpp, src/bm25.cpp; oracle/Makefile target
`dropin`) with this engine as its GpuBruteForceIndex, driven by the reference's own GPU test
cases (tests/python/test_gpu.py:17-70) with the bar raised to bit equality against the
reference's own CPU BruteForceIndex in the same module.

A separate process because the module registers vecsearch::GpuBruteForceIndex with pybind11
under the same C++ type as vecsearch_amd._vecsearch does. Prints one line per case and
"dropin ok" at the end; exits non-zero on the first mismatch.
"""
import importlib.util
import sys

import numpy as np


def load(path):
    spec = importlib.util.spec_from_file_location("_vecsearch", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def data(n, d, seed, normalize=False):
    x = np.random.default_rng(seed).standard_normal((n, d), dtype=np.float32)
    if normalize:
        x /= np.maximum(np.linalg.norm(x, axis=1, keepdims=True), 1e-12)
    return x


def same(a, b):
    (ia, da), (ib, db) = a, b
    return ia.shape == ib.shape and np.array_equal(ia, ib) and np.array_equal(da.view(np.uint32), db.view(np.uint32))


def main(path):
    vs = load(path)
    assert vs.cuda_available(), "the module sees no GPU"
    for metric in ("l2", "ip"):
        for d in (37, 128):
            for nq in (1, 5, 300):
                x, q = data(3000, d, 1, metric == "ip"), data(nq, d, 2, metric == "ip")
                cpu = vs.BruteForceIndex(d, metric)
                cpu.add(x)
                gpu = vs.GpuBruteForceIndex(d, metric)
                gpu.add(x[:1000])
                gpu.add(x[1000:])
                want = cpu.search(q, 10)
                for method in ("naive", "skinny", "tiled", "auto"):
                    got = gpu.search(q, 10, method=method)
                    assert same(got, want), f"{metric} d={d} nq={nq} method={method}"
                    assert same(gpu.search(q, 10, method), want), "positional method"
                print(f"case {metric} d={d} nq={nq}: 4 methods bit-identical to BruteForceIndex", flush=True)
    x, q = data(20000, 64, 3), data(40, 64, 4)
    cpu, gpu = vs.BruteForceIndex(64), vs.GpuBruteForceIndex(64)
    cpu.add(x)
    gpu.add(x)
    for k in (1, 16, 17, 100, 128):
        assert same(gpu.search(q, k), cpu.search(q, k)), f"k={k}"
    print("case k buckets 1/16/17/100/128: bit-identical", flush=True)
    gpu = vs.GpuBruteForceIndex(8)
    ids, dists = gpu.search(np.zeros((2, 8), np.float32), 5)
    assert (ids == -1).all() and np.isinf(dists).all()
    gpu.add(data(3, 8, 7))
    ids, dists = gpu.search(np.zeros(8, np.float32), 5)
    assert (ids[0, :3] >= 0).all() and (ids[0, 3:] == -1).all() and np.isinf(dists[0, 3:]).all()
    for bad in (dict(k=129), dict(k=5, method="fast")):
        try:
            gpu.search(np.zeros(8, np.float32), **bad)
        except ValueError:
            continue
        raise AssertionError(f"no ValueError for {bad}")
    print("case missing results and limits: as the reference", flush=True)
    print("dropin ok", flush=True)


if __name__ == "__main__":
    main(sys.argv[1])
