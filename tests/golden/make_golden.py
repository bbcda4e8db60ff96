"""Writes the golden vectors in tests/golden/*.npz from the compiled reference CPU index.

    make -C oracle ref          # builds oracle/_ref/libvsref.so from the reference sources
    python tests/golden/make_golden.py

Every fixture holds seeded inputs and the reference's outputs (ids and float32 distances, the
latter compared bit for bit):
  - distances.npz   reference dot / l2_sq / distance for dims 1..49 and 128, 383-385, 767-771
                    (covers the 16-dim steps, the 8-dim step and every scalar-tail length)
  - search_d{d}.npz BruteForceIndex::search_batch (1 and 3 threads) for l2 and ip
  - filtered.npz    HNSWIndex::search(filter=mask, exact_below=n): the exact filtered scan
  - line.npz        the known answer of tests/cpp/test_brute_force.cpp:9-20
  - hnsw_small.vshnsw / hnsw_small.npz   HNSWIndex::save of 300 x 16 rows (ip) by the reference,
                    and those rows: the saved-index format the engine reads vectors from
                    (`python tests/golden/make_golden.py hnsw` writes only these two)
Run in the development container only (the reference tree is not on the GPU machines).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402


def hnsw_file(ref):
    rng = np.random.default_rng(20260102)
    x = rng.standard_normal((300, 16)).astype(np.float32)
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    ref.hnsw_save(x, os.path.join(HERE, "hnsw_small.vshnsw"), "ip")
    np.savez_compressed(os.path.join(HERE, "hnsw_small.npz"), x=x)


def main():
    ref = oracle.ref()
    if len(sys.argv) > 1 and sys.argv[1] == "hnsw":
        if ref is None:
            sys.exit("oracle/_ref/libvsref.so is missing: run `make -C oracle ref` first")
        hnsw_file(ref)
        return
    if ref is None:
        sys.exit("oracle/_ref/libvsref.so is missing: run `make -C oracle ref` first")
    rng = np.random.default_rng(20260101)

    dims = list(range(1, 50)) + [128, 383, 384, 385, 767, 768, 771]
    out = {}
    for d in dims:
        a = rng.standard_normal((8, d)).astype(np.float32)
        b = rng.standard_normal((8, d)).astype(np.float32)
        out[f"a{d}"], out[f"b{d}"] = a, b
        out[f"dot{d}"] = np.array([ref.dot(a[i], b[i]) for i in range(8)], np.float32)
        out[f"l2{d}"] = np.array([ref.l2sq(a[i], b[i]) for i in range(8)], np.float32)
        out[f"ipd{d}"] = np.array([ref.distance("ip", a[i], b[i]) for i in range(8)], np.float32)
    out["dims"] = np.array(dims)
    np.savez_compressed(os.path.join(HERE, "distances.npz"), **out)

    for d, n, nq, k in [(1, 300, 5, 10), (7, 300, 5, 10), (13, 300, 5, 20), (37, 500, 13, 10),
                        (64, 500, 7, 17), (128, 160, 5, 10), (384, 100, 5, 10), (768, 60, 3, 10),
                        (385, 80, 4, 7)]:
        x = rng.standard_normal((n, d)).astype(np.float32)
        q = rng.standard_normal((nq, d)).astype(np.float32)
        xn = x / np.linalg.norm(x, axis=1, keepdims=True)
        qn = q / np.linalg.norm(q, axis=1, keepdims=True)
        res = {"x": x, "q": q, "xn": xn.astype(np.float32), "qn": qn.astype(np.float32), "k": k}
        for metric, xx, qq in (("l2", x, q), ("ip", xn, qn)):
            idx = ref.index(d, metric)
            idx.add(xx)
            for th in (1, 3):
                i, dd = idx.search(qq, k, th)
                res[f"{metric}_ids_t{th}"], res[f"{metric}_dists_t{th}"] = i, dd
            # k > n: padding (test_bindings.py:36-37)
            i, dd = idx.search(qq[:1], n + 3, 2)
            res[f"{metric}_pad_ids"], res[f"{metric}_pad_dists"] = i, dd
        np.savez_compressed(os.path.join(HERE, f"search_d{d}.npz"), **res)

    # exact filtered scan through the HNSW binding's path (test_bindings.py:76-90)
    x = rng.standard_normal((1500, 64)).astype(np.float32)
    q = rng.standard_normal((20, 64)).astype(np.float32)
    masks = {"third": np.arange(1500) % 3 == 0, "sparse": rng.random(1500) < 0.02,
             "two": np.isin(np.arange(1500), [7, 1400])}
    res = {"x": x, "q": q}
    for name, m in masks.items():
        res[f"mask_{name}"] = m
        for metric in ("l2", "ip"):
            i, dd = ref.hnsw_exact(x, q, 10, m, metric, threads=2)
            res[f"{name}_{metric}_ids"], res[f"{name}_{metric}_dists"] = i, dd
    np.savez_compressed(os.path.join(HERE, "filtered.npz"), **res)

    # known answer on a line (test_brute_force.cpp:9-20)
    pts = np.array([[float(i), 0.0] for i in range(10)], np.float32)
    idx = ref.index(2, "l2")
    idx.add(pts)
    i, dd = idx.search_one(np.array([3.2, 0.0], np.float32), 3)
    hnsw_file(ref)
    np.savez_compressed(os.path.join(HERE, "line.npz"), x=pts, q=np.array([3.2, 0.0], np.float32),
                        ids=i, dists=dd)
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
