"""vecsearch_amd: exact k-NN on an AMD Instinct MI355X, drop-in for vecsearch's GPU index.

Mirrors the reference's Python surface for this path (python/vecsearch/__init__.py:1-24,
python/bindings.cpp:151-197):

    GpuBruteForceIndex(dim, metric="l2", device=0)
        .add(x)                                    # (n, dim) float32 (other dtypes are cast)
        .search(queries, k=10, method="auto", filter=None)
            -> (ids int64[nq, k], dists float32[nq, k]), nearest first, missing = -1 / inf
            (a mask given as the third argument is the filter: search(queries, k, filter))
        .search_batch(queries, ks, ranges, allow_ids)      # per-query k and allowed ids
    cuda_available(), hip_available(), normalize(x), recall_at_k(found, truth, k)

Distances are the reference CPU index's FP32 values bit for bit (1 - <q, x> for "ip" /
"cosine", squared L2 for "l2"), and ties go to the lower id.

The compiled engine (libvsg_hip.so + _vecsearch) is required: importing this package raises
ImportError when it has not been built (``python vecsearch_amd/build.py``). There is no CPU
fallback.
"""
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

try:
    from . import _vecsearch
except ImportError as e:  # no silent fallback: the product is the HIP engine
    raise ImportError(
        "vecsearch_amd: the HIP engine is not built (run `python vecsearch_amd/build.py`): "
        f"{e}") from e

from ._vecsearch import (ALL_ROWS, GpuBruteForceIndex, Metric, SearchGraph, Server, cuda_available,
                         hip_available, hnsw_file_info)

LIBRARY_PATH = os.path.join(_HERE, "libvsg_hip.so")

__all__ = ["ALL_ROWS", "GpuBruteForceIndex", "Metric", "SearchGraph", "Server", "hnsw_file_info", "cuda_available", "hip_available", "normalize",
           "recall_at_k", "LIBRARY_PATH"]
__version__ = "0.1.0"


def normalize(x: np.ndarray) -> np.ndarray:
    """L2-normalize rows so inner product == cosine similarity (python/vecsearch/__init__.py:15-18)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    return x / np.maximum(np.linalg.norm(x, axis=1, keepdims=True), 1e-12)


def recall_at_k(found: np.ndarray, truth: np.ndarray, k: int = 10) -> float:
    """Fraction of the true top-k found in the returned top-k (python/vecsearch/__init__.py:21-24)."""
    hits = sum(len(set(f[:k]) & set(t[:k])) for f, t in zip(found, truth))
    return hits / (len(found) * k)
