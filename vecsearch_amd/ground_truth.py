"""Exact top-k of every query -> gt.npy, on the GPU engine.

    python -m vecsearch_amd.ground_truth --data data/tweets --k 100

Replaces bench/ground_truth.py, which used FAISS IndexFlatIP because the CPU BruteForceIndex
was too slow: reads base.npy and queries.npy from --data and writes gt.npy (int64, nq x k,
nearest first) next to them, the file the ANN benchmarks read (ground_truth.py:25-29). The ids
are the CPU BruteForceIndex's exactly (ties to the lower id), which the GPU tests check against
the oracle; the original script's FAISS sample comparison (ground_truth.py:32-41) is not needed.
"""
import argparse
import os
import time

import numpy as np

from . import GpuBruteForceIndex


def ground_truth(xb, xq, k, metric="ip", batch=4096):
    """(nq, k) int64 exact neighbour ids of xq in xb."""
    index = GpuBruteForceIndex(xb.shape[1], metric)
    index.add(xb)
    out = np.empty((len(xq), k), np.int64)
    for i in range(0, len(xq), batch):
        out[i:i + batch] = index.search(xq[i:i + batch], k)[0]
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--data", default="data/tweets")
    ap.add_argument("--k", type=int, default=100)
    ap.add_argument("--metric", default="ip", choices=["ip", "l2"])
    args = ap.parse_args(argv)

    xb = np.load(os.path.join(args.data, "base.npy"))
    xq = np.load(os.path.join(args.data, "queries.npy"))
    t = time.time()
    gt = ground_truth(xb, xq, args.k, args.metric)
    print(f"exact search {len(xq):,} x {len(xb):,} in {time.time() - t:.1f}s")
    np.save(os.path.join(args.data, "gt.npy"), gt)


if __name__ == "__main__":
    main()
