"""QPS-vs-batch sweep of the engine, written in the reference's CSV format.

    python -m vecsearch_amd.gpu_bench --backends hip,hip-naive,hip-skinny,hip-bf16,hip-fp8 \\
        --out results/gpu/mi355x.csv

Rows have the columns of bench/gpu_bench.py:121-133 (backend, device, n, dim, metric, k, batch,
ms_per_batch, qps, recall_vs_cpu) and the same data and timing rules, so they read next to the
reference's results/gpu/*.csv:
  - base and queries: seeded i.i.d. Gaussian rows, L2-normalised (gpu_bench.py:25-30);
  - each batch: one warm-up call, then the median of at least 3 calls and `--min-time`
    seconds; a call is the host API (host queries in, host results out, gpu_bench.py:75-85).
Backends: "hip" (method auto), "hip-naive", "hip-skinny", "hip-tiled", "hip-bf16", "hip-fp8";
"cpu" (the reference's BruteForceIndex) when the reference's `vecsearch` package is importable.
recall_vs_cpu is measured against that CPU index when it is available, else against the
engine's exact streaming scan ("hip-naive"), whose ids are the CPU index's.
"""
import argparse
import csv
import os
import platform
import subprocess
import time

import numpy as np

from . import GpuBruteForceIndex, normalize, recall_at_k

BACKENDS = ["hip", "hip-naive", "hip-skinny", "hip-tiled", "hip-bf16", "hip-fp8", "cpu"]


def unit_vectors(n, d, seed):
    """gpu_bench.py:25-30: Gaussian rows drawn in slices of 100,000, then normalised."""
    rng = np.random.default_rng(seed)
    x = np.empty((n, d), np.float32)
    for i in range(0, n, 100_000):
        x[i:i + 100_000] = rng.standard_normal((min(100_000, n - i), d), dtype=np.float32)
    return normalize(x)


def cpu_name():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "cpu"


def gpu_name():
    try:
        out = subprocess.run(["rocm-smi", "--showproductname"], capture_output=True, text=True,
                             timeout=20).stdout
        for line in out.splitlines():
            if "Card Series" in line or "Card series" in line:
                return line.split(":", 2)[-1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return "AMD Instinct (gfx950)"


def reference_cpu_index():
    """The reference's CPU BruteForceIndex, if its Python package is installed."""
    try:
        import vecsearch  # the reference package (python/vecsearch)
    except ImportError:
        return None
    return getattr(vecsearch, "BruteForceIndex", None)


def make_backend(name, xb, metric, threads):
    """Returns (search(q, k) -> ids, device description)."""
    d = xb.shape[1]
    if name == "cpu":
        cls = reference_cpu_index()
        if cls is None:
            raise SystemExit("backend 'cpu' needs the reference's vecsearch package")
        idx = cls(d, metric)
        idx.add(xb)
        return (lambda q, k: idx.search(q, k, num_threads=threads)[0]), f"{cpu_name()}, {threads} threads"
    idx = GpuBruteForceIndex(d, metric)
    idx.add(xb)
    method = name.split("-", 1)[1] if "-" in name else "auto"
    return (lambda q, k: idx.search(q, k, method=method)[0]), gpu_name()


def time_batch(search, q, k, min_time):
    """Median seconds per call: at least 3 calls and min_time seconds, unless one call is longer."""
    search(q, k)
    times = []
    while len(times) < 3 or sum(times) < min_time:
        t = time.perf_counter()
        search(q, k)
        times.append(time.perf_counter() - t)
        if times[0] > min_time:
            break
    return float(np.median(times))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--backends", default="hip,hip-skinny,hip-bf16,hip-fp8", help=f"comma list of {BACKENDS}")
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--dim", type=int, default=384)
    ap.add_argument("--metric", default="ip", choices=["ip", "l2"])
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--batches", default="1,8,32,128,512,1024,4096")
    ap.add_argument("--max-batch", default="cpu=1024,hip-naive=1024",
                    help="per-backend batch caps (backend=batch, comma separated)")
    ap.add_argument("--threads", type=int, default=os.cpu_count())
    ap.add_argument("--min-time", type=float, default=1.0)
    ap.add_argument("--check", type=int, default=128, help="queries compared with the truth")
    ap.add_argument("--out", required=True)
    args = ap.parse_args(argv)

    backends = args.backends.split(",")
    unknown = set(backends) - set(BACKENDS)
    if unknown:
        ap.error(f"unknown backends {sorted(unknown)}")
    batches = [int(b) for b in args.batches.split(",")]
    caps = {kv.split("=")[0]: int(kv.split("=")[1]) for kv in args.max_batch.split(",") if kv}

    print(f"data: {args.n:,} x {args.dim} ({args.metric}), queries: {max(batches):,}", flush=True)
    xb = unit_vectors(args.n, args.dim, 0)
    xq = unit_vectors(max(max(batches), args.check), args.dim, 1)

    cls = reference_cpu_index()
    if cls is not None:
        cpu = cls(args.dim, args.metric)
        cpu.add(xb)
        truth = cpu.search(xq[:args.check], args.k, num_threads=args.threads)[0]
        del cpu
    else:  # the engine's exact scan returns the CPU index's ids
        exact = GpuBruteForceIndex(args.dim, args.metric)
        exact.add(xb)
        truth = exact.search(xq[:args.check], args.k, method="naive")[0]
        del exact

    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    fields = ["backend", "device", "n", "dim", "metric", "k", "batch", "ms_per_batch", "qps", "recall_vs_cpu"]
    with open(args.out, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=fields)
        w.writeheader()
        for name in backends:
            search, device = make_backend(name, xb, args.metric, args.threads)
            recall = recall_at_k(search(xq[:args.check], args.k), truth, args.k)
            print(f"{name} on {device}: recall vs cpu {recall:.4f}", flush=True)
            for b in batches:
                if b > caps.get(name, b):
                    continue
                sec = time_batch(search, xq[:b], args.k, args.min_time)
                w.writerow(dict(backend=name, device=device, n=args.n, dim=args.dim, metric=args.metric,
                                k=args.k, batch=b, ms_per_batch=round(sec * 1e3, 3), qps=round(b / sec, 1),
                                recall_vs_cpu=round(recall, 4)))
                f.flush()
                print(f"  batch {b:>5}: {sec * 1e3:10.2f} ms  {b / sec:12,.0f} QPS", flush=True)
            del search
    print(f"wrote {args.out}")


if __name__ == "__main__":
    main()
