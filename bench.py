"""Exact top-k QPS of the MI355X engine on the 1M x 384 cosine workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B] [--sweep 1,128,4096]

A step is one search of a batch of B queries (drawn in turn from 10,000 held-out queries)
against the resident base: 1,000,000 x 384 float32 unit rows from a seeded 1,000-cluster
Gaussian mixture, cosine (inner product), k = 10. Queries and results stay in HBM during
the timed region; the PCIe-inclusive rate (host queries in, host results out, as the
reference's bench/gpu_bench.py:75-85 times it) is reported per batch size in "sweep".

One JSON line on stdout carries the contract fields plus:
  roofline      the dominant kernel (the exact streaming scan) against HBM bandwidth: algorithmic
                bytes per launch = N*d*4 + b*d*4 + b*k*8 (b = queries per pass) over its average
                launch time, measured with HIP events on the launch stream during the timed steps;
  cpu_baseline  the reference CPU index (oracle/_ref, BruteForceIndex::search_batch) or the C
                oracle timed on this host's cores over a bounded sample of the same workload;
  sweep         per batch size: device-resident QPS, and PCIe-inclusive QPS / p50 / p99 ms.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E
FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 (vector packed / matrix)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=1)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--dim", type=int, default=384)
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--queries", type=int, default=10_000)
    ap.add_argument("--method", default="auto")
    ap.add_argument("--sweep", default="1,8,32,128,512,1024,4096",
                    help="batch sizes for the per-batch table ('' to skip)")
    ap.add_argument("--sweep-seconds", type=float, default=0.5)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU baseline budget (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = min(16, os.cpu_count())")
    return ap.parse_args()


def traffic_from_profiles(name):
    """HBM bytes per launch of the scan, as measured by rocprofv3 --pmc and committed in profiles/."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.strip().startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1].strip().split()[0])
    return None


def cpu_baseline(base, queries, k, seconds, threads):
    import oracle

    ref = oracle.ref()
    kind = "reference" if ref is not None else "port"
    nq = 16
    q = queries[:nq]
    if ref is not None:
        idx = ref.index(base.shape[1], "ip")
        idx.add(base)
        run = lambda: idx.search(q, k, threads)  # noqa: E731
    else:
        run = lambda: oracle.search(base, q, k, "ip", threads=threads)  # noqa: E731
    run()  # warm-up
    t0 = time.perf_counter()
    calls = 0
    while True:
        run()
        calls += 1
        el = time.perf_counter() - t0
        if el >= seconds or calls >= 200:
            break
    qps = calls * nq / el
    return {"value": round(qps, 2), "unit": "queries/s", "cores": threads, "kind": kind,
            "sample": f"{calls} calls x {nq} queries against all {base.shape[0]:,} x {base.shape[1]} "
                      f"rows, k={k}, {el:.1f} s ({'BruteForceIndex::search_batch from the reference sources' if ref else 'C oracle'})"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(the exact scan would shard rows across GPUs and merge top-k lists; see DESIGN.md).",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import vecsearch_amd as vs
    from vecsearch_amd import datagen

    assert vs.hip_available(), "no GPU visible"
    n, d, k, B = args.n, args.dim, args.k, args.batch
    nqt = max(args.queries, B)
    log(f"data: {n:,} x {d} mixture base, {nqt:,} queries")
    t = time.time()
    base = datagen.mixture(n, d, seed=0)
    queries = datagen.mixture(nqt, d, seed=1)
    log(f"  generated in {time.time() - t:.1f} s")

    idx = vs.GpuBruteForceIndex(d, "cosine")
    t = time.time()
    idx.add(base)
    log(f"  uploaded in {time.time() - t:.1f} s")

    dev = torch.device("cuda:0")
    dq = torch.from_numpy(queries).to(dev)
    maxb = max([B] + [int(b) for b in args.sweep.split(",") if b])
    ids = torch.empty((maxb, k), dtype=torch.int32, device=dev)
    dists = torch.empty((maxb, k), dtype=torch.float32, device=dev)
    # a dedicated (default-priority) stream: kernels and timing events share it
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.set_stream(stream)
    sptr = stream.cuda_stream
    assert sptr != 0

    def step(i, b):
        off = (i * b) % (nqt - b + 1)
        idx.search_device(dq.data_ptr() + off * d * 4, b, k, ids.data_ptr(), dists.data_ptr(),
                          sptr, args.method)

    # ---- timed steps at the default batch ----
    for i in range(args.warmup):
        step(i, B)
    torch.cuda.synchronize()
    idx.kernel_timing(True)
    idx.kernel_times()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for i in range(args.steps):
        step(args.warmup + i, B)
    e1.record(stream)
    torch.cuda.synchronize()
    ms_total = e0.elapsed_time(e1)
    kt = idx.kernel_times()
    idx.kernel_timing(False)
    ms_per_step = ms_total / args.steps
    value = B * args.steps / (ms_total / 1e3)
    log(f"batch {B}: {ms_per_step:.3f} ms/step, {value:,.0f} QPS; kernels {kt}")

    qb = 1 if B == 1 or args.method == "naive" else 2 if B == 2 else 4 if B <= 4 else 8
    scan_ms = kt["scan_ms"] / max(1, kt["scan_launches"])  # (exact path; shortlist: see kernels)
    alg_bytes = n * d * 4 + qb * d * 4 + qb * k * 8
    alg_flops = 2.0 * n * d * qb
    t_hbm, t_fp = alg_bytes / (HBM_PEAK_GBS * 1e9), alg_flops / (FP32_PEAK_TFLOPS * 1e12)
    if t_hbm >= t_fp:
        achieved = alg_bytes / (scan_ms / 1e3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4)}
    else:
        achieved = alg_flops / (scan_ms / 1e3) / 1e12
        roof = {"bound": "mfma", "achieved": round(achieved, 2), "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4)}
    roof["traffic"] = traffic_from_profiles(f"scan_exact_{n // 1000}kx{d}_b{qb}_k{k}_traffic.md")
    roof["kernel"] = "vsg::scan_exact_kernel"
    roof["kernel_ms"] = round(scan_ms, 4)
    roof["alg_bytes_per_launch"] = alg_bytes
    roof["merge_ms"] = round(kt["merge_ms"] / max(1, kt["merge_launches"]), 4)

    # ---- sweep: device-resident QPS and host (PCIe-inclusive) latency per batch ----
    sweep = {}
    for b in [int(x) for x in args.sweep.split(",") if x]:
        if b > nqt:
            continue
        step(0, b)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        step(1, b)
        torch.cuda.synchronize()
        one = time.perf_counter() - t0
        reps = int(min(200, max(3, args.sweep_seconds / 2 / max(one, 1e-6))))
        e0.record(stream)
        for r in range(reps):
            step(r, b)
        e1.record(stream)
        torch.cuda.synchronize()
        dev_qps = b * reps / (e0.elapsed_time(e1) / 1e3)
        lat = []
        qh = queries
        t0 = time.perf_counter()
        while len(lat) < 3 or sum(lat) < args.sweep_seconds:
            off = (len(lat) * b) % (nqt - b + 1)
            t1 = time.perf_counter()
            idx.search(qh[off:off + b], k)
            lat.append(time.perf_counter() - t1)
        lat = np.array(lat)
        sweep[str(b)] = {"qps_device": round(dev_qps, 1), "qps_host": round(b / float(np.median(lat)), 1),
                         "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
                         "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3), "calls": len(lat)}
        log(f"  sweep batch {b}: {sweep[str(b)]}")

    cpu = None
    if args.cpu_seconds > 0:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"cpu baseline: {threads} threads, ~{args.cpu_seconds:.0f} s")
        cpu = cpu_baseline(base, queries, k, args.cpu_seconds, threads)
        log(f"  {cpu}")

    out = {
        "metric": "exact top-10 QPS, 1M x 384 cosine (plain FP32 exact path)",
        "value": round(value, 2),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded 1,000-cluster Gaussian mixture, unit rows; queries held out)",
        "config": {"workload": f"exact cosine top-{k}, {n:,} x {d} fp32 base resident in HBM, batch {B} "
                               f"device-resident queries", "n": n, "dim": d, "k": k, "batch": B,
                   "queries": nqt, "method": args.method, "parallelism": "1 GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "sweep": sweep,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
