"""QPS by batch size for each search method, as CSV rows in the reference's result format.

    python -m vecsearch_amd.gpu_bench --out results/gpu/mi355x.csv [--methods auto,skinny,bf16,fp8]

The file has the ten columns of the reference's results/gpu/*.csv (bench/gpu_bench.py:121):
backend, device, n, dim, metric, k, batch, ms_per_batch, qps, recall_vs_cpu. A backend is
"hip-<method>" ("hip" for method auto). Two rules are kept so the rows compare with the
reference's: the data are i.i.d. Gaussian rows, L2-normalised (datagen.unit_vectors), and a
cell's time is the median over repeated host calls (queries in from host memory, results out
to it) after one untimed call, repeated until there are at least three and `--min-time`
seconds of them (bench/gpu_bench.py:75-85).

recall_vs_cpu compares each method's ids with the exact streaming scan ("naive"), whose ids
and distances equal the reference CPU index's bit for bit; a CPU timing baseline is bench.py's.
"""
import argparse
import csv
import os
import statistics
import subprocess
import time
from dataclasses import asdict, dataclass

from . import GpuBruteForceIndex, datagen, recall_at_k

METHODS = ("auto", "naive", "skinny", "tiled", "bf16", "fp8")


@dataclass
class Row:
    backend: str
    device: str
    n: int
    dim: int
    metric: str
    k: int
    batch: int
    ms_per_batch: float
    qps: float
    recall_vs_cpu: float


COLUMNS = list(Row.__dataclass_fields__)


def device_label():
    """The GPU's marketing name from rocm-smi, else the architecture."""
    try:
        text = subprocess.run(["rocm-smi", "--showproductname"], capture_output=True, text=True,
                              timeout=20).stdout
    except (OSError, subprocess.SubprocessError):
        text = ""
    names = [ln.split(":")[-1].strip() for ln in text.splitlines() if "card series" in ln.lower()]
    return names[0] if names else "AMD Instinct (gfx950)"


def median_call_seconds(call, min_time):
    """One untimed call, then timed calls until there are >= 3 and they sum to >= min_time
    (a single call longer than min_time ends the cell after three)."""
    call()
    spent = []
    while len(spent) < 3 or (sum(spent) < min_time and spent[0] <= min_time):
        t0 = time.perf_counter()
        call()
        spent.append(time.perf_counter() - t0)
    return statistics.median(spent)


def sweep(index, queries, k, methods, batches, caps, truth_rows, min_time, device):
    """Yields one Row per (method, batch) cell; truth_rows are the exact ids of the first
    len(truth_rows) queries."""
    n, dim = len(index), index.dim
    check = queries[:len(truth_rows)]
    for method in methods:
        found = index.search(check, k, method)[0]
        recall = recall_at_k(found, truth_rows, k)
        for b in batches:
            if b > caps.get(method, b):
                continue
            block = queries[:b]
            sec = median_call_seconds(lambda: index.search(block, k, method), min_time)
            yield Row("hip" if method == "auto" else f"hip-{method}", device, n, dim, index.metric, k, b,
                      round(sec * 1e3, 3), round(b / sec, 1), round(recall, 4))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", required=True)
    ap.add_argument("--methods", default="auto,skinny,bf16,fp8", help=f"comma list of {', '.join(METHODS)}")
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--dim", type=int, default=384)
    ap.add_argument("--metric", choices=("ip", "l2"), default="ip")
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--batches", default="1,8,32,128,512,1024,4096")
    ap.add_argument("--caps", default="naive=1024", help="method=largest batch, comma separated")
    ap.add_argument("--min-time", type=float, default=1.0)
    ap.add_argument("--check", type=int, default=128, help="queries whose recall is measured")
    args = ap.parse_args(argv)

    methods = [m for m in args.methods.split(",") if m]
    bad = [m for m in methods if m not in METHODS]
    if bad:
        ap.error(f"unknown methods {bad}")
    batches = sorted(int(b) for b in args.batches.split(","))
    caps = dict((m, int(b)) for m, b in (kv.split("=") for kv in args.caps.split(",") if kv))

    base = datagen.unit_vectors(args.n, args.dim, seed=0)
    queries = datagen.unit_vectors(max(batches[-1], args.check), args.dim, seed=1)
    index = GpuBruteForceIndex(args.dim, args.metric)
    index.add(base)
    del base
    truth = index.search(queries[:args.check], args.k, "naive")[0]
    device = device_label()
    print(f"{args.n:,} x {args.dim} {args.metric}, k={args.k}, on {device}", flush=True)

    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w", newline="") as f:
        out = csv.DictWriter(f, fieldnames=COLUMNS)
        out.writeheader()
        for row in sweep(index, queries, args.k, methods, batches, caps, truth, args.min_time, device):
            out.writerow(asdict(row))
            f.flush()
            print(f"{row.backend:>10} batch {row.batch:>5}: {row.ms_per_batch:9.3f} ms {row.qps:12,.0f} QPS "
                  f"recall {row.recall_vs_cpu}", flush=True)
    print(f"wrote {args.out}")


if __name__ == "__main__":
    main()
