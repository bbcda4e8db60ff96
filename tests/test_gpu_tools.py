"""The reference-format tools on the engine: the QPS sweep writes bench/gpu_bench.py's CSV
columns, and ground_truth writes gt.npy with the oracle's ids."""
import csv

import numpy as np
import pytest

import oracle
from vecsearch_amd import datagen, gpu_bench, ground_truth

pytestmark = pytest.mark.gpu


def test_gpu_bench_csv(tmp_path):
    out = tmp_path / "sweep.csv"
    gpu_bench.main(["--n", "30000", "--dim", "64", "--batches", "1,8,64", "--methods", "auto,naive,skinny",
                    "--min-time", "0.05", "--check", "32", "--out", str(out)])
    rows = list(csv.DictReader(open(out)))
    # the columns of the reference's results/gpu/*.csv (bench/gpu_bench.py:121)
    assert list(rows[0].keys()) == ["backend", "device", "n", "dim", "metric", "k", "batch", "ms_per_batch",
                                    "qps", "recall_vs_cpu"]
    assert len(rows) == 9 and {r["backend"] for r in rows} == {"hip", "hip-naive", "hip-skinny"}
    assert all(float(r["recall_vs_cpu"]) == 1.0 and float(r["qps"]) > 0 for r in rows)


def test_median_call_seconds_rule():
    """One untimed call, then at least three timed calls and at least min_time seconds."""
    calls = []
    gpu_bench.median_call_seconds(lambda: calls.append(1), 0.0)
    assert len(calls) == 4


def test_ground_truth_file(tmp_path):
    xb = datagen.mixture(20000, 32, seed=90)
    xq = datagen.mixture(50, 32, seed=91)
    np.save(tmp_path / "base.npy", xb)
    np.save(tmp_path / "queries.npy", xq)
    ground_truth.main(["--data", str(tmp_path), "--k", "20"])
    gt = np.load(tmp_path / "gt.npy")
    assert gt.dtype == np.int64 and gt.shape == (50, 20)
    np.testing.assert_array_equal(gt, oracle.search(xb, xq, 20, "ip", threads=8)[0].astype(np.int64))
