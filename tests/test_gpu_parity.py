"""HIP engine vs the CPU oracle: ids identical and distances bit-identical.

Mirrors the reference's GPU tests (tests/python/test_gpu.py:17-70) with the bar raised from
"allclose, ids > 99 %" to exact equality, plus the filtered paths (test_bindings.py:76-90,
tests/cpp/test_filter.cpp:87-121) and the mixture workload of the benchmark at 100k rows.
"""
import numpy as np
import pytest

import oracle
import vecsearch_amd as vs
from vecsearch_amd import datagen as dg

pytestmark = pytest.mark.gpu
METHODS = ["naive", "skinny", "tiled", "auto"]


def data(n, d, seed, normalize=False):
    x = np.random.default_rng(seed).standard_normal((n, d), dtype=np.float32)
    return vs.normalize(x) if normalize else x


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def as_i64(ids):
    return np.where(ids == oracle.NONE, -1, ids.astype(np.int64))


def assert_same(got, exp):
    gi, gd = got
    ei, ed = exp
    assert gi.shape == ei.shape
    np.testing.assert_array_equal(gi, as_i64(ei))
    np.testing.assert_array_equal(bits(gd), bits(ed))


@pytest.mark.parametrize("metric", ["l2", "ip"])
@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("d", [37, 128])
@pytest.mark.parametrize("nq", [1, 5, 300])
def test_matches_oracle(metric, method, d, nq):
    x, q, k = data(3000, d, 1, metric == "ip"), data(nq, d, 2, metric == "ip"), 10
    gpu = vs.GpuBruteForceIndex(d, metric)
    gpu.add(x[:1000])  # two adds: the device buffer grows and keeps the first part
    gpu.add(x[1000:])
    assert len(gpu) == 3000 and gpu.dim == d
    got = gpu.search(q, k, method=method)
    assert_same(got, oracle.search(x, q, k, metric, threads=8))
    assert np.all(np.diff(got[1], axis=1) >= 0)


@pytest.mark.parametrize("d", [1, 3, 7, 8, 9, 15, 16, 17, 24, 31, 45, 384, 385, 768])
@pytest.mark.parametrize("metric", ["l2", "ip"])
def test_every_tail_length(d, metric):
    x, q = data(700, d, 3, metric == "ip"), data(9, d, 4, metric == "ip")
    gpu = vs.GpuBruteForceIndex(d, metric)
    gpu.add(x)
    assert_same(gpu.search(q, 12), oracle.search(x, q, 12, metric, threads=8))


@pytest.mark.parametrize("k", [1, 16, 17, 64, 65, 100, 128])
def test_k_buckets(k):
    x, q = data(20000, 64, 3), data(40, 64, 4)
    gpu = vs.GpuBruteForceIndex(64)
    gpu.add(x)
    assert_same(gpu.search(q, k), oracle.search(x, q, k, "l2", threads=8))


def test_many_queries_and_rounds():
    x, q = data(70000, 16, 5), data(5000, 16, 6)
    gpu = vs.GpuBruteForceIndex(16)
    gpu.add(x)
    assert_same(gpu.search(q, 10), oracle.search(x, q, 10, "l2", threads=8))


def test_exact_ties_go_to_lower_id():
    x = data(4000, 32, 7)
    x[100:140] = x[5]  # 41 identical rows
    gpu = vs.GpuBruteForceIndex(32)
    gpu.add(x)
    q = x[5:6] + 0.01
    ids, dists = gpu.search(q, 50)
    assert_same((ids, dists), oracle.search(x, q, 50, "l2"))
    assert list(ids[0, :41]) == [5] + list(range(100, 140))


def test_missing_results_and_limits():
    gpu = vs.GpuBruteForceIndex(8)
    ids, dists = gpu.search(np.zeros((2, 8), np.float32), 5)  # empty index
    assert (ids == -1).all() and np.isinf(dists).all()
    gpu.add(data(3, 8, 7))
    ids, dists = gpu.search(np.zeros(8, np.float32), 5)
    assert (ids[0, :3] >= 0).all() and (ids[0, 3:] == -1).all() and np.isinf(dists[0, 3:]).all()
    with pytest.raises(ValueError):
        gpu.search(np.zeros(8, np.float32), 129)
    with pytest.raises(ValueError):
        gpu.search(np.zeros(8, np.float32), 5, method="fast")
    with pytest.raises(ValueError):
        gpu.search(np.zeros((2, 9), np.float32), 5)
    ids, dists = gpu.search(np.zeros((0, 8), np.float32), 5)
    assert ids.shape == (0, 5)


def test_forcecast_float64():
    x, q = data(500, 24, 8), data(4, 24, 9)
    gpu = vs.GpuBruteForceIndex(24, "cosine")
    gpu.add(x.astype(np.float64))
    assert_same(gpu.search(q.astype(np.float64), 7), oracle.search(x, q, 7, "ip"))


@pytest.mark.parametrize("frac", [0.5, 0.3, 0.01, 0.0005, 0.0])
@pytest.mark.parametrize("metric", ["l2", "ip"])
def test_mask_filter(frac, metric):
    x, q = data(30000, 64, 10, metric == "ip"), data(17, 64, 11, metric == "ip")
    mask = np.random.default_rng(12).random(len(x)) < frac
    gpu = vs.GpuBruteForceIndex(64, metric)
    gpu.add(x)
    got = gpu.search(q, 10, filter=mask)
    assert_same(got, oracle.search(x, q, 10, metric, threads=8, mask=mask))
    ids = got[0]
    assert mask[ids[ids >= 0]].all()
    # test_bindings.py:76-90: equal to brute force over x[mask] mapped through flatnonzero
    if mask.any():
        sub = vs.GpuBruteForceIndex(64, metric)
        sub.add(x[mask])
        gi, gd = sub.search(q, 10)
        mapped = np.where(gi >= 0, np.flatnonzero(mask)[np.maximum(gi, 0)], -1)
        np.testing.assert_array_equal(ids, mapped)
    with pytest.raises(ValueError):
        gpu.search(q, 10, filter=np.ones(5, bool))


@pytest.mark.parametrize("frac", [0.9, 0.5, 0.3])
@pytest.mark.parametrize("metric", ["l2", "ip"])
def test_mask_filter_large_batch(frac, metric):
    # a dense filter over 256 queries and more runs the filtered scan with one shared id list
    # (the tile autotuner's first call included): the oracle's bits, as a small batch gets them
    x, q = data(70000, 64, 30, metric == "ip"), data(300, 64, 31, metric == "ip")
    mask = np.random.default_rng(32).random(len(x)) < frac
    gpu = vs.GpuBruteForceIndex(64, metric)
    gpu.add(x)
    exp = oracle.search(x, q, 10, metric, threads=8, mask=mask)
    for _ in range(2):
        assert_same(gpu.search(q, 10, filter=mask), exp)
    small = gpu.search(q[:8], 10, filter=mask)
    np.testing.assert_array_equal(small[0], exp[0][:8])


@pytest.mark.parametrize("metric", ["l2", "ip", "cosine"])
@pytest.mark.parametrize("nq", [24, 40, 100])
def test_exact_batch_on_filtered_scan(metric, nq):
    # 24 queries and more of the exact methods run as the filtered scan over every row
    # (block-shared tile below 64 queries, one lane per row above): the oracle's bits
    x, q = data(70000, 64, 40, metric != "l2"), data(nq, 64, 41, metric != "l2")
    gpu = vs.GpuBruteForceIndex(64, metric)
    gpu.add(x)
    exp = oracle.search(x, q, 10, "l2" if metric == "l2" else "ip", threads=8)
    assert_same(gpu.search(q, 10, method="skinny"), exp)
    assert_same(gpu.search(q, 10, method="naive"), exp)


def test_filter_fewer_allowed_than_k():
    x, q = data(5000, 32, 13), data(3, 32, 14)
    mask = np.zeros(len(x), bool)
    mask[[7, 4000]] = True
    gpu = vs.GpuBruteForceIndex(32)
    gpu.add(x)
    ids, dists = gpu.search(q, 10, filter=mask)
    assert set(ids[0, :2]) == {7, 4000} and (ids[:, 2:] == -1).all() and np.isinf(dists[:, 2:]).all()
    assert_same((ids, dists), oracle.search(x, q, 10, "l2", mask=mask))


@pytest.mark.parametrize("metric", ["l2", "ip"])
def test_search_batch_lists(metric):
    n, d = 40000, 96
    x = dg.mixture(n, d, seed=20)
    groups = dg.zipf_groups(n, 50, 1.1, seed=21)
    off, ids_g = dg.group_csr(groups, 50)
    q = dg.mixture(300, d, seed=22)
    ks, qg = dg.query_filters(300, 50, seed=23, ks=(1, 10, 100))
    ranges = dg.query_ranges(qg, off)
    gpu = vs.GpuBruteForceIndex(d, metric)
    gpu.add(x)
    got = gpu.search_batch(q, ks, ranges, ids_g)
    exp = oracle.search_lists(x, q, ks, ranges, ids_g, metric, threads=8, kmax=100)
    assert_same(got, exp)


@pytest.mark.parametrize("metric", ["l2", "ip"])
def test_search_batch_mixed_unfiltered(metric):
    # unfiltered queries (ALL_ROWS) run in the same persistent launch as filtered ones
    n, d = 30000, 64
    x = dg.mixture(n, d, seed=24)
    groups = dg.zipf_groups(n, 20, 1.1, seed=25)
    off, ids_g = dg.group_csr(groups, 20)
    q = dg.mixture(90, d, seed=26)
    ks, qg = dg.query_filters(90, 20, seed=27, ks=(1, 10, 100))
    ranges = dg.query_ranges(qg, off)
    ranges[::3] = [vs.ALL_ROWS, 0]
    gpu = vs.GpuBruteForceIndex(d, metric)
    gpu.add(x)
    got = gpu.search_batch(q, ks, ranges, ids_g)
    assert_same(got, oracle.search_lists(x, q, ks, ranges, ids_g, metric, threads=8, kmax=100))
    # and an unfiltered query equals the plain search
    plain = gpu.search(q[0:1], int(ks[0]), method="skinny")
    np.testing.assert_array_equal(got[0][0, :ks[0]], plain[0][0])


@pytest.mark.parametrize("d", [128, 384, 768])
@pytest.mark.parametrize("qb,r", [(8, 4), (8, 2), (16, 2), (32, 8)])
@pytest.mark.parametrize("metric", ["l2", "ip"])
def test_search_batch_eight_lanes_per_row(d, qb, r, metric):
    # the eight-lanes-per-row tree (row_half_mirror, xor 2, xor 1) gives the reference's bits;
    # qb = 32: one item per block, rows and query dims streamed through the LDS ring
    n = 12000
    x = dg.mixture(n, d, seed=34)
    groups = dg.zipf_groups(n, 8, 1.1, seed=35)
    off, ids_g = dg.group_csr(groups, 8)
    q = dg.mixture(70, d, seed=36)
    ks, qg = dg.query_filters(70, 8, seed=37, ks=(1, 10, 100))
    ranges = dg.query_ranges(qg, off)
    ranges[::5] = [vs.ALL_ROWS, 0]
    gpu = vs.GpuBruteForceIndex(d, metric)
    gpu.add(x)
    for name, v in (("lists_qb", qb), ("lists_r", r), ("lists_chunk", 512), ("lists_lanes_per_row", 8)):
        gpu.set_param(name, v)
    got = gpu.search_batch(q, ks, ranges, ids_g)
    assert gpu.get_param("lists_last_lanes_per_row") == 8
    assert_same(got, oracle.search_lists(x, q, ks, ranges, ids_g, metric, threads=8, kmax=100))


@pytest.mark.parametrize("d", [64, 128, 384, 768])
@pytest.mark.parametrize("kset", [(1, 10, 100), (5, 64)])
@pytest.mark.parametrize("metric", ["l2", "ip"])
def test_search_batch_one_lane_per_row(d, kset, metric):
    # one lane per row, queries as scalar operands, eight waves per item, lists merged by rank;
    # 70 queries over 8 groups give items of 8 and of fewer queries; the 2,048-row chunks send
    # whole 64-key steps through the sort-and-merge insert; 300 duplicated rows tie exactly
    n = 14000
    x = dg.mixture(n, d, seed=38)
    x[7000:7300] = x[100]
    groups = dg.zipf_groups(n, 8, 1.1, seed=39)
    off, ids_g = dg.group_csr(groups, 8)
    q = dg.mixture(70, d, seed=40)
    q[3] = x[100]
    ks, qg = dg.query_filters(70, 8, seed=41, ks=kset)
    ranges = dg.query_ranges(qg, off)
    ranges[::7] = [vs.ALL_ROWS, 0]
    gpu = vs.GpuBruteForceIndex(d, metric)
    gpu.add(x)
    for chunk in (2048, 320):
        gpu.set_param("lists_lanes_per_row", 1)
        gpu.set_param("lists_chunk", chunk)
        got = gpu.search_batch(q, ks, ranges, ids_g)
        assert gpu.get_param("lists_last_lanes_per_row") == 1
        assert_same(got, oracle.search_lists(x, q, ks, ranges, ids_g, metric, threads=8, kmax=int(max(kset))))


def test_search_batch_autotuner_keeps_results():
    # the first large call of a shape class times every candidate tile on its own inputs; the
    # class then reuses the fastest; results are the same bits whichever tile runs
    n, d = 60000, 64
    x = dg.mixture(n, d, seed=28)
    q = dg.mixture(300, d, seed=29)
    ks = np.full(300, 10, np.uint32)
    ranges = np.zeros((300, 2), np.uint64)
    ranges[:, 0] = vs.ALL_ROWS
    allow = np.zeros(1, np.uint32)
    gpu = vs.GpuBruteForceIndex(d, "ip")
    gpu.add(x)
    assert gpu.get_param("lists_tuned_classes") == 0
    first = gpu.search_batch(q, ks, ranges, allow)
    assert gpu.get_param("lists_tuned_classes") == 1
    assert_same(first, oracle.search_lists(x, q, ks, ranges, allow, "ip", threads=8, kmax=10))
    second = gpu.search_batch(q, ks, ranges, allow)
    assert_same(second, (np.where(first[0] < 0, oracle.NONE, first[0]).astype(np.uint32), first[1]))
    assert gpu.get_param("lists_last_qb") in (1, 2, 4, 8, 16, 32)
    gpu.set_param("lists_autotune", 0)
    assert gpu.get_param("lists_tuned_classes") == 0
    gpu.set_param("lists_qb", 4)
    gpu.set_param("lists_r", 4)
    third = gpu.search_batch(q, ks, ranges, allow)
    assert gpu.get_param("lists_last_qb") == 4 and gpu.get_param("lists_last_r") == 4
    np.testing.assert_array_equal(third[0], first[0])


def test_search_batch_edge_cases():
    x = data(3000, 20, 30)
    gpu = vs.GpuBruteForceIndex(20)
    gpu.add(x)
    q = data(5, 20, 31)
    allow = np.array([3, 9, 10, 2999], np.uint32)
    ranges = np.array([[0, 4], [0, 0], [1, 3], [0, 4], [3, 4]], np.uint64)
    ks = np.array([10, 5, 1, 0, 2], np.uint32)
    got = gpu.search_batch(q, ks, ranges, allow)
    assert_same(got, oracle.search_lists(x, q, ks, ranges, allow, "l2", kmax=10))
    assert (got[0][1] == -1).all() and (got[0][3] == -1).all()
    with pytest.raises(ValueError):  # ids must be increasing
        gpu.search_batch(q[:1], np.array([3], np.uint32), np.array([[0, 2]], np.uint64),
                         np.array([9, 3], np.uint32))
    with pytest.raises(ValueError):  # id out of range
        gpu.search_batch(q[:1], np.array([3], np.uint32), np.array([[0, 1]], np.uint64),
                         np.array([3000], np.uint32))
    with pytest.raises(ValueError):  # range past the id array
        gpu.search_batch(q[:1], np.array([3], np.uint32), np.array([[0, 5]], np.uint64), allow)


@pytest.mark.parametrize("batch", [1, 8, 100])
def test_mixture_100k_384(batch):
    # the CPU-runnable benchmark shape: 100,000 x 384 unit rows, cosine, top-10
    x = dg.mixture(100_000, 384, seed=0)
    q = dg.mixture(batch, 384, seed=1)
    gpu = vs.GpuBruteForceIndex(384, "cosine")
    gpu.add(x)
    assert_same(gpu.search(q, 10), oracle.search(x, q, 10, "ip", threads=8))


# ---- certified shortlist (large batches): must equal the exact results bit for bit ---------

@pytest.fixture(scope="module")
def mixture_70k():
    x = dg.mixture(70_000, 384, seed=40)
    q = dg.mixture(600, 384, seed=41)  # above 512 queries: the rows-stationary pass 2
    gpu = vs.GpuBruteForceIndex(384, "cosine")
    gpu.add(x)
    return x, q, gpu


@pytest.mark.parametrize("method", ["tiled", "bf16", "fp8", "auto"])
@pytest.mark.parametrize("k", [1, 10, 100, 128])
def test_shortlist_matches_oracle(mixture_70k, method, k):
    x, q, gpu = mixture_70k
    exp = oracle.search(x, q, k, "ip", threads=8)
    got = gpu.search(q, k, method=method)
    assert_same(got, exp)


@pytest.mark.parametrize("method", ["bf16", "fp8"])
@pytest.mark.parametrize("nq", [1, 4, 16, 64, 100, 128])
@pytest.mark.parametrize("k", [10, 100])
def test_shortlist_small_batches(mixture_70k, method, nq, k):
    # up to 128 queries (d <= 384) take the queries-stationary pass 2: same bits as the oracle,
    # and as the same queries inside a larger batch (rows-stationary pass 2); up to 128 queries
    # also rescore each query in several blocks (128, 100, 64, 16, 4: 2, 2, 4, 16, 32 shares)
    # and merge them
    x, q, gpu = mixture_70k
    got = gpu.search(q[:nq], k, method=method)
    assert_same(got, oracle.search(x, q[:nq], k, "ip", threads=8))
    big = gpu.search(q, k, method=method)
    np.testing.assert_array_equal(got[0], big[0][:nq])
    np.testing.assert_array_equal(bits(got[1]), bits(big[1][:nq]))


@pytest.mark.parametrize("method", ["bf16", "fp8"])
@pytest.mark.parametrize("nq", [384, 512])
def test_shortlist_three_and_four_query_groups(mixture_70k, method, nq):
    # 3 and 4 query groups of the queries-stationary pass 2 (FP8: each stage's row scales in
    # its ring slot): the oracle's bits, and the same as inside the 600-query rows-stationary batch
    x, q, gpu = mixture_70k
    got = gpu.search(q[:nq], 10, method=method)
    assert_same(got, oracle.search(x, q[:nq], 10, "ip", threads=8))
    big = gpu.search(q, 10, method=method)
    np.testing.assert_array_equal(got[0], big[0][:nq])
    np.testing.assert_array_equal(bits(got[1]), bits(big[1][:nq]))


@pytest.mark.parametrize("n,d,nq", [(5003, 128, 37), (1000, 64, 5), (70_001, 768, 64), (70_001, 768, 65)])
def test_shortlist_small_batch_shapes(n, d, nq):
    # ragged row counts (a partial last fragment, blocks of a few rows) and d = 768, where 64
    # queries are one query group and 65 are not
    x = dg.mixture(n, d, seed=90)
    q = dg.mixture(nq, d, seed=91)
    gpu = vs.GpuBruteForceIndex(d, "cosine")
    gpu.add(x)
    for method in ("bf16", "fp8"):
        assert_same(gpu.search(q, 10, method=method), oracle.search(x, q, 10, "ip", threads=8))


@pytest.mark.parametrize("n", [5, 65_537, 65_551])
@pytest.mark.parametrize("nq", [3, 129, 300])
def test_shortlist_ragged_and_tiny_bases(n, nq):
    # the shortlist serves bases of 65,536 rows and up: one row past a whole fragment, 15 past
    # (a partial last fragment), with one and several query groups of the queries-stationary
    # pass 2 and the rows-stationary one (300 queries); a 5-row base takes the exact path, and
    # k > n pads with -1 / inf
    d = 128
    x = dg.mixture(n, d, seed=92)
    q = dg.mixture(nq, d, seed=93)
    gpu = vs.GpuBruteForceIndex(d, "cosine")
    gpu.add(x)
    for method in ("bf16", "fp8"):
        for k in (1, 10):
            assert_same(gpu.search(q, k, method=method), oracle.search(x, q, k, "ip", threads=8))


@pytest.mark.parametrize("d", [32, 64, 100, 128, 256, 768])
def test_shortlist_other_dims(d):
    # widths of 64 and below take FP8's 32-dim MFMA instruction, the others its 128-dim one
    x = dg.mixture(66_000, d, seed=42)
    q = dg.mixture(70, d, seed=43)
    gpu = vs.GpuBruteForceIndex(d, "ip")
    gpu.add(x)
    for method in ("bf16", "fp8"):
        assert_same(gpu.search(q, 10, method=method), oracle.search(x, q, 10, "ip", threads=8))


@pytest.mark.parametrize("stride", [1, 7, 64])
def test_shortlist_sample_stride(stride):
    # any first-pass tile stride gives a valid (looser or tighter) threshold: the same results
    x = dg.mixture(80_000, 384, seed=52)
    q = dg.mixture(96, 384, seed=53)
    gpu = vs.GpuBruteForceIndex(384, "cosine")
    gpu.add(x)
    gpu.set_param("short_sample_stride", stride)
    assert gpu.get_param("short_sample_stride") == stride
    exp = oracle.search(x, q, 10, "ip", threads=8)
    for method in ("bf16", "fp8"):
        assert_same(gpu.search(q, 10, method=method), exp)
    with pytest.raises(ValueError):
        gpu.set_param("short_sample_stride", 1000)


@pytest.mark.parametrize("scale", [1e-3, 40.0])
def test_shortlist_unnormalized_ip(scale):
    # inner products far from 1 in magnitude: the thresholds and score cuts must stay
    # conservative where a float ulp is much larger (or smaller) than the fixed cushions
    rng = np.random.default_rng(48)
    x = (dg.mixture(66_000, 128, seed=48) * rng.uniform(0.5, 2.0, (66_000, 1))).astype(np.float32) * scale
    q = (dg.mixture(80, 128, seed=49) * scale).astype(np.float32)
    gpu = vs.GpuBruteForceIndex(128, "ip")
    gpu.add(x)
    for method in ("bf16", "fp8"):
        assert_same(gpu.search(q, 10, method=method), oracle.search(x, q, 10, "ip", threads=8))


@pytest.mark.parametrize("dups", [3000, 9000])
def test_shortlist_many_ties_and_overflow(dups):
    # `dups` identical rows tie with each other. 3,000 fit the candidate buffer (8,192) and are
    # rescored and ranked in LDS; 9,000 overflow it, and the queries near them are answered by
    # the exact scan. Either way ties go to the lower id.
    x = dg.mixture(70_000, 128, seed=44)
    x[10_000:10_000 + dups] = x[5]
    near = x[5:6] * 0.999 + 0.001 * dg.mixture(4, 128, seed=46)
    q = np.concatenate([near[:1], dg.mixture(40, 128, seed=45), near[1:],
                        dg.mixture(20, 128, seed=47)]).astype(np.float32)
    q = vs.normalize(q)
    gpu = vs.GpuBruteForceIndex(128, "cosine")
    gpu.add(x)
    got = gpu.search(q, 50, method="bf16")
    assert_same(got, oracle.search(x, q, 50, "ip", threads=8))
    fallbacks = gpu.short_stats()["fallbacks"]
    assert fallbacks >= 4 if dups > 8192 else fallbacks == 0
    assert list(got[0][0, :50]) == [5] + list(range(10_000, 10_049))


def test_shortlist_grows_with_adds(mixture_70k):
    # rows added after the low-precision copies exist are converted before the next search
    x, q, _ = mixture_70k
    gpu = vs.GpuBruteForceIndex(384, "cosine")
    gpu.add(x[:66_000])
    assert_same(gpu.search(q[:64], 10, method="bf16"), oracle.search(x[:66_000], q[:64], 10, "ip", threads=8))
    gpu.add(x[66_000:])
    assert_same(gpu.search(q[:64], 10, method="bf16"), oracle.search(x, q[:64], 10, "ip", threads=8))


@pytest.mark.parametrize("method", ["bf16", "fp8", "auto"])
@pytest.mark.parametrize("d,nq,k", [(64, 64, 10), (384, 300, 100), (384, 700, 10), (128, 40, 1), (768, 130, 37)])
def test_l2_shortlist(method, d, nq, k):
    # the L2 shortlist: estimates carry each row's |x|^2 / 2; rows of different norms (scaled by
    # 0.5 .. 2) so that term matters; ids and distances are the exact scan's, and the shortlist
    # (not only its exact fallback) answers some queries
    x = dg.mixture(66_000, d, seed=46)
    x *= np.exp(np.random.default_rng(48).uniform(np.log(0.5), np.log(2.0), len(x)))[:, None].astype(np.float32)
    q = dg.mixture(nq, d, seed=47)
    gpu = vs.GpuBruteForceIndex(d, "l2")
    gpu.add(x)
    got = gpu.search(q, k, method)
    assert_same(got, oracle.search(x, q, k, "l2", threads=8))
    if method != "fp8":  # FP8's wider bound may send every query of a case to the exact scan
        assert gpu.short_stats()["fallbacks"] < nq
