"""Every benchmark workload at its full size, through the HIP engine, against the oracle.

- P2: 1,000,000 x 384 cosine (1,000-cluster mixture), top-10, every method at batches 1 .. 4,096.
- Q: the same base through the BF16 and FP8 certified shortlists at batches 128 / 1,024 / 4,096,
  k = 10 and 100.
- F: 892,800 x 384, rows in 100 groups (Zipf exponent 1.1, 12 % .. 0.01 % of the rows), 4,096
  queries each filtered to one group with k in {10, 100}, one batched call.
- S: 2,000,000 x 768 through the serving loop: a mix of requests of 1 .. 512 queries, k in
  {10, 100}, half filtered, each equal alone and in the mix.

Query rows are checked against the oracle (oracle.search / oracle.search_lists: the reference's
CPU arithmetic, bit for bit) on a sample that covers every batch; whole batches are checked
against each other, since every method returns the same bits. The reference's own tests do
the same at toy sizes (tests/python/test_gpu.py:17-33, test_bindings.py:76-90).
"""
import os

import numpy as np
import pytest

import oracle
import vecsearch_amd as vs
from vecsearch_amd import datagen as dg

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)
BATCHES = [1, 8, 32, 128, 512, 1024, 4096]
# sampled query rows: all of the first 16, then spread over each larger batch
SAMPLE = np.unique(np.concatenate([np.arange(16), np.linspace(16, 127, 16).astype(int),
                                   np.linspace(128, 1023, 16).astype(int), np.linspace(1024, 4095, 16).astype(int)]))


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def as_i64(ids):
    return np.where(ids == oracle.NONE, -1, ids.astype(np.int64))


def check_rows(got_ids, got_d, rows, exp_ids, exp_d, what):
    """got_* are full batch outputs; rows index them; exp_* are the oracle's rows (k >= got k)."""
    k = got_ids.shape[1]
    gi, gd = got_ids[rows], got_d[rows]
    ei, ed = as_i64(exp_ids[:, :k]), exp_d[:, :k]
    bad = np.flatnonzero(~(np.all(gi == ei, axis=1) & np.all(bits(gd) == bits(ed), axis=1)))
    assert bad.size == 0, f"{what}: {bad.size} of {len(rows)} sampled queries differ (first: row {rows[bad[0]]})"


@pytest.fixture(scope="module")
def p2():
    print("\n[P2] 1,000,000 x 384 mixture, 4,096 held-out queries", flush=True)
    x = dg.mixture(1_000_000, 384, seed=0, workers=THREADS)
    q = dg.mixture(4096, 384, seed=1)
    idx = vs.GpuBruteForceIndex(384, "cosine")
    idx.add(x)
    ti, td = oracle.search(x, q[SAMPLE], 100, "ip", threads=THREADS)  # top-10 is the prefix
    plain = idx.search(q, 10, "skinny")
    yield dict(x=x, q=q, idx=idx, truth=(ti, td), plain=plain)
    del idx


@pytest.mark.parametrize("method", ["naive", "skinny"])
@pytest.mark.parametrize("batch", [1, 8, 48, 128, 4096])
def test_p2_exact_scan(p2, method, batch):
    # skinny from 32 queries runs the filtered scan's block-shared (32-63) or one-lane-per-row
    # (64 and up) tile over every row; naive and smaller batches the streaming scan
    ids, d = p2["idx"].search(p2["q"][:batch], 10, method)
    rows = SAMPLE[SAMPLE < batch]
    check_rows(ids, d, rows, p2["truth"][0][:len(rows)], p2["truth"][1][:len(rows)], f"P2 {method} b={batch}")
    np.testing.assert_array_equal(ids, p2["plain"][0][:batch])
    np.testing.assert_array_equal(bits(d), bits(p2["plain"][1][:batch]))


@pytest.mark.parametrize("batch", BATCHES)
def test_p2_auto_every_batch(p2, batch):
    ids, d = p2["idx"].search(p2["q"][:batch], 10)
    rows = SAMPLE[SAMPLE < batch]
    check_rows(ids, d, rows, p2["truth"][0][:len(rows)], p2["truth"][1][:len(rows)], f"P2 auto b={batch}")
    np.testing.assert_array_equal(ids, p2["plain"][0][:batch])
    np.testing.assert_array_equal(bits(d), bits(p2["plain"][1][:batch]))


@pytest.mark.parametrize("k", [10, 100])
@pytest.mark.parametrize("batch", [128, 1024, 4096])
@pytest.mark.parametrize("prec", ["bf16", "fp8"])
def test_q_shortlist(p2, prec, batch, k):
    ids, d = p2["idx"].search(p2["q"][:batch], k, prec)
    rows = SAMPLE[SAMPLE < batch]
    check_rows(ids, d, rows, p2["truth"][0][:len(rows)], p2["truth"][1][:len(rows)], f"Q {prec} b={batch} k={k}")
    if k == 10:
        np.testing.assert_array_equal(ids, p2["plain"][0][:batch])
        np.testing.assert_array_equal(bits(d), bits(p2["plain"][1][:batch]))
    else:  # the whole batch against the exact streaming scan at k = 100
        ei, ed = p2["idx"].search(p2["q"][:batch], k, "skinny")
        np.testing.assert_array_equal(ids, ei)
        np.testing.assert_array_equal(bits(d), bits(ed))


@pytest.fixture(scope="module")
def p2_l2(p2):
    idx = vs.GpuBruteForceIndex(384, "l2")
    idx.add(p2["x"])
    ti, td = oracle.search(p2["x"], p2["q"][SAMPLE], 100, "l2", threads=THREADS)
    yield dict(idx=idx, truth=(ti, td))
    del idx


@pytest.mark.parametrize("k", [10, 100])
@pytest.mark.parametrize("batch", [128, 4096])
@pytest.mark.parametrize("prec", ["bf16", "fp8"])
def test_l2_shortlist_full_size(p2, p2_l2, prec, batch, k):
    """Squared L2 at 1,000,000 x 384 through the certified shortlist."""
    q = p2["q"][:batch]
    ids, d = p2_l2["idx"].search(q, k, prec)
    rows = SAMPLE[SAMPLE < batch]
    check_rows(ids, d, rows, p2_l2["truth"][0][:len(rows)], p2_l2["truth"][1][:len(rows)], f"L2 {prec} b={batch} k={k}")
    ei, ed = p2_l2["idx"].search(q, k, "skinny")
    np.testing.assert_array_equal(ids, ei)
    np.testing.assert_array_equal(bits(d), bits(ed))


def test_q_fp64_error_bound(p2):
    """The returned distances are the FP32 values of the reference order; against FP64 their error
    stays within the sequential-sum bound gamma_384 * sum |q_i x_i| (unit rows: <= gamma_384)."""
    q, x = p2["q"][:16], p2["x"]
    ids, d = p2["idx"].search(q, 10, "bf16")
    u = 2.0 ** -24
    gamma = 384 * u / (1 - 384 * u)
    err = max(abs(float(d[i, j]) - oracle.distance_f64("ip", q[i], x[ids[i, j]])) for i in range(16) for j in range(10))
    assert err <= gamma + u


@pytest.fixture(scope="module")
def fwork():
    n = 892_800
    print("\n[F] 892,800 x 384 mixture, 100 groups (12 % .. 0.01 %), 4,096 filtered queries", flush=True)
    x = dg.mixture(n, 384, seed=4, workers=THREADS)
    groups = dg.filtered_groups(n, seed=2)
    off, allow = dg.group_csr(groups, dg.F_GROUPS)
    ks, qg = dg.query_filters(4096, dg.F_GROUPS, seed=3)
    ranges = dg.query_ranges(qg, off)
    q = dg.mixture(4096, 384, seed=5)
    idx = vs.GpuBruteForceIndex(384, "cosine")
    idx.add(x)
    yield dict(x=x, q=q, idx=idx, ks=ks, qg=qg, ranges=ranges, allow=allow, off=off)
    del idx


def test_f_batched_filtered_scan(fwork):
    w = fwork
    sizes = np.diff(w["off"].astype(np.int64))
    assert sizes.max() / len(w["x"]) > 0.11 and sizes.min() / len(w["x"]) < 0.0002
    ids, d = w["idx"].search_batch(w["q"], w["ks"], w["ranges"], w["allow"])
    ei, ed = oracle.search_lists(w["x"], w["q"], w["ks"], w["ranges"], w["allow"], "ip", threads=THREADS)
    bad = np.flatnonzero(~(np.all(ids == as_i64(ei), axis=1) & np.all(bits(d) == bits(ed), axis=1)))
    small = np.argsort(sizes)[:5]
    assert np.isin(w["qg"], small).any(), "the sample holds no query of the five smallest groups"
    assert bad.size == 0, f"F: {bad.size} of 4096 queries differ (first {bad[0]}, group size {sizes[w['qg'][bad[0]]]})"


def test_f_graph_replay_equals_eager(fwork):
    torch = pytest.importorskip("torch")
    w = fwork
    dq = torch.from_numpy(w["q"]).cuda()
    kmax = int(w["ks"].max())
    di = torch.empty((4096, kmax), dtype=torch.int32, device="cuda")
    dd = torch.empty((4096, kmax), dtype=torch.float32, device="cuda")
    g = vs.SearchGraph.lists(w["idx"], dq.data_ptr(), w["ks"], w["ranges"], w["allow"], di.data_ptr(),
                             dd.data_ptr(), kmax)
    ei, ed = w["idx"].search_batch(w["q"], w["ks"], w["ranges"], w["allow"])
    for _ in range(2):
        di.fill_(0)
        g.launch()
        torch.cuda.synchronize()
        gi = di.cpu().numpy().view(np.uint32)
        np.testing.assert_array_equal(np.where(gi == oracle.NONE, -1, gi.astype(np.int64)), ei)
        np.testing.assert_array_equal(bits(dd.cpu().numpy()), bits(ed))
    g.close()


def test_s_serving_alone_and_mixed():
    n, d = 2_000_000, 768
    print("\n[S] 2,000,000 x 768 mixture in the serving loop", flush=True)
    x = dg.mixture(n, d, seed=6, workers=THREADS)
    groups = dg.filtered_groups(n, seed=7)
    off, allow = dg.group_csr(groups, dg.F_GROUPS)
    idx = vs.GpuBruteForceIndex(d, "cosine")
    idx.add(x)
    rng = np.random.default_rng(8)
    reqs = []
    for i in range(40):
        nq = int(np.exp(rng.uniform(0, np.log(512))))
        k = int(rng.choice([10, 100]))
        g = int(rng.integers(0, dg.F_GROUPS)) if i % 2 else None
        q = dg.mixture(nq, d, seed=100 + i)
        reqs.append((q, k, None if g is None else allow[off[g]:off[g + 1]]))
    with vs.Server(idx, slots=2, tune=False) as srv:
        tickets = [srv.submit(q, k, a) for q, k, a in reqs]  # all in flight together
        mixed = [srv.wait(t)[:2] for t in tickets]
        alone = []
        for q, k, a in reqs[:12]:
            alone.append(srv.wait(srv.submit(q, k, a))[:2])
    for (mi, md), (ai, ad) in zip(mixed, alone):
        np.testing.assert_array_equal(mi, ai)
        np.testing.assert_array_equal(bits(md), bits(ad))
    # the oracle on the first query of every request (and the last of the filtered ones)
    for (q, k, a), (mi, md) in zip(reqs, mixed):
        rows = [0] if a is None else sorted({0, len(q) - 1})
        if a is None:
            ei, ed = oracle.search(x, q[rows], k, "ip", threads=THREADS)
        else:
            r = np.array([[0, len(a)]] * len(rows), np.uint64)
            ei, ed = oracle.search_lists(x, q[rows], np.full(len(rows), k, np.uint32), r, a, "ip", threads=THREADS)
        np.testing.assert_array_equal(mi[rows], as_i64(ei))
        np.testing.assert_array_equal(bits(md[rows]), bits(ed))
