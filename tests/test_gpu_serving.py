"""Serving loop (vsg_server_*): continuous batching returns, for every request, the same bits
alone and in a mix of concurrent filtered / unfiltered requests, and those bits are the
oracle's (the reference CPU index's exact answer over the request's allowed rows)."""
import numpy as np
import pytest

import oracle
import vecsearch_amd as vs
from vecsearch_amd import datagen as dg

pytestmark = pytest.mark.gpu


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def as_i64(ids):
    return np.where(ids == oracle.NONE, -1, ids.astype(np.int64))


@pytest.fixture(scope="module")
def served():
    n, d = 70_000, 128
    x = dg.mixture(n, d, seed=60)
    groups = dg.zipf_groups(n, 30, 1.1, seed=61)
    off, gids = dg.group_csr(groups, 30)
    gpu = vs.GpuBruteForceIndex(d, "cosine")
    gpu.add(x)
    rng = np.random.default_rng(62)
    reqs = []
    for i in range(36):
        nq = int(np.exp(rng.uniform(0, np.log(200))))
        if i % 9 == 0:
            nq = 64  # unfiltered groups large enough for the shortlist path
        q = dg.mixture(nq, d, seed=100 + i)
        k = int(rng.choice([1, 10, 100]))
        allow = None
        if i % 2:
            g = int(rng.integers(0, 30))
            allow = np.sort(gids[off[g]:off[g + 1]]).astype(np.uint32)
        reqs.append((q, k, allow))
    return x, gpu, reqs


def expected(x, q, k, allow):
    if allow is None:
        return oracle.search(x, q, k, "ip", threads=8)
    ranges = np.tile(np.array([[0, len(allow)]], np.uint64), (len(q), 1))
    return oracle.search_lists(x, q, np.full(len(q), k, np.uint32), ranges, allow, "ip", threads=8, kmax=k)


def test_same_alone_and_in_the_mix(served):
    x, gpu, reqs = served
    with vs.Server(gpu, max_batch_queries=512, slots=2) as srv:
        alone = []
        for q, k, allow in reqs:
            ids, dists, _ = srv.wait(srv.submit(q, k, allow))
            alone.append((ids, dists))
        tickets = [srv.submit(q, k, allow) for q, k, allow in reqs]  # all in flight together
        mixed = [srv.wait(t) for t in tickets]
        st = srv.stats()
    assert st["requests"] == 2 * len(reqs)
    assert st["max_batch_requests"] > 1  # the concurrent pass was batched
    for (q, k, allow), (ai, ad), (mi, md, lat) in zip(reqs, alone, mixed):
        np.testing.assert_array_equal(ai, mi)
        np.testing.assert_array_equal(bits(ad), bits(md))
        ei, ed = expected(x, q, k, allow)
        np.testing.assert_array_equal(mi, as_i64(ei))
        np.testing.assert_array_equal(bits(md), bits(ed))
        assert lat >= 0.0
        if allow is not None:
            assert np.isin(mi[mi >= 0], allow).all()


def test_one_persistent_launch_mode(served):
    # method "naive": unfiltered requests join the filtered scan's launch as ALL_ROWS ranges
    x, gpu, reqs = served
    with vs.Server(gpu, max_batch_queries=512, slots=1, method="naive") as srv:
        tickets = [srv.submit(q, k, allow) for q, k, allow in reqs[:10]]
        got = [srv.wait(t) for t in tickets]
    for (q, k, allow), (gi, gd, _) in zip(reqs[:10], got):
        ei, ed = expected(x, q, k, allow)
        np.testing.assert_array_equal(gi, as_i64(ei))
        np.testing.assert_array_equal(bits(gd), bits(ed))


def test_admission_by_workspace_budget(served):
    x, gpu, reqs = served
    # a budget below two requests' estimates: requests are admitted one batch at a time
    with vs.Server(gpu, max_batch_queries=4096, workspace_bytes=1 << 20, slots=2) as srv:
        tickets = [srv.submit(q, k, allow) for q, k, allow in reqs[:12]]
        got = [srv.wait(t) for t in tickets]
        st = srv.stats()
    assert st["batches"] >= 6
    for (q, k, allow), (gi, gd, _) in zip(reqs[:12], got):
        ei, ed = expected(x, q, k, allow)
        np.testing.assert_array_equal(gi, as_i64(ei))
        np.testing.assert_array_equal(bits(gd), bits(ed))


def test_empty_filter_and_errors(served):
    x, gpu, reqs = served
    q = reqs[0][0][:3]
    with vs.Server(gpu) as srv:
        ids, dists, _ = srv.wait(srv.submit(q, 5, np.zeros(0, np.uint32)))
        assert (ids == -1).all() and np.isinf(dists).all()
        with pytest.raises(ValueError):
            srv.submit(q, 129)
        with pytest.raises(ValueError):
            srv.submit(q, 5, np.array([9, 3], np.uint32))  # not increasing
        with pytest.raises(ValueError):
            srv.submit(q, 5, np.array([len(x)], np.uint32))  # out of range
        with pytest.raises(ValueError):
            srv.wait(12345)
