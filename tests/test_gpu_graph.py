"""Captured exact searches (vsg_graph_*, vsg_graph_create_lists): every replay equals the eager
search of the queries the buffer holds at that moment, with eager work of the index between the
replays. (AUTO captures
the streaming scan; eager AUTO of 13 queries takes the shortlist: the same bits.)"""
import numpy as np
import pytest

import oracle
import vecsearch_amd as vs
from vecsearch_amd import datagen as dg

pytestmark = pytest.mark.gpu


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("nq,k,method", [(1, 10, "naive"), (8, 10, "skinny"), (13, 100, "skinny"), (13, 10, "auto")])
def test_graph_replays_match_eager(nq, k, method):
    import torch

    d = 384
    x = dg.mixture(50_000, d, seed=70)
    gpu = vs.GpuBruteForceIndex(d, "cosine")
    gpu.add(x)
    dev = torch.device("cuda:0")
    d_q = torch.empty((nq, d), dtype=torch.float32, device=dev)
    d_ids = torch.empty((nq, k), dtype=torch.int32, device=dev)
    d_dists = torch.empty((nq, k), dtype=torch.float32, device=dev)
    stream = torch.cuda.Stream(device=dev)
    graph = vs.SearchGraph(gpu, d_q.data_ptr(), nq, k, d_ids.data_ptr(), d_dists.data_ptr(), method=method)
    for rep in range(4):
        q = dg.mixture(nq, d, seed=80 + rep)
        d_q.copy_(torch.from_numpy(q))
        torch.cuda.synchronize()
        graph.launch(stream.cuda_stream)
        stream.synchronize()
        got_ids = d_ids.cpu().numpy().view(np.uint32)
        got_d = d_dists.cpu().numpy()
        ei, ed = gpu.search(q, k, method=method)  # eager work between replays
        np.testing.assert_array_equal(np.where(got_ids == oracle.NONE, -1, got_ids.astype(np.int64)), ei)
        np.testing.assert_array_equal(bits(got_d), bits(ed))
        if rep == 0:
            oi, od = oracle.search(x, q, k, "ip", threads=8)
            np.testing.assert_array_equal(got_ids, oi)
            np.testing.assert_array_equal(bits(got_d), bits(od))
    graph.close()


def test_graph_limits():
    import torch

    d = 64
    x = dg.mixture(70_000, d, seed=71)
    gpu = vs.GpuBruteForceIndex(d, "cosine")
    gpu.add(x)
    dev = torch.device("cuda:0")
    d_q = torch.zeros((64, d), dtype=torch.float32, device=dev)
    d_ids = torch.empty((64, 10), dtype=torch.int32, device=dev)
    d_dists = torch.empty((64, 10), dtype=torch.float32, device=dev)
    with pytest.raises(ValueError):  # k > 128
        vs.SearchGraph(gpu, d_q.data_ptr(), 64, 200, d_ids.data_ptr(), d_dists.data_ptr(), method="skinny")
    g = vs.SearchGraph(gpu, d_q.data_ptr(), 64, 10, d_ids.data_ptr(), d_dists.data_ptr(), method="skinny")
    g.launch(0)
    gpu.synchronize()
    gpu.add(x[:16])  # the index grew: the captured plan no longer covers it
    with pytest.raises(RuntimeError):
        g.launch(0)
    g.close()


@pytest.mark.parametrize("metric,tile", [("ip", None), ("l2", None), ("ip", (32, 8, 1024, 8)), ("ip", (16, 2, 2048, 8))])
def test_lists_graph_replays_match_eager(metric, tile):
    # one persistent launch of the filtered scan (filtered and unfiltered queries, k in {1, 10, 100})
    # captured once; each replay reads new queries under the filters fixed at capture
    import torch

    n, d, nq = 40_000, 384, 120
    x = dg.mixture(n, d, seed=73)
    groups = dg.zipf_groups(n, 30, 1.1, seed=74)
    off, ids_g = dg.group_csr(groups, 30)
    ks, qg = dg.query_filters(nq, 30, seed=75, ks=(1, 10, 100))
    ranges = dg.query_ranges(qg, off)
    ranges[::4] = [vs.ALL_ROWS, 0]
    gpu = vs.GpuBruteForceIndex(d, metric)
    gpu.add(x)
    if tile:
        for name, v in zip(("lists_qb", "lists_r", "lists_chunk", "lists_lanes_per_row"), tile):
            gpu.set_param(name, v)
    dev = torch.device("cuda:0")
    d_q = torch.empty((nq, d), dtype=torch.float32, device=dev)
    d_ids = torch.empty((nq, 100), dtype=torch.int32, device=dev)
    d_dists = torch.empty((nq, 100), dtype=torch.float32, device=dev)
    stream = torch.cuda.Stream(device=dev)
    graph = vs.SearchGraph.lists(gpu, d_q.data_ptr(), ks, ranges, ids_g, d_ids.data_ptr(), d_dists.data_ptr(),
                                 kmax=100)
    for rep in range(3):
        q = dg.mixture(nq, d, seed=90 + rep)
        d_q.copy_(torch.from_numpy(q))
        torch.cuda.synchronize()
        graph.launch(stream.cuda_stream)
        stream.synchronize()
        got_ids = d_ids.cpu().numpy().view(np.uint32)
        got_d = d_dists.cpu().numpy()
        ei, ed = gpu.search_batch(q, ks, ranges, ids_g, kmax=100)  # eager work between replays
        np.testing.assert_array_equal(np.where(got_ids == oracle.NONE, -1, got_ids.astype(np.int64)), ei)
        np.testing.assert_array_equal(bits(got_d), bits(ed))
        if rep == 0:
            oi, od = oracle.search_lists(x, q, ks, ranges, ids_g, metric, threads=8, kmax=100)
            np.testing.assert_array_equal(got_ids, oi)
            np.testing.assert_array_equal(bits(got_d), bits(od))
    graph.close()


def test_lists_graph_limits():
    import torch

    d = 64
    x = dg.mixture(5_000, d, seed=76)
    gpu = vs.GpuBruteForceIndex(d, "cosine")
    gpu.add(x)
    dev = torch.device("cuda:0")
    d_q = torch.zeros((4, d), dtype=torch.float32, device=dev)
    d_ids = torch.empty((4, 10), dtype=torch.int32, device=dev)
    d_dists = torch.empty((4, 10), dtype=torch.float32, device=dev)
    ks = np.full(4, 10, np.uint32)
    ranges = np.array([[0, 2]] * 4, np.uint64)
    with pytest.raises(ValueError):  # allowed ids must be increasing and < size()
        vs.SearchGraph.lists(gpu, d_q.data_ptr(), ks, ranges, np.array([7, 3], np.uint32), d_ids.data_ptr(),
                             d_dists.data_ptr())
    with pytest.raises(ValueError):  # k > 128
        vs.SearchGraph.lists(gpu, d_q.data_ptr(), np.full(4, 200, np.uint32), ranges,
                             np.array([3, 7], np.uint32), d_ids.data_ptr(), d_dists.data_ptr())
    g = vs.SearchGraph.lists(gpu, d_q.data_ptr(), ks, ranges, np.array([3, 7], np.uint32), d_ids.data_ptr(),
                             d_dists.data_ptr())
    g.launch(0)
    gpu.synchronize()
    ids = d_ids.cpu().numpy().view(np.uint32)
    assert set(ids[:, :2].ravel()) == {3, 7} and (ids[:, 2:] == oracle.NONE).all()
    gpu.add(x[:16])  # the index grew
    with pytest.raises(RuntimeError):
        g.launch(0)
    g.close()


def _replay_and_check(gpu, x, make_q, nq, k, method, reps=3, metric="ip"):
    import torch

    d = x.shape[1]
    dev = torch.device("cuda:0")
    d_q = torch.empty((nq, d), dtype=torch.float32, device=dev)
    d_ids = torch.empty((nq, k), dtype=torch.int32, device=dev)
    d_dists = torch.empty((nq, k), dtype=torch.float32, device=dev)
    stream = torch.cuda.Stream(device=dev)
    graph = vs.SearchGraph(gpu, d_q.data_ptr(), nq, k, d_ids.data_ptr(), d_dists.data_ptr(), method=method)
    for rep in range(reps):
        q = make_q(rep)
        d_q.copy_(torch.from_numpy(q))
        torch.cuda.synchronize()
        graph.launch(stream.cuda_stream)
        stream.synchronize()
        got_ids = d_ids.cpu().numpy().view(np.uint32)
        got_d = d_dists.cpu().numpy()
        ei, ed = gpu.search(q, k, method=method)  # eager work between replays
        np.testing.assert_array_equal(np.where(got_ids == oracle.NONE, -1, got_ids.astype(np.int64)), ei)
        np.testing.assert_array_equal(bits(got_d), bits(ed))
        if rep == 0:
            oi, od = oracle.search(x, q, k, metric, threads=8)
            np.testing.assert_array_equal(got_ids, oi)
            np.testing.assert_array_equal(bits(got_d), bits(od))
    graph.close()


@pytest.mark.parametrize("nq,k,method", [(13, 10, "bf16"), (128, 10, "bf16"), (300, 100, "bf16"),
                                         (128, 10, "fp8"), (100, 100, "fp8"), (64, 10, "auto"), (40, 10, "tiled")])
def test_shortlist_graph_replays_match_eager(nq, k, method):
    x = dg.mixture(80_000, 384, seed=77)
    gpu = vs.GpuBruteForceIndex(384, "cosine")
    gpu.add(x)
    _replay_and_check(gpu, x, lambda rep: dg.mixture(nq, 384, seed=100 + rep), nq, k, method)


def test_shortlist_graph_candidate_overflow():
    # 9,000 tied rows overflow the candidate buffer of the queries near them: inside the graph
    # those queries are answered by the device-planned exact fallback, the others by the shortlist
    x = dg.mixture(70_000, 128, seed=44)
    x[10_000:19_000] = x[5]
    gpu = vs.GpuBruteForceIndex(128, "cosine")
    gpu.add(x)

    def make_q(rep):
        near = x[5:6] * 0.999 + 0.001 * dg.mixture(4, 128, seed=46 + rep)
        q = np.concatenate([near[:1], dg.mixture(40, 128, seed=45 + rep), near[1:], dg.mixture(20, 128, seed=47 + rep)])
        return np.ascontiguousarray(vs.normalize(q.astype(np.float32)), dtype=np.float32)

    _replay_and_check(gpu, x, make_q, 64, 50, "bf16")
    gpu.search(make_q(0), 50, method="bf16")
    assert gpu.short_stats()["fallbacks"] >= 4


def test_shortlist_region_overflow_eager_and_graph():
    # 600 queries (the rows-stationary pass 2) all near 3,000 tied rows: the tiles holding those
    # rows emit more candidates than their regions hold, and the whole batch is answered by the
    # exact scan (eager: on the host's decision; captured: by the device-planned fallback of
    # every query)
    x = dg.mixture(70_000, 128, seed=50)
    x[20_000:23_000] = x[7]
    gpu = vs.GpuBruteForceIndex(128, "cosine")
    gpu.add(x)

    def make_q(rep):
        q = x[7:8] * 0.999 + 0.001 * dg.mixture(600, 128, seed=51 + rep)
        return np.ascontiguousarray(vs.normalize(q.astype(np.float32)), dtype=np.float32)

    q = make_q(0)
    got = gpu.search(q, 10, method="bf16")
    assert gpu.short_stats()["fallbacks"] == 600
    oi, od = oracle.search(x, q, 10, "ip", threads=8)
    np.testing.assert_array_equal(got[0], oi.astype(np.int64))
    np.testing.assert_array_equal(bits(got[1]), bits(od))
    _replay_and_check(gpu, x, make_q, 600, 10, "bf16", reps=2)


def test_shortlist_graph_several_rounds():
    # 4,100 queries: a full round of 4,096 and a ragged round of 4, each ending with the
    # device-planned fallback sized for a full round
    import torch

    nq, k, d = 4100, 10, 128
    x = dg.mixture(70_000, d, seed=78)
    gpu = vs.GpuBruteForceIndex(d, "cosine")
    gpu.add(x)
    dev = torch.device("cuda:0")
    q = dg.mixture(nq, d, seed=79)
    d_q = torch.from_numpy(q).to(dev)
    d_ids = torch.empty((nq, k), dtype=torch.int32, device=dev)
    d_dists = torch.empty((nq, k), dtype=torch.float32, device=dev)
    graph = vs.SearchGraph(gpu, d_q.data_ptr(), nq, k, d_ids.data_ptr(), d_dists.data_ptr(), method="bf16")
    for _ in range(2):
        graph.launch(0)
        gpu.synchronize()
        got_ids = d_ids.cpu().numpy().view(np.uint32)
        got_d = d_dists.cpu().numpy()
        ei, ed = gpu.search(q, k, method="bf16")
        np.testing.assert_array_equal(np.where(got_ids == oracle.NONE, -1, got_ids.astype(np.int64)), ei)
        np.testing.assert_array_equal(bits(got_d), bits(ed))
    sel = np.r_[0:40, nq - 24:nq]  # both rounds against the oracle
    oi, od = oracle.search(x, q[sel], k, "ip", threads=8)
    np.testing.assert_array_equal(got_ids[sel], oi)
    np.testing.assert_array_equal(bits(got_d[sel]), bits(od))
    graph.close()
