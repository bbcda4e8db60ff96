"""Captured exact searches (vsg_graph_*): every replay equals the eager search of the queries the
buffer holds at that moment, with eager work of the index between the replays. (AUTO captures
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
    with pytest.raises(ValueError):  # the shortlist is not capturable
        vs.SearchGraph(gpu, d_q.data_ptr(), 64, 10, d_ids.data_ptr(), d_dists.data_ptr(), method="bf16")
    g = vs.SearchGraph(gpu, d_q.data_ptr(), 64, 10, d_ids.data_ptr(), d_dists.data_ptr(), method="skinny")
    g.launch(0)
    gpu.synchronize()
    gpu.add(x[:16])  # the index grew: the captured plan no longer covers it
    with pytest.raises(RuntimeError):
        g.launch(0)
    g.close()
