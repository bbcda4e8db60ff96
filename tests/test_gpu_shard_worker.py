"""GPU shard workers behind the reference's gRPC service (distributed/search.proto): two shards
of one base, each served from its saved-index file, queried concurrently over gRPC; each answer
is the oracle's exact top-k of its shard in global ids, and merging the shards' answers as the
reference's coordinator does (coordinator.py:31-40) gives the oracle's top-k of the whole base."""
import heapq
import json
import struct
import threading

import numpy as np
import pytest

import oracle
from vecsearch_amd import datagen as dg
from vecsearch_amd import search_proto as sp
from vecsearch_amd import shard_worker

pytestmark = pytest.mark.gpu
grpc = pytest.importorskip("grpc")


def write_vshnsw(path, x, metric):
    """The header and vector block of a saved HNSWIndex (hnsw.cpp:369-389); the graph that follows
    them in a real file is not read by the worker."""
    n, d = x.shape
    with open(path, "wb") as f:
        f.write(b"VSHNSW01")
        f.write(struct.pack("<QQIQQiI", d, n, metric, 16, 200, 0, 0))
        f.write(np.ascontiguousarray(x, np.float32).tobytes())


def test_two_gpu_shards_over_grpc(tmp_path):
    x = dg.mixture(40_000, 64, seed=80)
    q = dg.mixture(48, 64, seed=81)
    cuts = [0, 25_000, 40_000]
    manifest = {"dim": 64, "n": len(x), "shards": []}
    for i in range(2):
        name = f"shard_{i}.bin"
        write_vshnsw(tmp_path / name, x[cuts[i]:cuts[i + 1]], 1)
        manifest["shards"].append({"file": name, "offset": cuts[i], "n": cuts[i + 1] - cuts[i]})
    (tmp_path / "manifest.json").write_text(json.dumps(manifest))
    servers = [shard_worker.serve(str(tmp_path / "manifest.json"), i, 0, threads=16, host="127.0.0.1")
               for i in range(2)]
    try:
        stubs = [sp.SearcherStub(grpc.insecure_channel(f"127.0.0.1:{port}")) for _, _, port in servers]
        infos = [s.Info(sp.InfoRequest(), timeout=30) for s in stubs]
        assert [i.num_vectors for i in infos] == [25_000, 15_000] and all(i.dim == 64 for i in infos)

        ks = [10 if j % 3 else 37 for j in range(len(q))]
        out = {}

        def ask(j, s):
            out[(j, s)] = stubs[s].Search(sp.SearchRequest(query=q[j].tobytes(), k=ks[j], ef=64), timeout=60)

        threads = [threading.Thread(target=ask, args=(j, s)) for j in range(len(q)) for s in range(2)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        for s in range(2):
            part = x[cuts[s]:cuts[s + 1]]
            for j in range(len(q)):
                ei, ed = oracle.search(part, q[j:j + 1], ks[j], "ip")
                r = out[(j, s)]
                assert list(r.ids) == [int(i) + cuts[s] for i in ei[0]]
                assert np.array_equal(np.asarray(r.distances, np.float32).view(np.uint32), ed[0].view(np.uint32))
                assert f"shard_{s}" in r.timings_ms
        # the coordinator's merge of the two shard lists equals the whole base's top-k
        for j in range(len(q)):
            lists = [zip(out[(j, s)].distances, out[(j, s)].ids) for s in range(2)]
            top = heapq.nsmallest(ks[j], heapq.merge(*lists))
            ei, ed = oracle.search(x, q[j:j + 1], ks[j], "ip")
            assert [i for _, i in top] == list(ei[0].astype(np.int64))
        with pytest.raises(grpc.RpcError) as err:
            stubs[0].Search(sp.SearchRequest(query=q[0][:10].tobytes(), k=10), timeout=30)
        assert err.value.code() == grpc.StatusCode.INVALID_ARGUMENT
    finally:
        for server, servicer, _ in servers:
            server.stop(None)
            servicer.close()
