"""A GPU shard worker behind the reference's gRPC search service (distributed/search.proto).

    python -m vecsearch_amd.shard_worker --manifest shards/manifest.json --shard-id 0 --port 50051

The reference's worker (distributed/shard_worker.py:18-47) loads its shard's saved HNSWIndex
and answers each Search RPC with a CPU graph search. This one serves the same service, on the
same manifest (build_shards.py:14-31: {"dim", "n", "shards": [{"file", "offset", "n"}]}):

  - the shard's vectors are read from its saved HNSW file (the VSHNSW01 vector block; the graph
    is not needed) into a GPU index with the file's dim and metric;
  - Search: every concurrent RPC is submitted to the engine's continuous-batching Server, so
    requests in flight on the gRPC thread pool share device batches instead of taking turns;
    the answer is exact (the reference CPU index's ids and distances for the shard), `ef` is
    accepted and ignored (it tunes the graph search this replaces);
  - the response has the reference's shape: global ids (local id + the shard's offset) with
    their distances ascending, missing results dropped, and timings_ms {"shard_<id>": ms};
  - Info: num_vectors, dim, num_shards = 1.
The reference's coordinator (distributed/coordinator.py) merges such workers' answers unchanged.
A request whose query has the wrong length is answered INVALID_ARGUMENT, as the reference's is.
"""
import argparse
import json
import os
import time
from concurrent import futures

import numpy as np

from . import GpuBruteForceIndex, Server
from . import search_proto as sp


class GpuShardServicer:
    """Search / Info of the vecsearch.Searcher service over one shard on one GPU."""

    def __init__(self, index_path, offset, shard_id, device=0, slots=2, max_batch_queries=4096):
        t = time.time()
        self.idx = GpuBruteForceIndex.from_hnsw_file(index_path, device)
        self.dim = self.idx.dim
        self.offset = int(offset)
        self.shard_id = int(shard_id)
        self.server = Server(self.idx, max_batch_queries=max_batch_queries, slots=slots, method="auto")
        print(f"[gpu shard {shard_id}] {len(self.idx):,} vectors of dim {self.dim} in {time.time() - t:.1f}s",
              flush=True)

    def Search(self, req, ctx):
        import grpc

        t = time.perf_counter()
        q = np.frombuffer(req.query, dtype=np.float32)
        if q.shape[0] != self.dim:
            ctx.abort(grpc.StatusCode.INVALID_ARGUMENT, f"expected dim {self.dim}, got {q.shape[0]}")
        k = int(req.k)
        if k <= 0 or k > 128:
            ctx.abort(grpc.StatusCode.INVALID_ARGUMENT, "k must be in 1..128")
        ids, d, _ = self.server.wait(self.server.submit(q, k))
        ids, d = ids[0], d[0]
        keep = ids >= 0
        return sp.SearchResponse(ids=(ids[keep] + self.offset).tolist(), distances=d[keep].tolist(),
                                 timings_ms={f"shard_{self.shard_id}": (time.perf_counter() - t) * 1e3})

    def Info(self, req, ctx):
        return sp.InfoResponse(num_vectors=len(self.idx), dim=self.dim, num_shards=1)

    def close(self):
        self.server.close()


def serve(manifest, shard_id, port, threads=16, device=0, host="[::]"):
    """Starts the worker's gRPC server; returns (server, servicer, bound port)."""
    import grpc

    shard = json.load(open(manifest))["shards"][shard_id]
    path = os.path.join(os.path.dirname(os.path.abspath(manifest)), shard["file"])
    servicer = GpuShardServicer(path, shard["offset"], shard_id, device)
    server = grpc.server(futures.ThreadPoolExecutor(max_workers=threads), options=[("grpc.so_reuseport", 0)])
    sp.add_searcher_to_server(servicer, server)
    bound = server.add_insecure_port(f"{host}:{port}")
    server.start()
    return server, servicer, bound


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--manifest", required=True)
    ap.add_argument("--shard-id", type=int, required=True)
    ap.add_argument("--port", type=int, required=True)
    ap.add_argument("--threads", type=int, default=16, help="gRPC handler threads (requests in flight)")
    ap.add_argument("--device", type=int, default=0)
    a = ap.parse_args(argv)
    server, servicer, bound = serve(a.manifest, a.shard_id, a.port, a.threads, a.device)
    print(f"[gpu shard {a.shard_id}] serving on :{bound}", flush=True)
    try:
        server.wait_for_termination()
    finally:
        servicer.close()


if __name__ == "__main__":
    main()
