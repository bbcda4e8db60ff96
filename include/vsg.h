/* vsg: the C ABI of the MI355X exact k-NN engine (libvsg_hip.so, built for gfx950).
 *
 * Plain pointers and sizes only. Every entry point returns a status code; on failure the
 * calling thread's message is in vsg_last_error(). This is the boundary the reference's
 * GpuBruteForceIndex (include/vecsearch/gpu_brute_force.h:26-50, src/gpu_brute_force.cu:
 * 334-488) is rebuilt behind: the C++ class in include/vecsearch/gpu_brute_force.h and the
 * pybind11 module call only these functions.
 *
 * Semantics shared by every search entry point (gpu_brute_force.h:40-43, brute_force.cpp:
 * 152-168, distance.h:1-3,71-73):
 *   - distances, not similarities: squared L2, or 1 - <q, x> for inner product / cosine;
 *   - nq x k outputs, row-major, nearest first, ties broken by the lower id;
 *   - a missing result (k > number of candidate rows) is id VSG_NONE with distance +inf;
 *   - every returned distance is bit-identical to the reference CPU index's FP32 value.
 */
#ifndef VSG_H
#define VSG_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vsg_index vsg_index;

/* status codes; the C++ class maps them onto the reference's exception types
 * (gpu_brute_force.cu:11-16, 297, 387, 408) */
#define VSG_OK 0
#define VSG_ERR_INVALID_ARGUMENT 1 /* std::invalid_argument -> Python ValueError */
#define VSG_ERR_LENGTH 2           /* std::length_error */
#define VSG_ERR_OUT_OF_MEMORY 3    /* std::runtime_error (device allocation failed) */
#define VSG_ERR_DEVICE 4           /* std::runtime_error (HIP runtime error) */
#define VSG_ERR_NO_DEVICE 5        /* std::runtime_error (no GPU visible) */

/* Metric (distance.h:15) */
#define VSG_METRIC_L2 0
#define VSG_METRIC_IP 1

/* Search method (gpu_brute_force.h:16-21). Every method returns the same ids and distances
 * (the reference CPU index's); they differ only in how the work is done:
 *   NAIVE   the exact streaming scan, one query per pass over the base;
 *   SKINNY  the exact streaming scan, up to 8 queries per pass (24 or more unmasked queries
 *           with dim a multiple of 32, on bases of at least 65,536 rows: the filtered scan over
 *           every row, which shares each chunk's row reads between the queries);
 *   TILED   the certified shortlist with a BF16 MFMA scan (large batches);
 *   BF16    the same as TILED;
 *   FP8     the certified shortlist with an FP8 E4M3 (per-row scale) MFMA scan;
 *   AUTO    SKINNY for small batches, TILED for large ones.
 * The shortlist applies to inner product / cosine and to squared L2 (its estimate carries each
 * row's |x|^2 / 2) with dim <= 768 on bases of at least 65,536 rows; elsewhere every method runs
 * the exact scan. */
#define VSG_METHOD_AUTO 0
#define VSG_METHOD_NAIVE 1
#define VSG_METHOD_SKINNY 2
#define VSG_METHOD_TILED 3
#define VSG_METHOD_BF16 4
#define VSG_METHOD_FP8 5

#define VSG_NONE 0xFFFFFFFFu /* gpu_brute_force.h:28 kNone */
#define VSG_MAX_K 128        /* gpu_brute_force.h:29 kMaxK */

/* Thread-local text of the last failure on this thread ("" if none). */
const char* vsg_last_error(void);

/* Number of visible HIP devices (0 without a driver); replaces cuda_available()
 * (gpu_brute_force.cu:334-341). */
int vsg_device_count(int* count);

/* GpuBruteForceIndex(dim, metric, device) (gpu_brute_force.cu:386-396). dim == 0 -> INVALID. */
int vsg_create(size_t dim, int metric, int device, vsg_index** out);
void vsg_destroy(vsg_index* h);

int vsg_size(const vsg_index* h, size_t* n);
int vsg_dim(const vsg_index* h, size_t* dim);
int vsg_metric(const vsg_index* h, int* metric);

/* Make room for n rows without copying on later adds. */
int vsg_reserve(vsg_index* h, size_t n);

/* add(data, n) (gpu_brute_force.cu:404-423): host rows, n x dim, appended with ids size()..;
 * size() + n >= 2^32 - 1 -> LENGTH. */
int vsg_add(vsg_index* h, const float* data, size_t n);

/* search(queries, nq, k, out_ids, out_dists, method) (gpu_brute_force.cu:425-488) with host
 * buffers, including the query upload and the result download. k > VSG_MAX_K -> INVALID.
 * mask (nullable) has size() entries; a nonzero entry allows that row for every query, the
 * convention of SearchFilter.allowed (hnsw.h:29-33) and of the `filter` argument of the
 * Python bindings (bindings.cpp:21-36). With a mask the result is the exact scan over the
 * allowed rows (hnsw.cpp:336-354); no allowed rows gives an all-missing result
 * (hnsw.cpp:326). */
int vsg_search(vsg_index* h, const float* queries, size_t nq, size_t k, int method,
               const uint8_t* mask, uint32_t* out_ids, float* out_dists);

/* The same search on device-resident buffers (queries nq x dim fp32, out_ids / out_dists
 * nq x k), enqueued on `stream` (a hipStream_t; NULL = the index's own stream). Returns after
 * enqueueing; the outputs are valid once the stream reaches that point. Two paths wait on the
 * stream before returning: the shortlist, once per round of up to 4,096 queries, to learn
 * whether a query overflowed its candidates (those are answered again exactly), and an exact
 * search of 24 or more unmasked queries, which runs as the filtered scan over every row and
 * uploads its plan first. */
int vsg_search_device(vsg_index* h, const float* d_queries, size_t nq, size_t k, int method,
                      const uint8_t* d_mask, uint32_t* d_ids, float* d_dists, void* stream);

/* No counterpart in the reference. For callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc. The search of
 * vsg_search_device (same arguments, same ids and distances for every method), answered by the
 * exact streaming scan whose only scratch memory is a static device array of
 * VSG_NOALLOC_SCRATCH_BYTES per GPU, loaded with the library and held until the process ends;
 * the call makes no device allocation and creates no stream or event. The scratch holds
 * q x W x k 8-byte keys for the q <= 8 queries of one pass (W = the scan's lists, at most 512
 * at the default 8 waves per CU on 256 CUs): a search whose single pass needs more (e.g.
 * scan_waves_per_cu 32 with k = 128) is rejected with VSG_ERR_INVALID_ARGUMENT. Larger batches
 * run in rounds. Every call on one GPU shares its array, and each call leaves it in its own
 * state: calls that overlap on one GPU (other streams, other threads, other indexes) race;
 * the caller orders them. The index must have been created (vsg_create sets the array up on
 * its device). */
#define VSG_NOALLOC_SCRATCH_BYTES (4u << 20)
int vsg_search_device_noalloc(vsg_index* h, const float* d_queries, size_t nq, size_t k, int method,
                              const uint8_t* d_mask, uint32_t* d_ids, float* d_dists, void* stream);

/* Batched filtered exact scan with a per-query k and per-query allowed-id lists: query i may
 * return only rows allow_ids[range[2i] .. range[2i+1]) (strictly increasing, < size()), and
 * returns ks[i] <= kmax results. Ranges may coincide: queries with identical ranges share
 * their row reads. n_ids = len(allow_ids) (every range must lie inside it). Outputs are
 * nq x kmax with columns ks[i].. padded. Host buffers. This is HNSWIndex::search's
 * exact_below path (hnsw.cpp:322-354) for many queries with different filters in one call.
 * A query whose range begins with VSG_ALL_ROWS (its end is ignored) is unfiltered: it scans
 * every row, in the same persistent launch as the filtered queries (GpuBruteForceIndex::search
 * semantics, gpu_brute_force.cu:425-488). */
#define VSG_ALL_ROWS 0xFFFFFFFFFFFFFFFFull
int vsg_search_lists(vsg_index* h, const float* queries, size_t nq, const uint32_t* ks,
                     size_t kmax, const uint64_t* range, const uint32_t* allow_ids, size_t n_ids,
                     uint32_t* out_ids, float* out_dists);

/* Launch-shape parameters. Known names:
 *   "scan_waves_per_cu"    waves per CU of the streaming scan (1..32, default 8);
 *   "auto_short_batch"     smallest batch that AUTO sends to the shortlist (default 2);
 *   "short_sample_stride"  the shortlist's first pass visits every stride-th tile of rows
 *                          (0 = chosen per call from the batch size and the tile count,
 *                          default; kept to at least k tiles);
 *   "lists_qb", "lists_r", "lists_chunk", "lists_blocks_per_cu", "lists_lanes_per_row"
 *                          a fixed tile of the filtered scan (queries per work item 1/2/4/8/16,
 *                          row groups per lane step 2/4, allowed rows per item, blocks per CU,
 *                          lanes per row 4 or 8 (8: d = 128 / 384 / 768, half the accumulator
 *                          registers per query-row pair)); qb 32 with r 8 and 8 lanes per row
 *                          (d = 128 / 384 / 768): one item per block, its 32 queries sharing
 *                          every row read; r 1 with 1 lane per row (d a multiple of 32): one
 *                          row per lane, 8 queries per item whatever qb; 0 = chosen per call
 *                          (default);
 *   "lists_autotune"       1 (default): with no fixed tile, the first large call of each shape
 *                          class times the candidate tiles on its own inputs and the class keeps
 *                          the fastest; 0: the planner's default tile, and the table is cleared.
 * vsg_get_param reads these, and "lists_tuned_classes" (size of the tuned table) and
 * "lists_last_qb" / "lists_last_r" / "lists_last_chunk" / "lists_last_lanes_per_row" (the tile
 * of the last filtered call). */
int vsg_set_param(vsg_index* h, const char* name, int64_t value);
int vsg_get_param(vsg_index* h, const char* name, int64_t* value);

/* Kernel timing for measurement: while enabled, every search records HIP events around its
 * kernel launches on the stream they run on, by kind. vsg_kernel_times waits for the recorded
 * events, writes the accumulated milliseconds and launch counts per kind since the last call
 * (arrays of VSG_KERNEL_KINDS entries), and resets them. */
#define VSG_KERNEL_SCAN 0    /* exact streaming scan passes */
#define VSG_KERNEL_MERGE 1   /* per-query merge of the scan's lists */
#define VSG_KERNEL_MFMA 2    /* shortlist MFMA scan over every row (candidate emission) */
#define VSG_KERNEL_RESCORE 3 /* shortlist exact rescoring */
#define VSG_KERNEL_SAMPLE 4  /* shortlist sampled MFMA pass, threshold select, candidate filing */
#define VSG_KERNEL_LISTS 5   /* filtered scan over allowed-id lists (persistent launch) */
#define VSG_KERNEL_KINDS 6
int vsg_kernel_timing(vsg_index* h, int enable);
int vsg_kernel_times(vsg_index* h, double* ms, uint64_t* launches);

/* Shortlist statistics of the last search that used it: candidates emitted (sum over queries),
 * candidates rescored exactly (sum), and queries that overflowed and took the exact scan. */
int vsg_short_stats(vsg_index* h, uint64_t* emitted, uint64_t* rescored, uint64_t* fallbacks);

/* Wait for all work the index has enqueued on its own stream. */
int vsg_synchronize(vsg_index* h);

/* Device pointer of the resident base (size() x padded_dim fp32, row-major, zero padded),
 * for measurement code that needs the exact HBM footprint. */
int vsg_base_info(const vsg_index* h, const float** d_base, size_t* padded_dim);

/* ---- saved HNSW indexes (hnsw.cpp:369-420, format "VSHNSW01") ----
 * A saved HNSWIndex holds a 52-byte header (magic, dim u64, n u64, metric u32, M u64,
 * ef_construction u64, max_level i32, entry_point u32) followed by its n x dim fp32 vectors
 * in id order, then the graph. vsg_hnsw_file_info reads the header; vsg_add_hnsw_file appends
 * the vector block to an index (streamed through pinned staging into HBM, ids continuing
 * from size()); the graph is not read. A wrong magic, a truncated file, or a dim / metric
 * different from the index's -> INVALID. */
int vsg_hnsw_file_info(const char* path, size_t* dim, size_t* n, int* metric);
int vsg_add_hnsw_file(vsg_index* h, const char* path, size_t* rows_added);

/* HNSWIndex::search's filter dispatch (hnsw.cpp:322-327): with a mask whose allowed count is
 * at most exact_below, the exact scan over the allowed rows answers the batch on the GPU and
 * *handled = 1 (no allowed row: every result missing, as hnsw.cpp:326); otherwise *handled = 0
 * and nothing is written (the caller runs its graph search). */
int vsg_search_exact_below(vsg_index* h, const float* queries, size_t nq, size_t k,
                           const uint8_t* mask, size_t exact_below, uint32_t* out_ids,
                           float* out_dists, int* handled);

/* ---- captured searches (no counterpart in the reference) ----
 * vsg_graph_create records one search of nq device-resident queries into a HIP graph with its
 * own workspace, on the path an eager call of the same method, nq and k takes: the streaming
 * scan and its merge (NAIVE / SKINNY, and AUTO below auto_short_batch) or the certified
 * shortlist (TILED / BF16 / FP8, and AUTO from auto_short_batch queries). Inside a captured
 * shortlist, queries whose candidates overflow (or every query of a round whose emission
 * regions overflow) are answered by an exact fallback that the device plans itself: a one-block
 * kernel picks them and one persistent launch of the filtered scan plus its merge rewrites
 * their rows, so the replay waits on nothing. vsg_graph_launch replays the graph on `stream`,
 * reading whatever d_queries holds at that point and writing d_ids / d_dists.
 * The buffers and the index must outlive the graph. The graph is bound to the base's size and
 * address at capture: after vsg_add, a launch returns INVALID (capture a new graph). */
typedef struct vsg_graph vsg_graph;
int vsg_graph_create(vsg_index* h, const float* d_queries, size_t nq, size_t k, int method,
                     uint32_t* d_ids, float* d_dists, vsg_graph** out);
/* vsg_graph_create_lists records one persistent launch of the batched filtered scan
 * (vsg_search_lists' arguments: per-query ks, ranges into allow_ids, VSG_ALL_ROWS for an
 * unfiltered query) over nq device-resident queries, writing nq x kmax results to d_ids /
 * d_dists. ks, ranges and allowed ids are host buffers, planned and copied into the graph's
 * workspace at capture (a replay reads new queries, not new filters). The tile is the one an
 * eager call would use: the fixed tile, else the tuned tile of the batch's shape class, else
 * the planner's. Replays are kernels only. */
int vsg_graph_create_lists(vsg_index* h, const float* d_queries, size_t nq, const uint32_t* ks, size_t kmax,
                           const uint64_t* range, const uint32_t* allow_ids, size_t n_ids, uint32_t* d_ids,
                           float* d_dists, vsg_graph** out);
int vsg_graph_launch(vsg_graph* g, void* stream);
void vsg_graph_destroy(vsg_graph* g);

/* ---- serving loop (no counterpart in the reference: its GPU index serialises every call
 * behind one mutex, gpu_brute_force.cu:351,430) ----
 * Requests (a block of queries, one k, optionally an allowed-id list) are queued by
 * vsg_server_submit and answered by continuous batching: each of `slots` worker threads owns a
 * HIP stream and a private workspace, takes as many queued requests as fit `max_batch_queries`
 * and the admission budget, runs them as one device batch, and completes them; requests that
 * arrive meanwhile join the next free slot's batch, so batches overlap on the device. A request
 * is admitted only while the estimated workspace bytes of the requests in flight stay within
 * `workspace_bytes` (the first request of an idle server is always admitted). Inside a batch,
 * filtered requests run in one persistent launch of the filtered scan and the unfiltered
 * queries, gathered into one block, take `method` (AUTO: the streaming scan for small blocks,
 * the certified shortlist for large ones). Results
 * equal the index's own search of the same request (bit-identical), alone or in any mix.
 * With method VSG_METHOD_NAIVE every request of a batch, filtered or not, runs in the one
 * persistent launch of the filtered scan (unfiltered ones as VSG_ALL_ROWS ranges).
 * The caller keeps queries, allow_ids and the output buffers alive until the request is
 * waited for. While a server exists, vsg_add / vsg_reserve wait for in-flight batches. */
typedef struct vsg_server vsg_server;
typedef struct {
  size_t max_batch_queries; /* queries per device batch (0: 4096) */
  size_t workspace_bytes;   /* admission budget of requests in flight (0: 4 GiB) */
  int slots;                /* batches in flight, one stream each (0: 2) */
  int method;               /* VSG_METHOD_* for unfiltered queries (default AUTO) */
  int tune_mode;            /* 0: a batch whose filtered shape class is not tuned yet is tuned
                               inline (vsg_get_param "lists_autotune"); 1: such a batch runs the
                               planner's tile, so no batch waits for candidate runs */
} vsg_server_config;

typedef struct {
  uint64_t requests, queries, batches; /* completed */
  uint64_t admission_waits;            /* times a free slot waited for workspace budget */
  uint64_t peak_inflight_bytes;        /* largest admitted estimate in flight */
  uint64_t max_batch_requests;         /* most requests in one batch */
  uint64_t shortlist_fallbacks;        /* unfiltered queries answered by the exact scan after
                                          their shortlist candidates overflowed */
} vsg_server_stats;

int vsg_server_create(vsg_index* h, const vsg_server_config* cfg, vsg_server** out);
/* allow_ids == NULL: unfiltered; else the request may return only rows allow_ids[0 .. n_allowed)
 * (strictly increasing, < size(); n_allowed == 0 returns nothing). k <= VSG_MAX_K.
 * Outputs: nq x k ids / distances. *ticket identifies the request for vsg_server_wait. */
int vsg_server_submit(vsg_server* s, const float* queries, size_t nq, size_t k,
                      const uint32_t* allow_ids, size_t n_allowed, uint32_t* out_ids,
                      float* out_dists, uint64_t* ticket);
/* Blocks until the request is done; returns its status (VSG_OK or the error of its batch) and,
 * if latency_ms is not NULL, the time from submit to completion. A ticket is waited for once. */
int vsg_server_wait(vsg_server* s, uint64_t ticket, double* latency_ms);
int vsg_server_get_stats(vsg_server* s, vsg_server_stats* out);
/* Finishes every queued request, then stops the workers. */
void vsg_server_destroy(vsg_server* s);

#ifdef __cplusplus
}
#endif
#endif
