// Internal interfaces between the C-ABI layer (vsg_api.hip) and the kernel files.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <string>

namespace vsg {

constexpr int kWavesPerBlock = 4;
constexpr int kBlock = 64 * kWavesPerBlock;
constexpr int kMergeHeads = 16;  // lists per merge thread: a query merges <= 256 * 16 lists
constexpr uint32_t kMaxListsPerQuery = 256 * kMergeHeads;
constexpr uint32_t kNoneId = 0xFFFFFFFFu;

// vsg_scan.hip
hipError_t launch_scan(int qb, bool l2, uint32_t k, uint32_t n_blocks, hipStream_t st,
                       const float* tiles, uint32_t n, uint32_t dp, uint32_t d, const float* q,
                       uint32_t q0, uint32_t nqp, uint32_t rpw, const uint8_t* mask,
                       uint64_t* part);
hipError_t launch_merge(uint32_t n_out, const uint64_t* part, const uint64_t* list_off, uint32_t W,
                        uint32_t L, const uint32_t* ks, uint32_t k, const uint32_t* qmap,
                        uint32_t* ids, float* dists, uint32_t ldo, hipStream_t st);
hipError_t launch_fill_missing(hipStream_t st, uint32_t* ids, float* dists, size_t count);
// rows of `width` 32-bit words: dst[i] = src[idx[i]] (gather) or dst[idx[i]] = src[i] (scatter)
hipError_t launch_rows_move(hipStream_t st, const void* src, void* dst, const uint32_t* idx,
                            uint32_t rows, uint32_t width, bool scatter);
hipError_t launch_tile_rows(hipStream_t st, const float* base, float* tiles, uint32_t r0, uint32_t r1,
                            uint32_t dp);

// vsg_short.hip: certified low-precision shortlist (prec 1 = bf16, 2 = fp8 e4m3)
struct ShortScanArgs {
  hipStream_t st;
  int KS;
  const void* xfrag;
  const float* xscale;
  const float* xhn = nullptr;  // L2: |x|^2 / 2 per row (the rows-stationary pass subtracts it)
  uint32_t n;
  const void* qfrag;
  const float* qaux;  // per query fragment: 16 score cuts, 16 scales, 16 folded FP8 cuts
  uint32_t nq;
  uint32_t n_blocks;  // one block per visited tile
  uint32_t n_space;   // visited tiles: pass 1 ceil(n_tiles / tstride), pass 2 n_tiles
  uint32_t tstride;   // pass 1 visits every tstride-th tile; pass 2 uses 1
  uint32_t qsplit = 1;  // pass 1: blocks per visited tile (n_blocks = n_space * qsplit)
  uint32_t* tilemin;
  uint64_t* emit;         // pass 2: per wave region (8 per tile), wave_cap entries
  uint32_t wave_cap;
  uint32_t* emit_count;   // per wave region
  uint64_t* cand;         // per query, cap keys (filled by the scatter kernel)
  uint32_t cap;
  uint32_t* count;
  uint32_t* overflow;     // blocks whose region overflowed
  uint32_t* slice;        // nq * short_scatter_slices() counters of the scatter
  // queries-stationary pass 2 (short_qs_plan): blocks per query group (0: rows-stationary
  // pass 2), query groups, rows per block
  uint32_t qs_blocks = 0, qs_groups = 0, qs_rows = 0;
};
size_t short_scatter_slices();
// Queries-stationary pass 2 for a batch of at most one query group: its launch shape (blocks of
// `rows` rows), or false where it does not apply.
bool short_qs_plan(int prec, int ks, uint32_t n, uint32_t nq, int num_cus, uint32_t* blocks, uint32_t* groups,
                   uint32_t* rows, uint32_t* queries_per_region);
hipError_t launch_mfma_qs(int prec, const ShortScanArgs& a);
int short_ks(uint32_t d);
int short_cand_cap();
int short_tile_rows_for(int prec, int ks);
int short_group(int prec, int ks);
int short_aux_pad();
int short_aux_floats();  // floats of a query fragment's aux block
hipError_t launch_quantize(hipStream_t st, const float* src, uint32_t ld, uint32_t d, uint32_t r0,
                           uint32_t r1, uint32_t KS, int prec, void* out, float* scale,
                           double* norms);
hipError_t launch_norm_max(hipStream_t st, const double* norms, uint32_t r0, uint32_t r1,
                           unsigned long long* maxes);
// queries [0, nq): fragments (out), scales, error bounds E, the pass-1 aux block; flags[0..1] = 0
hipError_t launch_query_prep(hipStream_t st, const float* queries, uint32_t d, uint32_t nq, uint32_t KS, int prec,
                             void* out, float* qscale, const unsigned long long* maxes, float* E, float* aux,
                             uint32_t* flags, bool l2);
// hn[r] = |x_r|^2 / 2 (fp32) from the copy's FP64 norms, rows [r0, r1)
hipError_t launch_half_norms(hipStream_t st, const double* norms, uint32_t r0, uint32_t r1, float* hn);
hipError_t launch_zero_u32(hipStream_t st, uint32_t* p, size_t n);
hipError_t launch_mfma_scan(int prec, int pass, const ShortScanArgs& a);
hipError_t launch_scatter(int prec, const ShortScanArgs& a);
int short_waves();  // waves per shortlist scan block (emission regions per tile)
// thresholds of queries [0, nq) and the pass-2 aux block of queries [0, nq_pad) (padding: +inf)
hipError_t launch_select_threshold(hipStream_t st, const uint32_t* tilemin, uint32_t n_tiles, uint32_t nq,
                                   uint32_t nq_pad, const uint32_t* ks, uint32_t k, const float* E,
                                   const float* qscale, float* thr, float* aux);
hipError_t launch_rescore(hipStream_t st, bool l2, const float* base, uint32_t dp, uint32_t d,
                          const float* queries, uint32_t nq, const uint32_t* ks, uint32_t k,
                          const float* E, const uint64_t* cand, const uint32_t* count, uint32_t* ids,
                          float* dists, uint32_t ldo, uint32_t* flags, uint32_t* n_overflow,
                          uint32_t* rescored, uint32_t* sig, uint32_t* qmap, uint32_t split,
                          uint64_t* rpart, size_t rpart_keys);
// sig: nq, qmap: nq + 7 (sig, qmap null: queries in batch order). split: blocks per query of a
// batch of at most 256 queries (1 where split * ldo exceeds short_cand_cap(), rpart is null or
// its rpart_keys keys hold fewer than nq * split * ldo).
uint32_t short_rescore_split(uint32_t nq, uint32_t k, int num_cus);

// vsg_lists.hip: the filtered exact scan over per-query allowed-id lists.
struct ListsWorkspace {
  void* dev[11] = {};     // device buffers (queries, ks, items, qperm, slot offsets, ids, part, out,
                          // per-query thresholds, merge query map, queries in item order)
  size_t dev_cap[11] = {};
  void* host[3] = {};     // pinned staging (queries, ids out, dists out)
  size_t host_cap[3] = {};
  uint32_t* counter = nullptr;
  void release();
};

// Tile shape of the filtered scan (0 = chosen by the planner): queries per work item (qb),
// row groups per lane step (r), allowed rows per item (chunk), blocks per CU, lanes per row.
struct ListsTiles {
  int qb = 0, r = 0;
  uint64_t chunk = 0;
  int blocks_per_cu = 0;
  int lanes_per_row = 0;  // 4 (any d) or 8 (d = 128 / 384 / 768); 0 = 4
};
// Supported (qb, r, lanes per row) tiles of the filtered scan kernel for dimension dim.
bool lists_tile_supported(int qb, int r, int lpr, size_t dim);

// Event timing of a group of launches (implemented by the index when kernel timing is on).
struct TimingSink {
  virtual ~TimingSink() = default;
  virtual void begin(hipStream_t st) = 0;
  virtual void end(hipStream_t st, int kind, uint32_t launches) = 0;
};

int search_lists_host(ListsWorkspace& ws, const float* base, size_t n, size_t dim, size_t dp,
                      int metric, int num_cus, hipStream_t st, const float* queries, size_t nq,
                      const uint32_t* ks, size_t kmax, const uint64_t* range,
                      const uint32_t* allow_ids, size_t n_ids, uint32_t* out_ids,
                      float* out_dists, std::string& err, TimingSink* timing,
                      const ListsTiles& tiles, ListsTiles* used);

// Launch shape of a planned filtered batch (its plan, ks and allowed ids already in the
// workspace).
struct ListsLaunch {
  uint32_t nq = 0, kmax = 0, n_items = 0, grid = 0, dp = 0, d = 0;
  size_t lds = 0;
  int qb = 0, r = 0, lpr = 0, kp = 0;
  bool l2 = false;
  uint64_t chunk = 0;          // rows per work item
  uint32_t rows = 0, nch = 0;  // an unfiltered batch: base rows, chunks per query
};
// Plans a filtered batch of nq device-resident queries and uploads its plan into ws; returns
// once the uploads are complete. The batch's kernels are then recorded by lists_enqueue.
int prepare_lists_device(ListsWorkspace& ws, size_t n, size_t dim, size_t dp, int metric, int num_cus,
                         hipStream_t st, size_t nq, const uint32_t* ks, size_t kmax, const uint64_t* range,
                         const uint32_t* allow_ids, size_t n_ids, std::string& err, const ListsTiles& tiles,
                         ListsLaunch& out);
// Kernels only (no memset, copy or allocation): queries d_q (nq x d), results nq x kmax.
hipError_t lists_enqueue(ListsWorkspace& ws, const ListsLaunch& L, const float* base, const float* d_q,
                         uint32_t* d_ids, float* d_dists, hipStream_t st, TimingSink* timing);
// The exact answer of a device-chosen subset of up to nq unfiltered queries, k results each:
// the filtered scan planned at capture for all nq queries (prepare_lists_fallback); at run time
// a one-block kernel picks the queries i < rq with flags[i] != 0 (every one when *all != 0) and
// writes their work items, then the scan and the merge write those queries' rows of
// d_ids / d_dists (row stride ldo). Kernels only.
int prepare_lists_fallback(ListsWorkspace& ws, size_t n, size_t dim, size_t dp, int metric, int num_cus,
                           hipStream_t st, size_t nq, size_t k, std::string& err, ListsLaunch& out);
hipError_t lists_fallback_enqueue(ListsWorkspace& ws, const ListsLaunch& L, const float* base, const float* d_q,
                                  uint32_t rq, const uint32_t* flags, const uint32_t* all, uint32_t* d_ids,
                                  float* d_dists, uint32_t ldo, hipStream_t st);

}  // namespace vsg
