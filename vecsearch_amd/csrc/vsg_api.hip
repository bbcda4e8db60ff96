// The C ABI (include/vsg.h): index lifetime, resident base, workspace and search planning.
// Host side of the engine; the kernels live in vsg_scan.hip and vsg_lists.hip.
//
// Replaces GpuBruteForceIndex::Impl / ctor / add / search of the reference
// (src/gpu_brute_force.cu:344-488) with: a row-major padded base in HBM that grows by
// doubling (gpu_brute_force.cu:404-423 semantics: ids are assigned in add order), a grow-only
// device workspace, pinned host staging, and one HIP stream per index.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>
#include <mutex>
#include <shared_mutex>
#include <set>
#include <tuple>
#include <utility>
#include <new>
#include <string>
#include <vector>

#include "vsg.h"
#include "vsg_internal.h"

// The list space of vsg_search_device_noalloc: a fixed device array, one copy per GPU, held
// until the process ends. Host code never passes `scratch` itself to a kernel (on the host the
// name is only a stand-in for the symbol); it passes the device address that
// hipGetSymbolAddress returned for the current device (scratch_address).
static __device__ __align__(256) uint64_t scratch[VSG_NOALLOC_SCRATCH_BYTES / sizeof(uint64_t)];

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int fail_hip(const char* what, hipError_t e) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory ? VSG_ERR_OUT_OF_MEMORY : VSG_ERR_DEVICE;
}

#define VSG_HIP(expr)                                  \
  do {                                                 \
    hipError_t e_ = (expr);                            \
    if (e_ != hipSuccess) return fail_hip(#expr, e_);  \
  } while (0)

// Device address of `scratch` on each device, looked up once per device (with that device
// current) by the first vsg_create on it.
std::mutex g_scratch_mu;
std::map<int, uint64_t*> g_scratch_addr;

int scratch_setup(int device) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (g_scratch_addr.count(device)) return VSG_OK;
  void* p = nullptr;
  VSG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return fail(VSG_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned a null address");
  g_scratch_addr[device] = static_cast<uint64_t*>(p);
  return VSG_OK;
}

uint64_t* scratch_address(int device) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  auto it = g_scratch_addr.find(device);
  return it == g_scratch_addr.end() ? nullptr : it->second;
}

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t cap = 0;  // elements
  hipError_t reserve(size_t n) {
    if (n <= cap) return hipSuccess;
    if (p) {
      hipError_t e = hipFree(p);
      p = nullptr;
      cap = 0;
      if (e != hipSuccess) return e;
    }
    const size_t want = std::max(n, cap * 2);
    hipError_t e = hipMalloc(&p, want * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      return e;
    }
    cap = want;
    return hipSuccess;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

template <class T>
struct HostBuf {  // pinned host staging
  T* p = nullptr;
  size_t cap = 0;
  hipError_t reserve(size_t n) {
    if (n <= cap) return hipSuccess;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    const size_t want = std::max(n, cap * 2);
    hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&p), want * sizeof(T), hipHostMallocDefault);
    if (e != hipSuccess) {
      p = nullptr;
      return e;
    }
    cap = want;
    return hipSuccess;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
  }
};

constexpr size_t kPartBudgetBytes = size_t(256) << 20;  // per-round candidate lists

}  // namespace

namespace {

// Per-search scratch: device workspace, pinned staging and statistics. Each eager call leases
// one (a Lane, below) for its duration; each serving slot and each captured graph owns another,
// so searches in flight on different streams never share scratch memory.
struct Workspace {
  // shortlist
  DevBuf<uint8_t> s_qfrag;
  DevBuf<float> s_qscale, s_E, s_thr, s_qaux;
  DevBuf<uint32_t> s_tilemin, s_count, s_flags, s_rescored, s_nover, s_sig, s_qmap;
  DevBuf<uint64_t> s_rpart;  // rescore shares' first-k keys (small batches)
  DevBuf<uint64_t> s_cand;
  DevBuf<uint64_t> s_emit;
  DevBuf<uint32_t> s_emit_count, s_slice;
  HostBuf<uint32_t> s_hover;
  DevBuf<uint32_t> f_sel;  // overflowed queries of a round (exact fallback)
  DevBuf<float> f_q;
  DevBuf<uint32_t> f_ids;
  DevBuf<float> f_dists;
  uint64_t st_emitted = 0, st_rescored = 0, st_fallback = 0;
  size_t st_nq = 0;
  // exact scan and host staging
  DevBuf<float> q;
  DevBuf<uint64_t> part;
  DevBuf<uint32_t> ids;
  DevBuf<float> dists;
  DevBuf<uint8_t> mask;
  HostBuf<float> hq;
  HostBuf<uint32_t> hids;
  HostBuf<float> hdists;
  vsg::ListsWorkspace lists;
  vsg::ListsTiles lists_last;  // tile of the last filtered call

  void release() {
    s_qfrag.release();
    s_qscale.release();
    s_E.release();
    s_thr.release();
    s_qaux.release();
    s_tilemin.release();
    s_count.release();
    s_flags.release();
    s_rescored.release();
    s_nover.release();
    s_sig.release();
    s_qmap.release();
    s_rpart.release();
    s_cand.release();
    s_emit.release();
    s_emit_count.release();
    s_slice.release();
    s_hover.release();
    f_sel.release();
    f_q.release();
    f_ids.release();
    f_dists.release();
    q.release();
    part.release();
    ids.release();
    dists.release();
    mask.release();
    hq.release();
    hids.release();
    hdists.release();
    lists.release();
  }
};

// One eager search in flight: a stream, a workspace, and an event marking the end of the last
// work that used the workspace, on whichever stream it was enqueued (vsg_search_device runs on
// the caller's stream). The next user of the lane makes its own stream wait for that event
// before touching the workspace, so two searches never overlap on the same scratch memory.
struct Lane {
  hipStream_t st = nullptr;
  Workspace w;
  hipEvent_t done = nullptr;
  bool pending = false;  // `done` was recorded and the next user must wait for it
};

constexpr size_t kStageBytes = size_t(64) << 20;  // one pinned upload slice

}  // namespace

struct vsg_index {
  size_t dim = 0, dp = 0;
  int metric = VSG_METRIC_L2;
  int device = 0;
  int num_cus = 256;
  size_t n = 0, cap = 0;
  float* base = nullptr;   // cap x dp, row-major, zero padded (gathers: filters, rescoring)
  float* tiles = nullptr;  // the same rows in 16-row step-major tiles (streaming scans)
  int scan_waves_per_cu = 8;
  int auto_short_batch = 2;  // measured at 1M x 384: the shortlist is faster from 2 queries (and at 1)
  int short_sample_stride = 0;  // pass 1 visits every stride-th tile (0: chosen per call)
  vsg::ListsTiles lists_tiles;  // filtered scan tile shape (0 = planner's choice or autotuner)
  bool lists_autotune = true;
  std::map<std::tuple<size_t, int, int, int>, vsg::ListsTiles> lists_tuned;  // by shape class

  // low-precision copies for the shortlist (index 1 = bf16, 2 = fp8), built on first use
  struct Prec {
    void* frag = nullptr;             // fragment layout, cap rows
    float* scale = nullptr;           // fp8 per-row scale
    double* norms = nullptr;          // per row |x|^2, |x^|^2, |x - x^|^2
    float* hn = nullptr;              // L2: per row |x|^2 / 2 (fp32)
    unsigned long long* maxes = nullptr;  // max |x|, |x^|, |x - x^| (fp64 bits)
    size_t rows = 0, cap = 0;
  } prec[3];
  int ks = 0;  // 32-dim k-steps of the copies (0: shortlist unsupported for this dim)

  hipStream_t stream = nullptr;  // adds, low-precision copies, and vsg_search_device(NULL)

  // Eager searches lease a Lane each (at most max_lanes at once; a caller beyond that waits for
  // one to come back), so concurrent callers of the host API run side by side on their own
  // streams instead of queueing on one lock as the reference's per-index mutex makes them do
  // (gpu_brute_force.cu:351,430).
  std::mutex lanes_mu;
  std::condition_variable lanes_cv;
  std::vector<std::unique_ptr<Lane>> lanes;
  std::vector<Lane*> idle;
  int max_lanes = 4;
  int lanes_busy = 0, lanes_peak = 0;

  // What vsg_short_stats and the lists_last_* parameters report: the last call's figures.
  std::mutex last_mu;
  uint64_t last_emitted = 0, last_rescored = 0, last_fallback = 0;
  vsg::ListsTiles last_tiles;

  // pinned upload slices (vsg_add, vsg_add_hnsw_file): one is filled while the other's copy runs
  float* stage[2] = {nullptr, nullptr};
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  bool stage_busy[2] = {false, false};

  // Locking: searches hold `rw` shared and writers (add, reserve, set_param) hold it
  // exclusively, so searches run concurrently with each other but never with a change of the
  // base. `mu` orders the writers and the parameter / timing calls among themselves; prec_mu
  // guards the lazy low-precision copies, tune_mu the tile table, timing_mu the event list.
  std::mutex mu;
  std::shared_mutex rw;
  std::mutex prec_mu, tune_mu, timing_mu;

  // kernel timing (vsg_kernel_timing)
  struct Timed {
    hipEvent_t a, b;
    int kind;  // VSG_KERNEL_* (vsg.h)
    uint32_t launches;
  };
  std::atomic<bool> timing{false};
  std::vector<Timed> timed;

  ~vsg_index() {
    for (auto& t : timed) {
      (void)hipEventDestroy(t.a);
      (void)hipEventDestroy(t.b);
    }
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& l : lanes) {
      if (l->st) (void)hipStreamSynchronize(l->st);
      if (l->done) (void)hipEventSynchronize(l->done);
    }
    for (auto& l : lanes) {
      l->w.release();
      if (l->done) (void)hipEventDestroy(l->done);
      if (l->st) (void)hipStreamDestroy(l->st);
    }
    for (int i = 0; i < 2; ++i) {
      if (stage[i]) (void)hipHostFree(stage[i]);
      if (stage_ev[i]) (void)hipEventDestroy(stage_ev[i]);
    }
    if (base) (void)hipFree(base);
    if (tiles) (void)hipFree(tiles);
    for (auto& p : prec) {
      if (p.frag) (void)hipFree(p.frag);
      if (p.scale) (void)hipFree(p.scale);
      if (p.norms) (void)hipFree(p.norms);
      if (p.hn) (void)hipFree(p.hn);
      if (p.maxes) (void)hipFree(p.maxes);
    }
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// Capacity grows by doubling (gpu_brute_force.cu:410-415): new zeroed buffers, old rows copied.
// cap is a multiple of 16 so the tiled copy holds whole tiles.
int grow(vsg_index* h, size_t need) {
  if (need <= h->cap) return VSG_OK;
  size_t cap = std::max({need, h->cap * 2, size_t(1024)});
  cap = (cap + 15) / 16 * 16;
  const size_t bytes = cap * h->dp * sizeof(float);
  float* nb = nullptr;
  float* nt = nullptr;
  hipError_t e = hipMalloc(&nb, bytes);
  if (e == hipSuccess) e = hipMalloc(&nt, bytes);
  if (e == hipSuccess) e = hipMemsetAsync(nb, 0, bytes, h->stream);
  if (e == hipSuccess) e = hipMemsetAsync(nt, 0, bytes, h->stream);
  if (e == hipSuccess && h->n) {
    e = hipMemcpyAsync(nb, h->base, h->n * h->dp * sizeof(float), hipMemcpyDeviceToDevice, h->stream);
    const size_t tiled = (h->n + 15) / 16 * 16 * h->dp * sizeof(float);
    if (e == hipSuccess) e = hipMemcpyAsync(nt, h->tiles, tiled, hipMemcpyDeviceToDevice, h->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) {
    if (nb) (void)hipFree(nb);
    if (nt) (void)hipFree(nt);
    return fail_hip("grow", e);
  }
  if (h->base) VSG_HIP(hipFree(h->base));
  if (h->tiles) VSG_HIP(hipFree(h->tiles));
  h->base = nb;
  h->tiles = nt;
  h->cap = cap;
  return VSG_OK;
}

// Appends m host rows to the base (the caller holds rw exclusively and has grown the base).
// fill(dst, r0, c) writes rows r0 .. r0 + c of the source into dst, packed (c x dim floats);
// they go through two pinned 64 MB slices: while one slice is copied to HBM on the index
// stream, the next is filled on the host. Rows are padded from dim to dp floats in the slice.
template <class Fill>
int upload_rows(vsg_index* h, size_t m, Fill&& fill) {
  const size_t row_floats = h->dp;
  const size_t per = std::max<size_t>(1, kStageBytes / (row_floats * sizeof(float)));
  for (int i = 0; i < 2; ++i) {
    if (!h->stage[i])
      VSG_HIP(hipHostMalloc(reinterpret_cast<void**>(&h->stage[i]), per * row_floats * sizeof(float),
                            hipHostMallocDefault));
    if (!h->stage_ev[i]) VSG_HIP(hipEventCreateWithFlags(&h->stage_ev[i], hipEventDisableTiming));
  }
  int s = 0;
  for (size_t r0 = 0; r0 < m; r0 += per, s ^= 1) {
    const size_t c = std::min(per, m - r0);
    if (h->stage_busy[s]) {
      VSG_HIP(hipEventSynchronize(h->stage_ev[s]));
      h->stage_busy[s] = false;
    }
    float* buf = h->stage[s];
    int rc = fill(buf, r0, c);
    if (rc) return rc;
    if (h->dim != row_floats) {  // spread the packed rows out to dp floats, last row first
      for (size_t i = c; i-- > 0;) {
        std::memmove(buf + i * row_floats, buf + i * h->dim, h->dim * sizeof(float));
        std::memset(buf + i * row_floats + h->dim, 0, (row_floats - h->dim) * sizeof(float));
      }
    }
    VSG_HIP(hipMemcpyAsync(h->base + (h->n + r0) * row_floats, buf, c * row_floats * sizeof(float),
                           hipMemcpyHostToDevice, h->stream));
    VSG_HIP(hipEventRecord(h->stage_ev[s], h->stream));
    h->stage_busy[s] = true;
  }
  VSG_HIP(vsg::launch_tile_rows(h->stream, h->base, h->tiles, uint32_t(h->n), uint32_t(h->n + m),
                                uint32_t(h->dp)));
  VSG_HIP(hipStreamSynchronize(h->stream));
  h->stage_busy[0] = h->stage_busy[1] = false;
  h->n += m;
  return VSG_OK;
}

// Set while this thread records a HIP graph: no timing events go into a capture.
thread_local bool t_capturing = false;

bool timing_on(const vsg_index* h) { return h->timing && !t_capturing; }

// A lane held for one eager search. acquire() takes an idle lane (or makes a new one, up to
// max_lanes, or waits) and makes `st` wait for the lane's previous work; the destructor marks
// the lane's workspace busy until `st` reaches the end of this search's work and returns it.
class LaneLease {
 public:
  explicit LaneLease(vsg_index* h) : h_(h) {}
  LaneLease(const LaneLease&) = delete;
  LaneLease& operator=(const LaneLease&) = delete;
  ~LaneLease() {
    if (!lane_) return;
    if (st_ && hipEventRecord(lane_->done, st_) == hipSuccess) lane_->pending = true;
    std::lock_guard<std::mutex> g(h_->lanes_mu);
    h_->idle.push_back(lane_);
    h_->lanes_busy--;
    h_->lanes_cv.notify_one();
  }
  // st == nullptr: the lane's own stream.
  int acquire(hipStream_t st = nullptr) {
    {
      std::unique_lock<std::mutex> lk(h_->lanes_mu);
      for (;;) {
        if (!h_->idle.empty()) {
          lane_ = h_->idle.back();
          h_->idle.pop_back();
          break;
        }
        if (int(h_->lanes.size()) < h_->max_lanes) {
          auto l = std::make_unique<Lane>();
          hipError_t e = hipStreamCreateWithFlags(&l->st, hipStreamNonBlocking);
          if (e == hipSuccess) e = hipEventCreateWithFlags(&l->done, hipEventDisableTiming);
          if (e != hipSuccess) {
            if (l->st) (void)hipStreamDestroy(l->st);
            return fail_hip("search lane", e);
          }
          lane_ = l.get();
          h_->lanes.push_back(std::move(l));
          break;
        }
        h_->lanes_cv.wait(lk);
      }
      h_->lanes_busy++;
      h_->lanes_peak = std::max(h_->lanes_peak, h_->lanes_busy);
    }
    st_ = st ? st : lane_->st;
    if (lane_->pending) VSG_HIP(hipStreamWaitEvent(st_, lane_->done, 0));
    return VSG_OK;
  }
  Workspace& w() { return lane_->w; }
  hipStream_t stream() const { return st_; }

 private:
  vsg_index* h_;
  Lane* lane_ = nullptr;
  hipStream_t st_ = nullptr;
};

// Plan of one streaming scan over the base: n_blocks workgroups of 4 waves, each wave owning
// rows_per_wave consecutive rows (a multiple of 64); each block leaves one k-list per query
// (W = n_blocks lists).
struct ScanPlan {
  uint32_t n_blocks, rows_per_wave, W;
};

ScanPlan plan_scan(const vsg_index* h) {
  const size_t n = h->n;
  // scan_waves_per_cu waves per CU (8 measured best at 1M x 384), at least 64 rows per wave.
  const size_t max_waves =
      std::min<size_t>(size_t(h->num_cus) * h->scan_waves_per_cu, vsg::kMaxListsPerQuery);
  size_t waves = std::min(max_waves, std::max<size_t>(1, (n + 63) / 64));
  size_t blocks = (waves + 3) / 4;
  size_t rpw = (n + blocks * 4 - 1) / (blocks * 4);
  rpw = (rpw + 63) / 64 * 64;
  blocks = std::max<size_t>(1, (n + rpw * 4 - 1) / (rpw * 4));
  ScanPlan p;
  p.n_blocks = uint32_t(blocks);
  p.rows_per_wave = uint32_t(rpw);
  p.W = uint32_t(blocks);
  return p;
}

// Queries per streaming pass: one for the legacy "naive" method, else up to 8, and no more
// than fit the pass's queries in 64 KB of LDS.
int queries_per_pass(int method, size_t nq, size_t dp) {
  int qb = 8;
  if (method == VSG_METHOD_NAIVE || nq <= 1) qb = 1;
  else if (nq <= 2) qb = 2;
  else if (nq <= 4) qb = 4;
  while (qb > 1 && size_t(qb) * dp * sizeof(float) > 65536) qb /= 2;
  return qb;
}

// Event pair around a group of launches when kernel timing is on.
struct TimedScope {
  vsg_index* h;
  hipStream_t st;
  int kind;
  uint32_t launches = 0;
  hipEvent_t a = nullptr, b = nullptr;
  TimedScope(vsg_index* h_, hipStream_t st_, int kind_) : h(h_), st(st_), kind(kind_) {
    if (!timing_on(h)) return;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess ||
        hipEventRecord(a, st) != hipSuccess) {
      h->timing = false;  // give up timing rather than fail the search
    }
  }
  ~TimedScope() {
    if (!a || !b) return;
    if (timing_on(h) && hipEventRecord(b, st) == hipSuccess) {
      std::lock_guard<std::mutex> g(h->timing_mu);
      h->timed.push_back({a, b, kind, launches});
    } else {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  }
};

// TimingSink over the index's event list (for kernels launched outside this file).
struct IndexTiming : vsg::TimingSink {
  vsg_index* h;
  hipEvent_t a = nullptr;
  explicit IndexTiming(vsg_index* h_) : h(h_) {}
  void begin(hipStream_t st) override {
    a = nullptr;
    if (!timing_on(h)) return;
    if (hipEventCreate(&a) != hipSuccess || hipEventRecord(a, st) != hipSuccess) {
      if (a) (void)hipEventDestroy(a);
      a = nullptr;
      h->timing = false;
    }
  }
  void end(hipStream_t st, int kind, uint32_t launches) override {
    if (!a) return;
    hipEvent_t b = nullptr;
    if (timing_on(h) && hipEventCreate(&b) == hipSuccess && hipEventRecord(b, st) == hipSuccess) {
      std::lock_guard<std::mutex> g(h->timing_mu);
      h->timed.push_back({a, b, kind, launches});
    } else {
      (void)hipEventDestroy(a);
      if (b) (void)hipEventDestroy(b);
    }
    a = nullptr;
  }
};

// Enqueue an exact streaming search of nq device-resident queries on stream st, its per-block
// lists in `part` (part_keys keys: every pass writes all W x k keys of its queries, so nothing
// in it needs clearing). Queries go in rounds of as many passes as `part` holds.
int enqueue_exact_in(vsg_index* h, uint64_t* part, size_t part_keys, const float* d_q, size_t nq, size_t k,
                     int method, const uint8_t* d_mask, uint32_t* d_ids, float* d_dists, hipStream_t st) {
  const ScanPlan p = plan_scan(h);
  const int qb = queries_per_pass(method, nq, h->dp);
  const size_t per_query = size_t(p.W) * k;
  size_t round = part_keys / per_query / qb * qb;
  if (round == 0)
    return fail(VSG_ERR_INVALID_ARGUMENT, "exact scan: one pass needs " + std::to_string(qb * per_query * 8) +
                                              " bytes of list space, more than the " +
                                              std::to_string(part_keys * 8) + " available");
  round = std::min(round, (nq + qb - 1) / qb * qb);
  hipError_t e;
  const bool l2 = h->metric == VSG_METRIC_L2;
  for (size_t r0 = 0; r0 < nq; r0 += round) {
    const size_t rq = std::min(round, nq - r0);
    const float* qr = d_q + r0 * h->dim;
    {
      TimedScope ts(h, st, VSG_KERNEL_SCAN);
      for (size_t q0 = 0; q0 < rq; q0 += qb) {
        const uint32_t nqp = uint32_t(std::min<size_t>(qb, rq - q0));
        e = vsg::launch_scan(qb, l2, uint32_t(k), p.n_blocks, st, h->tiles, uint32_t(h->n),
                             uint32_t(h->dp), uint32_t(h->dim), qr, uint32_t(q0), nqp,
                             p.rows_per_wave, d_mask, part);
        if (e != hipSuccess) return fail_hip("scan kernel", e);
        ts.launches++;
      }
    }
    {
      TimedScope ts(h, st, VSG_KERNEL_MERGE);
      e = vsg::launch_merge(uint32_t(rq), part, nullptr, p.W, uint32_t(k), nullptr,
                            uint32_t(k), nullptr, d_ids + r0 * k, d_dists + r0 * k, uint32_t(k), st);
      if (e != hipSuccess) return fail_hip("merge kernel", e);
      ts.launches++;
    }
  }
  return VSG_OK;
}

vsg::ListsTiles lists_tiles_for(vsg_index* h, size_t nq, size_t kmax, const uint64_t* range);

// Unmasked exact batches of at least kLaneExactBatch queries run as a filtered batch whose every
// query is unfiltered, on the one-lane-per-row tile from 64 queries and the block-shared tile
// (32 queries an item) below, unless a filtered-scan tile is fixed or tuned for the class: the
// work items are chunk-major, so a chunk's rows are read from HBM once and from L2 by the
// chunk's other query blocks, where the streaming scan reads the whole base once per 8 queries
// (at 1M x 384, host calls: 24 queries 0.89 -> 0.74 ms, 32 queries 1.18 -> 0.78 ms, 64 queries
// 2.32 -> 1.47 ms, 256 queries 8.98 -> 4.86 ms, the same bits; at 16 queries the two are even).
// Planning uploads the plan and synchronises the stream, so a captured step keeps the streaming
// scan.
constexpr size_t kLaneExactBatch = 24;

int enqueue_exact_lanes(vsg_index* h, Workspace& w, const float* d_q, size_t nq, size_t k, uint32_t* d_ids,
                        float* d_dists, hipStream_t st) {
  for (size_t r0 = 0; r0 < nq; r0 += 4096) {
    const size_t rq = std::min<size_t>(4096, nq - r0);
    const std::vector<uint64_t> range(2 * rq, VSG_ALL_ROWS);
    const std::vector<uint32_t> ks(rq, uint32_t(k));
    vsg::ListsTiles t = lists_tiles_for(h, rq, k, range.data());
    if (t.qb == 0 && t.r == 0 && t.chunk == 0 && t.blocks_per_cu == 0 && t.lanes_per_row == 0) {
      if (rq < 64 && vsg::lists_tile_supported(32, 8, 8, h->dim)) {
        t.qb = 32;
        t.r = 8;
        t.chunk = 2048;
        t.lanes_per_row = 8;
      } else {
        t.qb = 8;
        t.r = 1;
        t.lanes_per_row = 1;
      }
    }
    vsg::ListsLaunch L;
    const int rc = vsg::prepare_lists_device(w.lists, h->n, h->dim, h->dp, h->metric, h->num_cus, st, rq, ks.data(),
                                             k, range.data(), nullptr, 0, g_err, t, L);
    if (rc) return rc;
    IndexTiming tm(h);
    VSG_HIP(vsg::lists_enqueue(w.lists, L, h->base, d_q + r0 * h->dim, d_ids + r0 * k, d_dists + r0 * k, st, &tm));
  }
  return VSG_OK;
}

// The same with the lists in the workspace (up to kPartBudgetBytes per round).
int enqueue_exact(vsg_index* h, Workspace& w, const float* d_q, size_t nq, size_t k, int method,
                  const uint8_t* d_mask, uint32_t* d_ids, float* d_dists, hipStream_t st) {
  if (d_mask == nullptr && nq >= kLaneExactBatch && method != VSG_METHOD_NAIVE && k >= 1 && k <= VSG_MAX_K &&
      h->n >= 65536 && vsg::lists_tile_supported(8, 1, 1, h->dim)) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone)
      return enqueue_exact_lanes(h, w, d_q, nq, k, d_ids, d_dists, st);
  }
  const ScanPlan p = plan_scan(h);
  const int qb = queries_per_pass(method, nq, h->dp);
  const size_t per_query = size_t(p.W) * k * sizeof(uint64_t);
  size_t round = std::max<size_t>(qb, kPartBudgetBytes / per_query / qb * qb);
  round = std::min(round, (nq + qb - 1) / qb * qb);
  const size_t keys = round * size_t(p.W) * k;
  hipError_t e = w.part.reserve(keys);
  if (e != hipSuccess) return fail_hip("workspace", e);
  return enqueue_exact_in(h, w.part.p, keys, d_q, nq, k, method, d_mask, d_ids, d_dists, st);
}

// ---- certified shortlist (vsg_short.hip) ---------------------------------------------------

bool short_supported(const vsg_index* h, size_t k) {
  return vsg::short_ks(uint32_t(h->dim)) > 0 && h->n >= 65536 && k <= VSG_MAX_K &&
         size_t(vsg::short_cand_cap()) >= 4 * k;
}

// Quantized copy `p` (1 bf16, 2 fp8) covering every row of the base. Built on the index
// stream, which is synchronised before the copy is used from any other stream.
int ensure_prec(vsg_index* h, int p) {
  std::lock_guard<std::mutex> g(h->prec_mu);
  auto& c = h->prec[p];
  if (c.frag && c.cap == h->cap && c.rows == h->n) return VSG_OK;
  const int ks = vsg::short_ks(uint32_t(h->dim));
  if (c.frag && (c.cap != h->cap || ks != h->ks)) {  // the base grew: rebuild from scratch
    (void)hipFree(c.frag);
    if (c.scale) (void)hipFree(c.scale);
    (void)hipFree(c.norms);
    if (c.hn) (void)hipFree(c.hn);
    (void)hipFree(c.maxes);
    c = vsg_index::Prec();
  }
  h->ks = ks;
  if (!c.frag) {
    const size_t chunk = p == 1 ? 1024 : 512;
    const size_t bytes = h->cap / 16 * size_t(ks) * chunk;
    VSG_HIP(hipMalloc(&c.frag, bytes));
    VSG_HIP(hipMemsetAsync(c.frag, 0, bytes, h->stream));
    if (p == 2) VSG_HIP(hipMalloc(&c.scale, h->cap * sizeof(float)));
    VSG_HIP(hipMalloc(&c.norms, h->cap * 3 * sizeof(double)));
    if (h->metric == VSG_METRIC_L2) VSG_HIP(hipMalloc(&c.hn, h->cap * sizeof(float)));
    VSG_HIP(hipMalloc(&c.maxes, 3 * sizeof(unsigned long long)));
    VSG_HIP(hipMemsetAsync(c.maxes, 0, 3 * sizeof(unsigned long long), h->stream));
    c.cap = h->cap;
    c.rows = 0;
  }
  if (c.rows < h->n) {
    VSG_HIP(vsg::launch_quantize(h->stream, h->base, uint32_t(h->dp), uint32_t(h->dim), uint32_t(c.rows),
                                 uint32_t(h->n), uint32_t(ks), p, c.frag, c.scale, c.norms));
    VSG_HIP(vsg::launch_norm_max(h->stream, c.norms, uint32_t(c.rows), uint32_t(h->n), c.maxes));
    if (c.hn) VSG_HIP(vsg::launch_half_norms(h->stream, c.norms, uint32_t(c.rows), uint32_t(h->n), c.hn));
    c.rows = h->n;
  }
  VSG_HIP(hipStreamSynchronize(h->stream));
  return VSG_OK;
}

int enqueue_search(vsg_index* h, Workspace& w, const float* d_q, size_t nq, size_t k, int method,
                   const uint8_t* d_mask, uint32_t* d_ids, float* d_dists, hipStream_t st);

// Launch shape of a shortlist search and its workspace (short_reserve).
struct ShortPlan {
  uint32_t cap = 0, tile_rows = 0, n_tiles = 0, wave_cap = 0;
  size_t round = 0, qpad = 0;
  bool qs = false;
  uint32_t qs_blocks = 0, qs_groups = 0, qs_rows = 0, qs_qpr = 0;
};

// Plans a shortlist search of nq queries and sizes the workspace for it (no allocation when it
// already fits, as inside a captured step).
int short_reserve(vsg_index* h, Workspace& w, size_t nq, size_t k, int p, ShortPlan& sp) {
  const int ks = h->ks;
  const uint32_t cap = uint32_t(vsg::short_cand_cap());
  const uint32_t tile_rows = uint32_t(vsg::short_tile_rows_for(p, ks));
  const uint32_t n_tiles = uint32_t((h->n + tile_rows - 1) / tile_rows);
  const size_t round = std::min<size_t>(nq, 4096);
  const size_t chunk = p == 1 ? 1024 : 512;
  const size_t grp = size_t(vsg::short_group(p, ks)) * 16;  // queries per stage of the MFMA scan
  const size_t qpad = (round + grp - 1) / grp * grp;
  hipError_t e;
  if ((e = w.s_qfrag.reserve(qpad / 16 * ks * chunk)) != hipSuccess) return fail_hip("workspace", e);
  if ((e = w.s_qscale.reserve(qpad)) != hipSuccess) return fail_hip("workspace", e);
  if ((e = w.s_E.reserve(round)) != hipSuccess) return fail_hip("workspace", e);
  if ((e = w.s_thr.reserve(round)) != hipSuccess) return fail_hip("workspace", e);
  if ((e = w.s_qaux.reserve(qpad / 16 * size_t(vsg::short_aux_floats()) + vsg::short_aux_pad())) != hipSuccess)
    return fail_hip("workspace", e);
  if ((e = w.s_tilemin.reserve(round * n_tiles)) != hipSuccess) return fail_hip("workspace", e);
  if ((e = w.s_count.reserve(round)) != hipSuccess) return fail_hip("workspace", e);
  if ((e = w.s_flags.reserve(round)) != hipSuccess) return fail_hip("workspace", e);
  if ((e = w.s_rescored.reserve(round)) != hipSuccess) return fail_hip("workspace", e);
  if ((e = w.s_nover.reserve(2)) != hipSuccess) return fail_hip("workspace", e);
  if ((e = w.s_sig.reserve(round)) != hipSuccess) return fail_hip("workspace", e);
  if ((e = w.s_qmap.reserve(round + 7)) != hipSuccess) return fail_hip("workspace", e);
  // the rescore shares of a small batch (a round of nq queries takes at most max(nq, CUs)
  // shares in all)
  if ((e = w.s_rpart.reserve(std::max<size_t>(round, size_t(std::max(1, h->num_cus))) * k)) != hipSuccess)
    return fail_hip("workspace", e);
  if ((e = w.s_cand.reserve(round * cap)) != hipSuccess) return fail_hip("workspace", e);
  // Pass 2 runs queries-stationary when the batch fits one block's registers (one round),
  // else rows-stationary over 512-row tiles; L2 always rows-stationary (the pass that carries
  // the rows' |x|^2 / 2 terms).
  uint32_t qs_blocks = 0, qs_groups = 0, qs_rows = 0, qs_qpr = 0;
  const bool qs = h->metric != VSG_METRIC_L2 &&
                  vsg::short_qs_plan(p, ks, uint32_t(h->n), uint32_t(nq), h->num_cus, &qs_blocks, &qs_groups,
                                     &qs_rows, &qs_qpr);
  // pass-2 emission regions, one per wave: a region of R rows and Q queries receives about
  // Q * (candidates per query) * R / n entries (rows are in no particular order); sized for
  // 4x that at 2,048 candidates per query and at least 1,024 (4,096 for the few large regions
  // of the queries-stationary pass), the most a region can receive being R * Q (an overflow
  // sends the round to the exact scan)
  const size_t region_rows = qs ? qs_rows : tile_rows / size_t(vsg::short_waves());
  const size_t region_q = qs ? qs_qpr : round;
  const size_t expect = region_q * 2048 * region_rows / std::max<size_t>(1, h->n);
  const uint32_t wave_cap =
      uint32_t(std::min(region_rows * region_q, std::max<size_t>(qs ? 4096 : 1024, 4 * expect)));
  const size_t n_regions = qs ? size_t(qs_blocks) * qs_groups * 4 : size_t(n_tiles) * vsg::short_waves();
  if ((e = w.s_emit.reserve(n_regions * wave_cap)) != hipSuccess) return fail_hip("workspace", e);
  if ((e = w.s_emit_count.reserve(n_regions)) != hipSuccess) return fail_hip("workspace", e);
  if ((e = w.s_slice.reserve(round * vsg::short_scatter_slices())) != hipSuccess) return fail_hip("workspace", e);
  if ((e = w.s_hover.reserve(2)) != hipSuccess) return fail_hip("pinned staging", e);
  sp.cap = cap;
  sp.tile_rows = tile_rows;
  sp.n_tiles = n_tiles;
  sp.wave_cap = wave_cap;
  sp.round = round;
  sp.qpad = qpad;
  sp.qs = qs;
  sp.qs_blocks = qs_blocks;
  sp.qs_groups = qs_groups;
  sp.qs_rows = qs_rows;
  sp.qs_qpr = qs_qpr;
  return VSG_OK;
}

// Shortlist search of nq device-resident queries (prec 1 bf16, 2 fp8). Without `dev_fallback`
// it synchronises the stream once per round to learn whether any query overflowed its
// candidate buffer (or an emission region overflowed); those queries are answered again by the
// exact scan. With it (a captured step), every round ends with the device-planned exact
// fallback of vsg::lists_fallback_enqueue instead, and nothing waits on the host.
int enqueue_short(vsg_index* h, Workspace& w, const float* d_q, size_t nq, size_t k, int p, uint32_t* d_ids,
                  float* d_dists, hipStream_t st, const vsg::ListsLaunch* dev_fallback = nullptr) {
  int rc = ensure_prec(h, p);
  if (rc) return rc;
  const auto& c = h->prec[p];
  const int ks = h->ks;
  ShortPlan sp;
  if ((rc = short_reserve(h, w, nq, k, p, sp))) return rc;
  const uint32_t cap = sp.cap, n_tiles = sp.n_tiles, wave_cap = sp.wave_cap;
  const size_t round = sp.round, qpad = sp.qpad;
  const bool qs = sp.qs;
  const uint32_t qs_blocks = sp.qs_blocks, qs_groups = sp.qs_groups, qs_rows = sp.qs_rows;
  hipError_t e;
  w.st_emitted = w.st_rescored = w.st_fallback = 0;
  w.st_nq = 0;

  std::vector<uint32_t> flags, counts, rescored;
  for (size_t r0 = 0; r0 < nq; r0 += round) {
    const uint32_t rq = uint32_t(std::min(round, nq - r0));
    const float* qr = d_q + r0 * h->dim;
    uint32_t* ids = d_ids + r0 * k;
    float* dists = d_dists + r0 * k;
    // flags: [0] queries whose candidates overflowed, [1] emission regions that overflowed
    VSG_HIP(vsg::launch_query_prep(st, qr, uint32_t(h->dim), rq, uint32_t(ks), p, w.s_qfrag.p, w.s_qscale.p,
                                   c.maxes, w.s_E.p, w.s_qaux.p, w.s_nover.p, h->metric == VSG_METRIC_L2));

    vsg::ShortScanArgs a;
    a.st = st;
    a.KS = ks;
    a.xfrag = c.frag;
    a.xscale = c.scale;
    a.xhn = c.hn;
    a.n = uint32_t(h->n);
    a.qfrag = w.s_qfrag.p;
    a.qaux = w.s_qaux.p;
    a.nq = rq;
    // pass 1 visits every tstride-th tile (a subset's k-th smallest minimum still bounds Dk;
    // a looser bound only grows the candidate set, checked against its capacity): at least 4k
    // tiles, at most every 16th. The query stages of a visited tile are shared by qsplit blocks
    // so that the pass, one block's time, fills about one block per CU: at 1M rows 123 tiles x 2
    // shares, where one tile per CU (245 x 1) was 2 % slower at batch 4,096 (its tile ran every
    // query) and 1.5-2.5 % slower at batches 8-256 (a tighter bound that saved less).
    const uint32_t tstride =
        h->short_sample_stride > 0
            ? uint32_t(std::max<size_t>(1, std::min<size_t>(size_t(h->short_sample_stride), n_tiles / std::max<size_t>(1, k))))
            : uint32_t(std::max<size_t>(1, std::min<size_t>(16, n_tiles / (4 * k))));
    const uint32_t space1 = (n_tiles + tstride - 1) / tstride;
    const uint32_t nst1 = uint32_t(((rq + 15) / 16 + size_t(vsg::short_group(p, ks)) - 1) / size_t(vsg::short_group(p, ks)));
    const uint32_t qsplit = std::max(1u, std::min(nst1, uint32_t(h->num_cus) / space1));
    a.tilemin = w.s_tilemin.p;
    a.emit = w.s_emit.p;
    a.wave_cap = wave_cap;
    a.emit_count = w.s_emit_count.p;
    a.overflow = w.s_nover.p + 1;
    a.cand = w.s_cand.p;
    a.cap = cap;
    a.count = w.s_count.p;
    a.slice = w.s_slice.p;
    {
      TimedScope ts(h, st, VSG_KERNEL_SAMPLE);
      a.tstride = tstride;
      a.n_space = space1;
      a.qsplit = qsplit;
      a.n_blocks = space1 * qsplit;
      VSG_HIP(vsg::launch_mfma_scan(p, 1, a));
      VSG_HIP(vsg::launch_select_threshold(st, w.s_tilemin.p, space1, rq, uint32_t(qpad), nullptr, uint32_t(k),
                                           w.s_E.p, w.s_qscale.p, w.s_thr.p, w.s_qaux.p));
      ts.launches++;
    }
    {
      TimedScope ts(h, st, VSG_KERNEL_MFMA);
      a.tstride = 1;
      a.qsplit = 1;
      a.n_space = a.n_blocks = n_tiles;
      if (qs) {
        a.qs_blocks = qs_blocks;
        a.qs_groups = qs_groups;
        a.qs_rows = qs_rows;
        VSG_HIP(vsg::launch_mfma_qs(p, a));
      } else {
        VSG_HIP(vsg::launch_mfma_scan(p, 2, a));
      }
      ts.launches++;
    }
    {
      TimedScope ts(h, st, VSG_KERNEL_SAMPLE);
      VSG_HIP(vsg::launch_scatter(p, a));
    }
    {
      TimedScope ts(h, st, VSG_KERNEL_RESCORE);
      VSG_HIP(vsg::launch_rescore(st, h->metric == VSG_METRIC_L2, h->base, uint32_t(h->dp), uint32_t(h->dim),
                                  qr, rq, nullptr, uint32_t(k), w.s_E.p, w.s_cand.p, w.s_count.p, ids,
                                  dists, uint32_t(k), w.s_flags.p, w.s_nover.p, w.s_rescored.p,
                                  p == 2 ? w.s_sig.p : nullptr, p == 2 ? w.s_qmap.p : nullptr,
                                  vsg::short_rescore_split(rq, uint32_t(k), h->num_cus), w.s_rpart.p,
                                  w.s_rpart.cap));
      ts.launches++;
    }
    if (dev_fallback) {
      VSG_HIP(vsg::lists_fallback_enqueue(w.lists, *dev_fallback, h->base, qr, rq, w.s_flags.p, w.s_nover.p + 1,
                                          ids, dists, uint32_t(k), st));
      w.st_nq += rq;
      continue;
    }
    VSG_HIP(hipMemcpyAsync(w.s_hover.p, w.s_nover.p, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    VSG_HIP(hipStreamSynchronize(st));
    if (w.s_hover.p[1]) {  // an emission region overflowed: answer this round exactly
      w.st_fallback += rq;
      rc = enqueue_exact(h, w, qr, rq, k, VSG_METHOD_SKINNY, nullptr, ids, dists, st);
      if (rc) return rc;
      w.st_nq += rq;
      continue;
    }
    const uint32_t nover = w.s_hover.p[0];
    if (timing_on(h) || nover) {
      counts.resize(rq);
      rescored.resize(rq);
      VSG_HIP(hipMemcpy(counts.data(), w.s_count.p, rq * sizeof(uint32_t), hipMemcpyDeviceToHost));
      VSG_HIP(hipMemcpy(rescored.data(), w.s_rescored.p, rq * sizeof(uint32_t), hipMemcpyDeviceToHost));
      for (uint32_t i = 0; i < rq; ++i) {
        w.st_emitted += counts[i];
        w.st_rescored += rescored[i];
      }
    }
    w.st_nq += rq;
    if (nover) {
      // overflowed queries: gathered into one block, answered by the exact scan together,
      // scattered back into this round's output rows
      flags.resize(rq);
      VSG_HIP(hipMemcpy(flags.data(), w.s_flags.p, rq * sizeof(uint32_t), hipMemcpyDeviceToHost));
      std::vector<uint32_t> sel;
      for (uint32_t i = 0; i < rq; ++i)
        if (flags[i]) sel.push_back(i);
      const uint32_t nf = uint32_t(sel.size());
      w.st_fallback += nf;
      if ((e = w.f_sel.reserve(nf)) != hipSuccess) return fail_hip("workspace", e);
      if ((e = w.f_q.reserve(size_t(nf) * h->dim)) != hipSuccess) return fail_hip("workspace", e);
      if ((e = w.f_ids.reserve(size_t(nf) * k)) != hipSuccess) return fail_hip("workspace", e);
      if ((e = w.f_dists.reserve(size_t(nf) * k)) != hipSuccess) return fail_hip("workspace", e);
      VSG_HIP(hipMemcpy(w.f_sel.p, sel.data(), nf * sizeof(uint32_t), hipMemcpyHostToDevice));
      VSG_HIP(vsg::launch_rows_move(st, qr, w.f_q.p, w.f_sel.p, nf, uint32_t(h->dim), false));
      rc = enqueue_exact(h, w, w.f_q.p, nf, k, VSG_METHOD_SKINNY, nullptr, w.f_ids.p, w.f_dists.p, st);
      if (rc) return rc;
      VSG_HIP(vsg::launch_rows_move(st, w.f_ids.p, ids, w.f_sel.p, nf, uint32_t(k), true));
      VSG_HIP(vsg::launch_rows_move(st, w.f_dists.p, dists, w.f_sel.p, nf, uint32_t(k), true));
      VSG_HIP(hipStreamSynchronize(st));  // sel / f_* are reused by the next round
    }
  }
  return VSG_OK;
}

int check_search_args(const vsg_index* h, size_t k, int method) {
  if (!h) return fail(VSG_ERR_INVALID_ARGUMENT, "null index");
  if (k > VSG_MAX_K)
    return fail(VSG_ERR_INVALID_ARGUMENT, "k must be <= " + std::to_string(VSG_MAX_K));
  if (method < VSG_METHOD_AUTO || method > VSG_METHOD_FP8)
    return fail(VSG_ERR_INVALID_ARGUMENT, "unknown search method");
  return VSG_OK;
}

// The method -> path choice (see VSG_METHOD_* in vsg.h).
int enqueue_search(vsg_index* h, Workspace& w, const float* d_q, size_t nq, size_t k, int method,
                   const uint8_t* d_mask, uint32_t* d_ids, float* d_dists, hipStream_t st) {
  if (d_mask == nullptr && short_supported(h, k)) {
    int p = 0;
    if (method == VSG_METHOD_TILED || method == VSG_METHOD_BF16) p = 1;
    else if (method == VSG_METHOD_FP8) p = 2;
    else if (method == VSG_METHOD_AUTO && nq >= size_t(h->auto_short_batch)) p = 1;
    if (p) {
      const int rc = enqueue_short(h, w, d_q, nq, k, p, d_ids, d_dists, st);
      if (rc == VSG_OK) {
        std::lock_guard<std::mutex> g(h->last_mu);
        h->last_emitted = w.st_emitted;
        h->last_rescored = w.st_rescored;
        h->last_fallback = w.st_fallback;
      }
      return rc;
    }
  }
  return enqueue_exact(h, w, d_q, nq, k, method, d_mask, d_ids, d_dists, st);
}

// The tile of the last filtered call, for the lists_last_* parameters.
void note_lists(vsg_index* h, const Workspace& w) {
  std::lock_guard<std::mutex> g(h->last_mu);
  h->last_tiles = w.lists_last;
}

void fill_missing(uint32_t* ids, float* dists, size_t count) {
  std::fill(ids, ids + count, VSG_NONE);
  std::fill(dists, dists + count, INFINITY);
}

// ---- tile autotuner of the filtered scan ---------------------------------------------------
// A call's shape class is (dim, k > 64, log2 queries per distinct range, log2 rows per distinct
// range). The first call of a class that is large enough to time (>= 2^24 query-row pairs) runs
// once per candidate tile, each timed on the scan kernel alone, and the class keeps the fastest.
// Every candidate returns the same bits, so the outputs of that call are valid whichever ran
// last.

// Scan-kernel stopwatch for the tuner (search_lists_host synchronises before returning).
struct ScanStopwatch : vsg::TimingSink {
  hipEvent_t pending = nullptr, a = nullptr, b = nullptr;
  ~ScanStopwatch() override {
    for (hipEvent_t e : {pending, a, b})
      if (e) (void)hipEventDestroy(e);
  }
  void begin(hipStream_t st) override {
    if (!pending && hipEventCreate(&pending) != hipSuccess) pending = nullptr;
    if (pending) (void)hipEventRecord(pending, st);
  }
  void end(hipStream_t st, int kind, uint32_t) override {
    if (kind != VSG_KERNEL_LISTS || !pending) return;
    if (!b && hipEventCreate(&b) != hipSuccess) b = nullptr;
    if (!b) return;
    std::swap(a, pending);
    (void)hipEventRecord(b, st);
  }
  float ms() {
    float v = -1.f;
    if (a && b && hipEventSynchronize(b) == hipSuccess && hipEventElapsedTime(&v, a, b) == hipSuccess) return v;
    return -1.f;
  }
};

int ilog2_floor(uint64_t v) {
  int r = 0;
  while (v > 1) {
    v >>= 1;
    ++r;
  }
  return r;
}

// Shape class of a filtered batch in the tuned-tile table (dim, k > 64, queries per distinct
// allowed set, rows per distinct allowed set, both log2), and its allowed (query, row) pairs.
struct ListsClass {
  std::tuple<size_t, int, int, int> key;
  uint64_t pairs = 0;
};
ListsClass lists_class(const vsg_index* h, size_t nq, size_t kmax, const uint64_t* range) {
  std::set<std::pair<uint64_t, uint64_t>> distinct;
  uint64_t pairs = 0, rows = 0;
  for (size_t i = 0; i < nq; ++i) {
    const bool all = range[2 * i] == VSG_ALL_ROWS;
    const uint64_t len = all ? h->n : (range[2 * i + 1] >= range[2 * i] ? range[2 * i + 1] - range[2 * i] : 0);
    pairs += len;
    const std::pair<uint64_t, uint64_t> rk(all ? uint64_t(VSG_ALL_ROWS) : range[2 * i],
                                           all ? uint64_t(0) : range[2 * i + 1]);
    if (distinct.insert(rk).second) rows += len;
  }
  const size_t nd = std::max<size_t>(1, distinct.size());
  ListsClass c;
  c.key = std::make_tuple(h->dim, kmax > 64 ? 1 : 0, ilog2_floor(nq / nd + 1), ilog2_floor(rows / nd + 1));
  c.pairs = pairs;
  return c;
}

bool lookup_tuned(vsg_index* h, const std::tuple<size_t, int, int, int>& key, vsg::ListsTiles* out) {
  std::lock_guard<std::mutex> g(h->tune_mu);
  auto it = h->lists_tuned.find(key);
  if (it == h->lists_tuned.end()) return false;
  *out = it->second;
  return true;
}

// The tile an eager call of this batch runs with when it makes no candidate runs: the fixed
// tile, else the tuned tile of its shape class, else the planner's.
vsg::ListsTiles lists_tiles_for(vsg_index* h, size_t nq, size_t kmax, const uint64_t* range) {
  const vsg::ListsTiles& fixed = h->lists_tiles;
  const bool planner = fixed.qb == 0 && fixed.r == 0 && fixed.chunk == 0 && fixed.blocks_per_cu == 0 &&
                       fixed.lanes_per_row == 0;
  vsg::ListsTiles t = fixed;
  if (planner && h->lists_autotune) (void)lookup_tuned(h, lists_class(h, nq, kmax, range).key, &t);
  return t;
}

// Filtered / mixed batch through the persistent lists scan, with the tile autotuner; host
// buffers in and out, on stream st with workspace w. The caller holds h->rw shared.
// allow_tune = false (serving): an untuned class runs the planner's tile instead of pausing the
// batch for the candidate runs.
int run_lists(vsg_index* h, Workspace& w, hipStream_t st, const float* queries, size_t nq,
              const uint32_t* ks, size_t kmax, const uint64_t* range, const uint32_t* allow_ids,
              size_t n_ids, uint32_t* out_ids, float* out_dists, bool allow_tune = true) {
  IndexTiming tm(h);
  const vsg::ListsTiles& fixed = h->lists_tiles;
  const bool planner = fixed.qb == 0 && fixed.r == 0 && fixed.chunk == 0 && fixed.blocks_per_cu == 0 &&
                       fixed.lanes_per_row == 0;
  if (planner && h->lists_autotune) {
    const ListsClass cls = lists_class(h, nq, kmax, range);
    const auto& key = cls.key;
    const uint64_t pairs = cls.pairs;
    vsg::ListsTiles tuned;
    const bool have = lookup_tuned(h, key, &tuned);
    if (have)
      return vsg::search_lists_host(w.lists, h->base, h->n, h->dim, h->dp, h->metric, h->num_cus, st,
                                    queries, nq, ks, kmax, range, allow_ids, n_ids, out_ids, out_dists,
                                    g_err, &tm, tuned, &w.lists_last);
    if (allow_tune && pairs >= (uint64_t(1) << 24)) {
      // {queries per wave, row groups per lane step, chunk rows (0: planner), lanes per row}
      // (32 queries: one item per block, every row read shared by its four waves; one lane per
      // row: eight queries per wave through scalar loads)
      static const int cand[][4] = {{4, 4, 0, 4},     {4, 4, 1024, 4},  {4, 4, 4096, 4}, {8, 2, 1024, 4},
                                    {2, 4, 0, 4},     {8, 4, 1024, 8},  {8, 4, 4096, 8}, {8, 2, 2048, 8},
                                    {16, 2, 2048, 8}, {16, 2, 4096, 8}, {32, 8, 0, 8},   {32, 8, 2048, 8},
                                    {32, 8, 8192, 8}, {8, 1, 0, 1},     {8, 1, 1024, 1}, {8, 1, 2048, 1},
                                    {8, 1, 4096, 1},  {8, 1, 8192, 1}};
      vsg::ListsTiles best;
      float best_ms = -1.f;
      for (const auto& c : cand) {
        if (!vsg::lists_tile_supported(c[0], c[1], c[3], h->dim)) continue;
        vsg::ListsTiles t;
        t.qb = c[0];
        t.r = c[1];
        t.chunk = uint64_t(c[2]);
        t.lanes_per_row = c[3];
        ScanStopwatch sw;
        vsg::ListsTiles used;
        const int rc = vsg::search_lists_host(w.lists, h->base, h->n, h->dim, h->dp, h->metric,
                                              h->num_cus, st, queries, nq, ks, kmax, range,
                                              allow_ids, n_ids, out_ids, out_dists, g_err, &sw, t, &used);
        if (rc) return rc;
        const float ms = sw.ms();
        if (ms >= 0.f && (best_ms < 0.f || ms < best_ms)) {
          best_ms = ms;
          best = t;
          w.lists_last = used;
        }
      }
      if (best_ms >= 0.f) {
        std::lock_guard<std::mutex> g(h->tune_mu);
        h->lists_tuned[key] = best;
      }
      return VSG_OK;
    }
  }
  return vsg::search_lists_host(w.lists, h->base, h->n, h->dim, h->dp, h->metric, h->num_cus, st,
                                queries, nq, ks, kmax, range, allow_ids, n_ids, out_ids, out_dists,
                                g_err, &tm, fixed, &w.lists_last);
}

}  // namespace

extern "C" {

const char* vsg_last_error(void) { return g_err.c_str(); }

int vsg_device_count(int* count) {
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // no driver: clear the error, report zero devices
    c = 0;
  }
  if (count) *count = c;
  return VSG_OK;
}

int vsg_create(size_t dim, int metric, int device, vsg_index** out) {
  if (!out) return fail(VSG_ERR_INVALID_ARGUMENT, "null output handle");
  *out = nullptr;
  if (dim == 0) return fail(VSG_ERR_INVALID_ARGUMENT, "dim must be > 0");
  if (dim > 16384) return fail(VSG_ERR_INVALID_ARGUMENT, "dim must be <= 16384");
  if (metric != VSG_METRIC_L2 && metric != VSG_METRIC_IP)
    return fail(VSG_ERR_INVALID_ARGUMENT, "metric must be L2 (0) or inner product (1)");
  int count = 0;
  vsg_device_count(&count);
  if (count == 0) return fail(VSG_ERR_NO_DEVICE, "no HIP device is visible");
  if (device < 0 || device >= count)
    return fail(VSG_ERR_INVALID_ARGUMENT, "device index out of range");
  VSG_HIP(hipSetDevice(device));
  int rc = scratch_setup(device);
  if (rc) return rc;
  vsg_index* h = new (std::nothrow) vsg_index;
  if (!h) return fail(VSG_ERR_OUT_OF_MEMORY, "host allocation failed");
  h->dim = dim;
  h->dp = (dim + 15) / 16 * 16;
  h->metric = metric;
  h->device = device;
  hipDeviceProp_t prop;
  hipError_t e = hipGetDeviceProperties(&prop, device);
  if (e == hipSuccess) {
    h->num_cus = std::max(1, prop.multiProcessorCount);
    e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  }
  if (e != hipSuccess) {
    delete h;
    return fail_hip("vsg_create", e);
  }
  *out = h;
  return VSG_OK;
}

void vsg_destroy(vsg_index* h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  delete h;
}

int vsg_size(const vsg_index* h, size_t* n) {
  if (!h || !n) return fail(VSG_ERR_INVALID_ARGUMENT, "null argument");
  *n = h->n;
  return VSG_OK;
}

int vsg_dim(const vsg_index* h, size_t* dim) {
  if (!h || !dim) return fail(VSG_ERR_INVALID_ARGUMENT, "null argument");
  *dim = h->dim;
  return VSG_OK;
}

int vsg_metric(const vsg_index* h, int* metric) {
  if (!h || !metric) return fail(VSG_ERR_INVALID_ARGUMENT, "null argument");
  *metric = h->metric;
  return VSG_OK;
}

int vsg_reserve(vsg_index* h, size_t n) {
  if (!h) return fail(VSG_ERR_INVALID_ARGUMENT, "null index");
  std::lock_guard<std::mutex> g(h->mu);
  std::unique_lock<std::shared_mutex> wl(h->rw);
  if (n >= VSG_NONE) return fail(VSG_ERR_LENGTH, "an index holds at most 2^32 - 2 vectors");
  VSG_HIP(hipSetDevice(h->device));
  return grow(h, n);
}

int vsg_add(vsg_index* h, const float* data, size_t m) {
  if (!h) return fail(VSG_ERR_INVALID_ARGUMENT, "null index");
  std::lock_guard<std::mutex> g(h->mu);
  std::unique_lock<std::shared_mutex> wl(h->rw);
  if (m == 0) return VSG_OK;
  if (!data) return fail(VSG_ERR_INVALID_ARGUMENT, "null data");
  if (h->n + m >= VSG_NONE) return fail(VSG_ERR_LENGTH, "an index holds at most 2^32 - 2 vectors");
  VSG_HIP(hipSetDevice(h->device));
  int rc = grow(h, h->n + m);
  if (rc) return rc;
  return upload_rows(h, m, [&](float* dst, size_t r0, size_t c) {
    std::memcpy(dst, data + r0 * h->dim, c * h->dim * sizeof(float));
    return VSG_OK;
  });
}

int vsg_search(vsg_index* h, const float* queries, size_t nq, size_t k, int method,
               const uint8_t* mask, uint32_t* out_ids, float* out_dists) {
  int rc = check_search_args(h, k, method);
  if (rc) return rc;
  if (nq == 0 || k == 0) return VSG_OK;
  if (!queries || !out_ids || !out_dists) return fail(VSG_ERR_INVALID_ARGUMENT, "null buffer");
  std::shared_lock<std::shared_mutex> rl(h->rw);
  if (h->n == 0) {
    fill_missing(out_ids, out_dists, nq * k);
    return VSG_OK;
  }
  VSG_HIP(hipSetDevice(h->device));
  size_t allowed = h->n;
  if (mask) {
    allowed = size_t(std::count_if(mask, mask + h->n, [](uint8_t v) { return v != 0; }));
    if (allowed == 0) {
      fill_missing(out_ids, out_dists, nq * k);  // hnsw.cpp:326: no allowed row, no result
      return VSG_OK;
    }
  }
  LaneLease lease(h);
  if ((rc = lease.acquire())) return rc;
  Workspace& w = lease.w();
  const hipStream_t st = lease.stream();
  // A sparse filter, or a dense one over a batch of 256 queries and more (d a multiple of 32):
  // the allowed rows as one id list that every query shares, through the filtered scan (a
  // sparse one gathers the exact_below regime's few rows instead of streaming the whole base;
  // a large batch reads each chunk of allowed rows from HBM once for all its query blocks, where
  // the masked streaming scan reads the base once per 8 queries: at 1M x 384 with half the rows
  // allowed, 256 queries in 6.3 ms a call instead of 9.4 ms as two calls of 128, 1,024 in 12.6
  // instead of 37.5 ms; at 128 queries the host's id list and plan cancel the gain).
  const bool dense_batch = mask && allowed * 4 >= h->n && nq >= 256 && vsg::lists_tile_supported(8, 1, 1, h->dim);
  if (mask && (allowed * 4 < h->n || dense_batch)) {
    std::vector<uint32_t> ids;
    ids.reserve(allowed);
    for (size_t i = 0; i < h->n; ++i)
      if (mask[i]) ids.push_back(uint32_t(i));
    std::vector<uint64_t> range(2 * nq);
    for (size_t i = 0; i < nq; ++i) {
      range[2 * i] = 0;
      range[2 * i + 1] = allowed;
    }
    std::vector<uint32_t> ks(nq, uint32_t(k));
    if (dense_batch) {  // the tile autotuner's choice for the class, as for a filtered batch
      rc = run_lists(h, w, st, queries, nq, ks.data(), k, range.data(), ids.data(), ids.size(), out_ids, out_dists);
    } else {
      IndexTiming tm(h);
      rc = vsg::search_lists_host(w.lists, h->base, h->n, h->dim, h->dp, h->metric, h->num_cus, st, queries,
                                  nq, ks.data(), k, range.data(), ids.data(), ids.size(), out_ids, out_dists,
                                  g_err, &tm, h->lists_tiles, &w.lists_last);
    }
    if (rc == VSG_OK) note_lists(h, w);
    return rc;
  }
  hipError_t e;
  if ((e = w.q.reserve(nq * h->dim)) != hipSuccess) return fail_hip("workspace", e);
  if ((e = w.ids.reserve(nq * k)) != hipSuccess) return fail_hip("workspace", e);
  if ((e = w.dists.reserve(nq * k)) != hipSuccess) return fail_hip("workspace", e);
  if ((e = w.hq.reserve(nq * h->dim)) != hipSuccess) return fail_hip("pinned staging", e);
  if ((e = w.hids.reserve(nq * k)) != hipSuccess) return fail_hip("pinned staging", e);
  if ((e = w.hdists.reserve(nq * k)) != hipSuccess) return fail_hip("pinned staging", e);
  const uint8_t* d_mask = nullptr;
  if (mask) {
    if ((e = w.mask.reserve(h->n)) != hipSuccess) return fail_hip("workspace", e);
    VSG_HIP(hipMemcpyAsync(w.mask.p, mask, h->n, hipMemcpyHostToDevice, st));
    d_mask = w.mask.p;
  }
  std::memcpy(w.hq.p, queries, nq * h->dim * sizeof(float));
  VSG_HIP(hipMemcpyAsync(w.q.p, w.hq.p, nq * h->dim * sizeof(float), hipMemcpyHostToDevice, st));
  rc = enqueue_search(h, w, w.q.p, nq, k, method, d_mask, w.ids.p, w.dists.p, st);
  if (rc) return rc;
  VSG_HIP(hipMemcpyAsync(w.hids.p, w.ids.p, nq * k * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  VSG_HIP(hipMemcpyAsync(w.hdists.p, w.dists.p, nq * k * sizeof(float), hipMemcpyDeviceToHost, st));
  VSG_HIP(hipStreamSynchronize(st));
  std::memcpy(out_ids, w.hids.p, nq * k * sizeof(uint32_t));
  std::memcpy(out_dists, w.hdists.p, nq * k * sizeof(float));
  return VSG_OK;
}

int vsg_search_device(vsg_index* h, const float* d_queries, size_t nq, size_t k, int method,
                      const uint8_t* d_mask, uint32_t* d_ids, float* d_dists, void* stream) {
  int rc = check_search_args(h, k, method);
  if (rc) return rc;
  if (nq == 0 || k == 0) return VSG_OK;
  if (!d_queries || !d_ids || !d_dists) return fail(VSG_ERR_INVALID_ARGUMENT, "null buffer");
  std::shared_lock<std::shared_mutex> rl(h->rw);
  VSG_HIP(hipSetDevice(h->device));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : h->stream;
  if (h->n == 0) {
    VSG_HIP(vsg::launch_fill_missing(st, d_ids, d_dists, nq * k));
    return VSG_OK;
  }
  // The lane's workspace is ordered on `st` both ways: `st` waits for the lane's previous
  // work, and the lane is marked busy until `st` has finished this search.
  LaneLease lease(h);
  if ((rc = lease.acquire(st))) return rc;
  return enqueue_search(h, lease.w(), d_queries, nq, k, method, d_mask, d_ids, d_dists, st);
}

int vsg_search_device_noalloc(vsg_index* h, const float* d_queries, size_t nq, size_t k, int method,
                              const uint8_t* d_mask, uint32_t* d_ids, float* d_dists, void* stream) {
  int rc = check_search_args(h, k, method);
  if (rc) return rc;
  if (nq == 0 || k == 0) return VSG_OK;
  if (!d_queries || !d_ids || !d_dists) return fail(VSG_ERR_INVALID_ARGUMENT, "null buffer");
  std::shared_lock<std::shared_mutex> rl(h->rw);
  VSG_HIP(hipSetDevice(h->device));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : h->stream;
  if (h->n == 0) {
    VSG_HIP(vsg::launch_fill_missing(st, d_ids, d_dists, nq * k));
    return VSG_OK;
  }
  uint64_t* part = scratch_address(h->device);
  if (!part) return fail(VSG_ERR_DEVICE, "no scratch address for this device");
  // every method has the exact streaming scan's results; NAIVE keeps one query per pass
  const int m = method == VSG_METHOD_NAIVE ? VSG_METHOD_NAIVE : VSG_METHOD_SKINNY;
  const bool timing = t_capturing;
  t_capturing = true;  // no timing events (creating one may allocate)
  rc = enqueue_exact_in(h, part, VSG_NOALLOC_SCRATCH_BYTES / sizeof(uint64_t), d_queries, nq, k, m, d_mask,
                        d_ids, d_dists, st);
  t_capturing = timing;
  return rc;
}

int vsg_search_lists(vsg_index* h, const float* queries, size_t nq, const uint32_t* ks,
                     size_t kmax, const uint64_t* range, const uint32_t* allow_ids, size_t n_ids,
                     uint32_t* out_ids, float* out_dists) {
  if (!h) return fail(VSG_ERR_INVALID_ARGUMENT, "null index");
  if (kmax > VSG_MAX_K)
    return fail(VSG_ERR_INVALID_ARGUMENT, "k must be <= " + std::to_string(VSG_MAX_K));
  if (nq == 0 || kmax == 0) return VSG_OK;
  if (!queries || !ks || !range || !out_ids || !out_dists)
    return fail(VSG_ERR_INVALID_ARGUMENT, "null buffer");
  std::shared_lock<std::shared_mutex> rl(h->rw);
  VSG_HIP(hipSetDevice(h->device));
  LaneLease lease(h);
  int rc = lease.acquire();
  if (rc) return rc;
  rc = run_lists(h, lease.w(), lease.stream(), queries, nq, ks, kmax, range, allow_ids, n_ids, out_ids,
                 out_dists);
  if (rc == VSG_OK) note_lists(h, lease.w());
  return rc;
}

int vsg_get_param(vsg_index* h, const char* name, int64_t* value) {
  if (!h || !name || !value) return fail(VSG_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> g(h->mu);
  const std::string key(name);
  if (key == "scan_waves_per_cu") *value = h->scan_waves_per_cu;
  else if (key == "auto_short_batch") *value = h->auto_short_batch;
  else if (key == "short_sample_stride") *value = h->short_sample_stride;
  else if (key == "lists_qb") *value = h->lists_tiles.qb;
  else if (key == "lists_r") *value = h->lists_tiles.r;
  else if (key == "lists_chunk") *value = int64_t(h->lists_tiles.chunk);
  else if (key == "lists_blocks_per_cu") *value = h->lists_tiles.blocks_per_cu;
  else if (key == "lists_autotune") *value = h->lists_autotune ? 1 : 0;
  else if (key == "lists_tuned_classes") {
    std::lock_guard<std::mutex> t(h->tune_mu);
    *value = int64_t(h->lists_tuned.size());
  } else if (key.rfind("lists_last_", 0) == 0 || key == "search_lanes_peak" || key == "search_lanes") {
    std::lock_guard<std::mutex> l(h->last_mu);
    std::lock_guard<std::mutex> p(h->lanes_mu);
    if (key == "lists_last_qb") *value = h->last_tiles.qb;
    else if (key == "lists_last_r") *value = h->last_tiles.r;
    else if (key == "lists_last_chunk") *value = int64_t(h->last_tiles.chunk);
    else if (key == "lists_last_lanes_per_row") *value = h->last_tiles.lanes_per_row;
    else if (key == "search_lanes_peak") *value = h->lanes_peak;
    else if (key == "search_lanes") *value = h->max_lanes;
    else return fail(VSG_ERR_INVALID_ARGUMENT, "unknown parameter: " + key);
  } else if (key == "lists_lanes_per_row") *value = h->lists_tiles.lanes_per_row;
  else return fail(VSG_ERR_INVALID_ARGUMENT, "unknown parameter: " + key);
  return VSG_OK;
}

int vsg_set_param(vsg_index* h, const char* name, int64_t value) {
  if (!h || !name) return fail(VSG_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> g(h->mu);
  std::unique_lock<std::shared_mutex> wl(h->rw);
  const std::string key(name);
  if (key == "scan_waves_per_cu") {
    if (value < 1 || value > 32) return fail(VSG_ERR_INVALID_ARGUMENT, "scan_waves_per_cu must be in 1..32");
    h->scan_waves_per_cu = int(value);
    return VSG_OK;
  }
  if (key == "short_sample_stride") {
    if (value < 0 || value > 256) return fail(VSG_ERR_INVALID_ARGUMENT, "short_sample_stride must be in 0..256");
    h->short_sample_stride = int(value);
    return VSG_OK;
  }
  if (key == "auto_short_batch") {
    if (value < 1) return fail(VSG_ERR_INVALID_ARGUMENT, "auto_short_batch must be >= 1");
    h->auto_short_batch = int(value);
    return VSG_OK;
  }
  if (key == "lists_qb") {
    if (value != 0 && value != 1 && value != 2 && value != 4 && value != 8 && value != 16 && value != 32)
      return fail(VSG_ERR_INVALID_ARGUMENT, "lists_qb must be 0 (auto), 1, 2, 4, 8, 16 or 32");
    h->lists_tiles.qb = int(value);
    return VSG_OK;
  }
  if (key == "lists_r") {
    if (value != 0 && value != 1 && value != 2 && value != 4 && value != 8)  // 1: the one-lane-per-row tile
      return fail(VSG_ERR_INVALID_ARGUMENT, "lists_r must be 0 (auto), 1, 2, 4 or 8");
    h->lists_tiles.r = int(value);
    return VSG_OK;
  }
  if (key == "lists_lanes_per_row") {
    if (value != 0 && value != 1 && value != 4 && value != 8)
      return fail(VSG_ERR_INVALID_ARGUMENT, "lists_lanes_per_row must be 0 (auto), 1, 4 or 8");
    h->lists_tiles.lanes_per_row = int(value);
    return VSG_OK;
  }
  if (key == "lists_chunk") {
    if (value < 0 || value > (int64_t(1) << 24))
      return fail(VSG_ERR_INVALID_ARGUMENT, "lists_chunk must be 0 (auto) or at most 2^24 rows");
    h->lists_tiles.chunk = uint64_t(value);
    return VSG_OK;
  }
  if (key == "lists_autotune") {
    if (value != 0 && value != 1) return fail(VSG_ERR_INVALID_ARGUMENT, "lists_autotune must be 0 or 1");
    h->lists_autotune = value != 0;
    if (!h->lists_autotune) {
      std::lock_guard<std::mutex> t(h->tune_mu);
      h->lists_tuned.clear();
    }
    return VSG_OK;
  }
  if (key == "lists_blocks_per_cu") {
    if (value < 0 || value > 16) return fail(VSG_ERR_INVALID_ARGUMENT, "lists_blocks_per_cu must be in 0..16");
    h->lists_tiles.blocks_per_cu = int(value);
    return VSG_OK;
  }
  return fail(VSG_ERR_INVALID_ARGUMENT, "unknown parameter: " + key);
}

int vsg_kernel_timing(vsg_index* h, int enable) {
  if (!h) return fail(VSG_ERR_INVALID_ARGUMENT, "null index");
  std::lock_guard<std::mutex> g(h->mu);
  h->timing = enable != 0;
  return VSG_OK;
}

int vsg_kernel_times(vsg_index* h, double* ms_out, uint64_t* launches_out) {
  if (!h) return fail(VSG_ERR_INVALID_ARGUMENT, "null index");
  std::lock_guard<std::mutex> g(h->mu);
  VSG_HIP(hipSetDevice(h->device));
  double ms[VSG_KERNEL_KINDS] = {};
  uint64_t cnt[VSG_KERNEL_KINDS] = {};
  hipError_t err = hipSuccess;
  std::vector<vsg_index::Timed> timed;
  {
    std::lock_guard<std::mutex> t(h->timing_mu);
    timed.swap(h->timed);
  }
  for (auto& t : timed) {
    float v = 0;
    hipError_t e = hipEventSynchronize(t.b);
    if (e == hipSuccess) e = hipEventElapsedTime(&v, t.a, t.b);
    if (e != hipSuccess && err == hipSuccess) err = e;
    ms[t.kind] += v;
    cnt[t.kind] += t.launches;
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  if (err != hipSuccess) return fail_hip("vsg_kernel_times", err);
  for (int i = 0; i < VSG_KERNEL_KINDS; ++i) {
    if (ms_out) ms_out[i] = ms[i];
    if (launches_out) launches_out[i] = cnt[i];
  }
  return VSG_OK;
}

int vsg_short_stats(vsg_index* h, uint64_t* emitted, uint64_t* rescored, uint64_t* fallbacks) {
  if (!h) return fail(VSG_ERR_INVALID_ARGUMENT, "null index");
  std::lock_guard<std::mutex> g(h->last_mu);
  if (emitted) *emitted = h->last_emitted;
  if (rescored) *rescored = h->last_rescored;
  if (fallbacks) *fallbacks = h->last_fallback;
  return VSG_OK;
}

int vsg_synchronize(vsg_index* h) {
  if (!h) return fail(VSG_ERR_INVALID_ARGUMENT, "null index");
  VSG_HIP(hipSetDevice(h->device));
  VSG_HIP(hipStreamSynchronize(h->stream));
  std::vector<hipEvent_t> evs;
  {
    std::lock_guard<std::mutex> g(h->lanes_mu);
    for (auto& l : h->lanes)
      if (l->pending) evs.push_back(l->done);
  }
  for (hipEvent_t ev : evs) VSG_HIP(hipEventSynchronize(ev));
  return VSG_OK;
}

int vsg_base_info(const vsg_index* h, const float** d_base, size_t* padded_dim) {
  if (!h) return fail(VSG_ERR_INVALID_ARGUMENT, "null index");
  if (d_base) *d_base = h->base;
  if (padded_dim) *padded_dim = h->dp;
  return VSG_OK;
}

}  // extern "C"

// ---- serving loop ------------------------------------------------------------------------

namespace {

using Clock = std::chrono::steady_clock;

struct ServeRequest {
  const float* q = nullptr;
  size_t nq = 0, k = 0;
  const uint32_t* allow = nullptr;
  size_t n_allow = 0;
  bool filtered = false;
  uint32_t* out_ids = nullptr;
  float* out_dists = nullptr;
  size_t bytes = 0;  // admission estimate
  Clock::time_point t_submit, t_done;
  bool done = false;
  int status = VSG_OK;
  std::string err;
};

struct ServeSlot {
  hipStream_t st = nullptr;
  Workspace w;
  std::thread th;
};

}  // namespace

struct vsg_server {
  vsg_index* h = nullptr;
  vsg_server_config cfg{};
  std::mutex m;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<ServeRequest>> queue;
  std::unordered_map<uint64_t, std::shared_ptr<ServeRequest>> live;
  uint64_t next_ticket = 1;
  size_t inflight = 0;  // admitted estimate in flight
  bool tune = true;     // filtered-scan tuning inside batches (off: untuned classes use the planner)
  int busy = 0;         // slots running a batch
  bool stop = false;
  vsg_server_stats stats{};
  std::vector<std::unique_ptr<ServeSlot>> slots;
};

namespace {

// Workspace estimate of one request: staging and device copies of its queries and results,
// the per-query scratch of the path it takes (shortlist candidate buffers and tile minima, or
// the filtered scan's per-chunk lists), and its allowed ids.
size_t request_bytes(const vsg_index* h, size_t nq, size_t k, bool filtered, size_t n_allow) {
  const size_t io = nq * (h->dim * 4 + k * 8) * 2;
  if (!filtered) {
    const size_t tiles = h->n / 256 + 1;
    return io + nq * (size_t(vsg::short_cand_cap()) * 8 + vsg::short_scatter_slices() * 4 + tiles * 4 + 64 * k * 8);
  }
  const size_t chunks = n_allow / 1024 + 1;
  return io + n_allow * 4 + nq * chunks * k * 8;
}

// One device batch on a slot. Returns the status shared by the batch's requests.
int run_batch(vsg_server* s, ServeSlot& sl, const std::vector<std::shared_ptr<ServeRequest>>& batch) {
  vsg_index* h = s->h;
  std::shared_lock<std::shared_mutex> rl(h->rw);
  VSG_HIP(hipSetDevice(h->device));
  Workspace& w = sl.w;
  size_t nu = 0, ku = 0;
  for (const auto& r : batch)
    if (!r->filtered) {
      nu += r->nq;
      ku = std::max(ku, r->k);
    }
  // Unfiltered queries take `method` (the streaming scan reads the tiled base, the shortlist
  // the low-precision copy); an unfiltered query joins the filtered scan's launch only when
  // the server was asked for it (method NAIVE keeps one persistent launch per batch).
  const bool unfiltered_in_lists = nu > 0 && s->cfg.method == VSG_METHOD_NAIVE;
  if (h->n == 0) {
    for (const auto& r : batch) fill_missing(r->out_ids, r->out_dists, r->nq * r->k);
    return VSG_OK;
  }
  if (nu > 0 && !unfiltered_in_lists) {
    hipError_t e;
    if ((e = w.q.reserve(nu * h->dim)) != hipSuccess) return fail_hip("workspace", e);
    if ((e = w.ids.reserve(nu * ku)) != hipSuccess) return fail_hip("workspace", e);
    if ((e = w.dists.reserve(nu * ku)) != hipSuccess) return fail_hip("workspace", e);
    if ((e = w.hq.reserve(nu * h->dim)) != hipSuccess) return fail_hip("pinned staging", e);
    if ((e = w.hids.reserve(nu * ku)) != hipSuccess) return fail_hip("pinned staging", e);
    if ((e = w.hdists.reserve(nu * ku)) != hipSuccess) return fail_hip("pinned staging", e);
    size_t at = 0;
    for (const auto& r : batch)
      if (!r->filtered) {
        std::memcpy(w.hq.p + at * h->dim, r->q, r->nq * h->dim * sizeof(float));
        at += r->nq;
      }
    VSG_HIP(hipMemcpyAsync(w.q.p, w.hq.p, nu * h->dim * sizeof(float), hipMemcpyHostToDevice, sl.st));
    int rc = enqueue_search(h, w, w.q.p, nu, ku, s->cfg.method, nullptr, w.ids.p, w.dists.p, sl.st);
    if (rc) return rc;
    VSG_HIP(hipMemcpyAsync(w.hids.p, w.ids.p, nu * ku * sizeof(uint32_t), hipMemcpyDeviceToHost, sl.st));
    VSG_HIP(hipMemcpyAsync(w.hdists.p, w.dists.p, nu * ku * sizeof(float), hipMemcpyDeviceToHost, sl.st));
    VSG_HIP(hipStreamSynchronize(sl.st));
    at = 0;
    bool filtered_follow = false;
    for (const auto& r : batch)
      if (!r->filtered) {
        for (size_t i = 0; i < r->nq; ++i) {  // top-k is the first k of the top-ku
          std::memcpy(r->out_ids + i * r->k, w.hids.p + (at + i) * ku, r->k * sizeof(uint32_t));
          std::memcpy(r->out_dists + i * r->k, w.hdists.p + (at + i) * ku, r->k * sizeof(float));
        }
        at += r->nq;
      } else {
        filtered_follow = true;
      }
    // the unfiltered requests are answered: release them now instead of after the filtered
    // launch that follows (their latency is their own part of the batch)
    if (filtered_follow) {
      const auto now = Clock::now();
      {
        std::lock_guard<std::mutex> lk(s->m);
        for (const auto& r : batch)
          if (!r->filtered) {
            r->status = VSG_OK;
            r->t_done = now;
            r->done = true;
          }
      }
      s->cv_done.notify_all();
    }
  }
  // filtered requests (and a small unfiltered group): one persistent launch
  size_t nl = 0, kl = 0, nid = 0;
  for (const auto& r : batch)
    if (r->filtered || unfiltered_in_lists) {
      nl += r->nq;
      kl = std::max(kl, r->k);
      if (r->filtered) nid += r->n_allow;
    }
  if (nl == 0) return VSG_OK;
  std::vector<float> q(nl * h->dim);
  std::vector<uint32_t> ks(nl), ids(std::max<size_t>(1, nid));
  std::vector<uint64_t> range(2 * nl);
  size_t at = 0, idat = 0;
  for (const auto& r : batch) {
    if (!(r->filtered || unfiltered_in_lists)) continue;
    std::memcpy(q.data() + at * h->dim, r->q, r->nq * h->dim * sizeof(float));
    if (r->filtered && r->n_allow) std::memcpy(ids.data() + idat, r->allow, r->n_allow * sizeof(uint32_t));
    for (size_t i = 0; i < r->nq; ++i) {
      ks[at + i] = uint32_t(r->k);
      range[2 * (at + i)] = r->filtered ? uint64_t(idat) : uint64_t(VSG_ALL_ROWS);
      range[2 * (at + i) + 1] = r->filtered ? uint64_t(idat + r->n_allow) : 0;
    }
    if (r->filtered) idat += r->n_allow;
    at += r->nq;
  }
  std::vector<uint32_t> oid(nl * kl);
  std::vector<float> odist(nl * kl);
  int rc = run_lists(h, w, sl.st, q.data(), nl, ks.data(), kl, range.data(), ids.data(), nid,
                     oid.data(), odist.data(), s->tune);
  if (rc) return rc;
  at = 0;
  for (const auto& r : batch) {
    if (!(r->filtered || unfiltered_in_lists)) continue;
    for (size_t i = 0; i < r->nq; ++i) {
      std::memcpy(r->out_ids + i * r->k, oid.data() + (at + i) * kl, r->k * sizeof(uint32_t));
      std::memcpy(r->out_dists + i * r->k, odist.data() + (at + i) * kl, r->k * sizeof(float));
    }
    at += r->nq;
  }
  return VSG_OK;
}

void serve_worker(vsg_server* s, ServeSlot* sl) {
  for (;;) {
    std::vector<std::shared_ptr<ServeRequest>> batch;
    size_t bytes = 0;
    {
      std::unique_lock<std::mutex> lk(s->m);
      bool counted = false;
      for (;;) {
        if (!s->queue.empty()) {
          const size_t head = s->queue.front()->bytes;
          if (s->inflight == 0 || s->inflight + head <= s->cfg.workspace_bytes) break;
          if (!counted) {
            s->stats.admission_waits++;
            counted = true;
          }
        } else if (s->stop) {
          return;
        }
        s->cv_work.wait(lk);
      }
      size_t nq = 0;
      while (!s->queue.empty()) {
        const auto& r = s->queue.front();
        if (!batch.empty() && (nq + r->nq > s->cfg.max_batch_queries ||
                               s->inflight + bytes + r->bytes > s->cfg.workspace_bytes))
          break;
        nq += r->nq;
        bytes += r->bytes;
        batch.push_back(r);
        s->queue.pop_front();
      }
      s->inflight += bytes;
      s->busy++;
      s->stats.peak_inflight_bytes = std::max<uint64_t>(s->stats.peak_inflight_bytes, s->inflight);
      s->stats.max_batch_requests = std::max<uint64_t>(s->stats.max_batch_requests, batch.size());
    }
    sl->w.st_fallback = 0;
    const int rc = run_batch(s, *sl, batch);
    const uint64_t fb = sl->w.st_fallback;
    const std::string err = rc ? g_err : std::string();
    const auto now = Clock::now();
    {
      std::lock_guard<std::mutex> lk(s->m);
      for (const auto& r : batch) {
        if (!r->done) {  // (unfiltered requests of a mixed batch were released by run_batch)
          r->status = rc;
          r->err = err;
          r->t_done = now;
          r->done = true;
        }
        s->stats.requests++;
        s->stats.queries += r->nq;
      }
      s->stats.batches++;
      s->stats.shortlist_fallbacks += fb;
      s->inflight -= bytes;
      s->busy--;
    }
    s->cv_done.notify_all();
    s->cv_work.notify_all();
  }
}

}  // namespace

extern "C" {

int vsg_server_create(vsg_index* h, const vsg_server_config* cfg, vsg_server** out) {
  if (!h || !out) return fail(VSG_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  vsg_server_config c{};
  if (cfg) c = *cfg;
  if (c.max_batch_queries == 0) c.max_batch_queries = 4096;
  if (c.workspace_bytes == 0) c.workspace_bytes = size_t(4) << 30;
  if (c.slots == 0) c.slots = 2;
  if (c.slots < 1 || c.slots > 8) return fail(VSG_ERR_INVALID_ARGUMENT, "slots must be in 1..8");
  if (c.method < VSG_METHOD_AUTO || c.method > VSG_METHOD_FP8)
    return fail(VSG_ERR_INVALID_ARGUMENT, "unknown search method");
  if (c.tune_mode != 0 && c.tune_mode != 1) return fail(VSG_ERR_INVALID_ARGUMENT, "tune_mode must be 0 or 1");
  VSG_HIP(hipSetDevice(h->device));
  vsg_server* s = new (std::nothrow) vsg_server;
  if (!s) return fail(VSG_ERR_OUT_OF_MEMORY, "host allocation failed");
  s->h = h;
  s->cfg = c;
  s->tune = c.tune_mode == 0;
  for (int i = 0; i < c.slots; ++i) {
    auto sl = std::make_unique<ServeSlot>();
    hipError_t e = hipStreamCreateWithFlags(&sl->st, hipStreamNonBlocking);
    if (e != hipSuccess) {
      vsg_server_destroy(s);
      return fail_hip("vsg_server_create", e);
    }
    s->slots.push_back(std::move(sl));
  }
  for (auto& sl : s->slots) sl->th = std::thread(serve_worker, s, sl.get());
  *out = s;
  return VSG_OK;
}

int vsg_server_submit(vsg_server* s, const float* queries, size_t nq, size_t k,
                      const uint32_t* allow_ids, size_t n_allowed, uint32_t* out_ids,
                      float* out_dists, uint64_t* ticket) {
  if (!s || !ticket) return fail(VSG_ERR_INVALID_ARGUMENT, "null argument");
  if (k > VSG_MAX_K) return fail(VSG_ERR_INVALID_ARGUMENT, "k must be <= " + std::to_string(VSG_MAX_K));
  if ((nq && !queries) || (nq * k && (!out_ids || !out_dists)))
    return fail(VSG_ERR_INVALID_ARGUMENT, "null buffer");
  if (allow_ids == nullptr && n_allowed != 0) return fail(VSG_ERR_INVALID_ARGUMENT, "null allow_ids");
  vsg_index* h = s->h;
  {
    std::shared_lock<std::shared_mutex> rl(h->rw);
    for (size_t i = 0; i < n_allowed; ++i)
      if (allow_ids[i] >= h->n || (i && allow_ids[i] <= allow_ids[i - 1]))
        return fail(VSG_ERR_INVALID_ARGUMENT, "allowed ids must be strictly increasing and < size()");
  }
  auto r = std::make_shared<ServeRequest>();
  r->q = queries;
  r->nq = nq;
  r->k = k;
  r->allow = allow_ids;
  r->n_allow = n_allowed;
  r->filtered = allow_ids != nullptr;
  r->out_ids = out_ids;
  r->out_dists = out_dists;
  r->bytes = request_bytes(h, nq, k, r->filtered, n_allowed);
  r->t_submit = Clock::now();
  std::lock_guard<std::mutex> lk(s->m);
  if (s->stop) return fail(VSG_ERR_INVALID_ARGUMENT, "server is shutting down");
  *ticket = s->next_ticket++;
  s->live[*ticket] = r;
  if (nq == 0 || k == 0) {
    r->done = true;
    r->t_done = r->t_submit;
    s->stats.requests++;
    return VSG_OK;
  }
  s->queue.push_back(r);
  s->cv_work.notify_one();
  return VSG_OK;
}

int vsg_server_wait(vsg_server* s, uint64_t ticket, double* latency_ms) {
  if (!s) return fail(VSG_ERR_INVALID_ARGUMENT, "null server");
  std::shared_ptr<ServeRequest> r;
  {
    std::unique_lock<std::mutex> lk(s->m);
    auto it = s->live.find(ticket);
    if (it == s->live.end()) return fail(VSG_ERR_INVALID_ARGUMENT, "unknown or already waited ticket");
    r = it->second;
    s->cv_done.wait(lk, [&] { return r->done; });
    s->live.erase(it);
  }
  if (latency_ms)
    *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
  if (r->status) return fail(r->status, r->err);
  return VSG_OK;
}

int vsg_server_get_stats(vsg_server* s, vsg_server_stats* out) {
  if (!s || !out) return fail(VSG_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> lk(s->m);
  *out = s->stats;
  return VSG_OK;
}

void vsg_server_destroy(vsg_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->m);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto& sl : s->slots)
    if (sl->th.joinable()) sl->th.join();
  (void)hipSetDevice(s->h->device);
  for (auto& sl : s->slots) {
    if (sl->st) (void)hipStreamSynchronize(sl->st);
    sl->w.release();
    if (sl->st) (void)hipStreamDestroy(sl->st);
  }
  delete s;
}

}  // extern "C"

// ---- captured searches -------------------------------------------------------------------

struct vsg_graph {
  vsg_index* h = nullptr;
  Workspace w;
  hipStream_t st = nullptr;  // capture stream
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  size_t n = 0;
  const float* base = nullptr;
};

extern "C" {

int vsg_graph_create(vsg_index* h, const float* d_queries, size_t nq, size_t k, int method,
                     uint32_t* d_ids, float* d_dists, vsg_graph** out) {
  if (!out) return fail(VSG_ERR_INVALID_ARGUMENT, "null output handle");
  *out = nullptr;
  int rc = check_search_args(h, k, method);
  if (rc) return rc;
  if (nq == 0 || k == 0 || !d_queries || !d_ids || !d_dists)
    return fail(VSG_ERR_INVALID_ARGUMENT, "a captured search needs nq > 0, k > 0 and device buffers");
  std::lock_guard<std::mutex> g(h->mu);
  std::shared_lock<std::shared_mutex> rl(h->rw);
  if (h->n == 0) return fail(VSG_ERR_INVALID_ARGUMENT, "cannot capture a search of an empty index");
  // the path an eager call takes (enqueue_search): the shortlist (p > 0) or the streaming scan
  int p = 0;
  if (short_supported(h, k)) {
    if (method == VSG_METHOD_TILED || method == VSG_METHOD_BF16) p = 1;
    else if (method == VSG_METHOD_FP8) p = 2;
    else if (method == VSG_METHOD_AUTO && nq >= size_t(h->auto_short_batch)) p = 1;
  }
  VSG_HIP(hipSetDevice(h->device));
  vsg_graph* gr = new (std::nothrow) vsg_graph;
  if (!gr) return fail(VSG_ERR_OUT_OF_MEMORY, "host allocation failed");
  gr->h = h;
  gr->n = h->n;
  gr->base = h->base;
  auto cleanup = [&](int code) {
    vsg_graph_destroy(gr);
    return code;
  };
  hipError_t e = hipStreamCreateWithFlags(&gr->st, hipStreamNonBlocking);
  if (e != hipSuccess) return cleanup(fail_hip("vsg_graph_create", e));
  // size the workspace before capture (no allocation may happen inside it)
  vsg::ListsLaunch fallback;
  if (p) {
    // the low-precision copy, the shortlist's buffers, and the device-planned exact fallback
    // (a filtered-scan launch sized for a whole round of queries)
    if ((rc = ensure_prec(h, p))) return cleanup(rc);
    ShortPlan sp;
    if ((rc = short_reserve(h, gr->w, nq, k, p, sp))) return cleanup(rc);
    if ((rc = vsg::prepare_lists_fallback(gr->w.lists, h->n, h->dim, h->dp, h->metric, h->num_cus, gr->st,
                                          sp.round, k, g_err, fallback)))
      return cleanup(rc);
  } else {
    const ScanPlan sc = plan_scan(h);
    if ((e = gr->w.part.reserve(std::min<size_t>(kPartBudgetBytes / sizeof(uint64_t),
                                                 ((nq + 7) / 8 * 8) * size_t(sc.W) * k))) != hipSuccess)
      return cleanup(fail_hip("workspace", e));
  }
  e = hipStreamBeginCapture(gr->st, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return cleanup(fail_hip("hipStreamBeginCapture", e));
  t_capturing = true;  // no event records inside the graph
  if (p)
    rc = enqueue_short(h, gr->w, d_queries, nq, k, p, d_ids, d_dists, gr->st, &fallback);
  else
    rc = enqueue_exact(h, gr->w, d_queries, nq, k, method == VSG_METHOD_AUTO ? VSG_METHOD_SKINNY : method,
                       nullptr, d_ids, d_dists, gr->st);
  t_capturing = false;
  hipGraph_t graph = nullptr;
  e = hipStreamEndCapture(gr->st, &graph);
  gr->graph = graph;
  if (rc) return cleanup(rc);
  if (e != hipSuccess) return cleanup(fail_hip("hipStreamEndCapture", e));
  e = hipGraphInstantiate(&gr->exec, gr->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) return cleanup(fail_hip("hipGraphInstantiate", e));
  *out = gr;
  return VSG_OK;
}

int vsg_graph_create_lists(vsg_index* h, const float* d_queries, size_t nq, const uint32_t* ks, size_t kmax,
                           const uint64_t* range, const uint32_t* allow_ids, size_t n_ids, uint32_t* d_ids,
                           float* d_dists, vsg_graph** out) {
  if (!out) return fail(VSG_ERR_INVALID_ARGUMENT, "null output handle");
  *out = nullptr;
  if (!h) return fail(VSG_ERR_INVALID_ARGUMENT, "null index");
  if (kmax > VSG_MAX_K) return fail(VSG_ERR_INVALID_ARGUMENT, "k must be <= " + std::to_string(VSG_MAX_K));
  if (nq == 0 || kmax == 0 || !d_queries || !ks || !range || !d_ids || !d_dists)
    return fail(VSG_ERR_INVALID_ARGUMENT, "a captured search needs nq > 0, kmax > 0 and device buffers");
  std::lock_guard<std::mutex> g(h->mu);
  std::shared_lock<std::shared_mutex> rl(h->rw);
  if (h->n == 0) return fail(VSG_ERR_INVALID_ARGUMENT, "cannot capture a search of an empty index");
  VSG_HIP(hipSetDevice(h->device));
  vsg_graph* gr = new (std::nothrow) vsg_graph;
  if (!gr) return fail(VSG_ERR_OUT_OF_MEMORY, "host allocation failed");
  gr->h = h;
  gr->n = h->n;
  gr->base = h->base;
  auto cleanup = [&](int code) {
    vsg_graph_destroy(gr);
    return code;
  };
  hipError_t e = hipStreamCreateWithFlags(&gr->st, hipStreamNonBlocking);
  if (e != hipSuccess) return cleanup(fail_hip("vsg_graph_create_lists", e));
  // plan and upload before capture: the graph holds the kernels only, reading the plan, ks and
  // allowed ids from its own workspace
  vsg::ListsLaunch L;
  int rc = vsg::prepare_lists_device(gr->w.lists, h->n, h->dim, h->dp, h->metric, h->num_cus, gr->st, nq, ks, kmax,
                                     range, allow_ids, n_ids, g_err, lists_tiles_for(h, nq, kmax, range), L);
  if (rc) return cleanup(rc);
  e = hipStreamBeginCapture(gr->st, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return cleanup(fail_hip("hipStreamBeginCapture", e));
  const hipError_t le = vsg::lists_enqueue(gr->w.lists, L, h->base, d_queries, d_ids, d_dists, gr->st, nullptr);
  hipGraph_t graph = nullptr;
  e = hipStreamEndCapture(gr->st, &graph);
  gr->graph = graph;
  if (le != hipSuccess) return cleanup(fail_hip("vsg_graph_create_lists", le));
  if (e != hipSuccess) return cleanup(fail_hip("hipStreamEndCapture", e));
  e = hipGraphInstantiate(&gr->exec, gr->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) return cleanup(fail_hip("hipGraphInstantiate", e));
  *out = gr;
  return VSG_OK;
}

int vsg_graph_launch(vsg_graph* g, void* stream) {
  if (!g) return fail(VSG_ERR_INVALID_ARGUMENT, "null graph");
  vsg_index* h = g->h;
  std::shared_lock<std::shared_mutex> rl(h->rw);
  if (h->n != g->n || h->base != g->base)
    return fail(VSG_ERR_INVALID_ARGUMENT, "the index changed since the graph was captured");
  VSG_HIP(hipSetDevice(h->device));
  VSG_HIP(hipGraphLaunch(g->exec, stream ? static_cast<hipStream_t>(stream) : h->stream));
  return VSG_OK;
}

void vsg_graph_destroy(vsg_graph* g) {
  if (!g) return;
  (void)hipSetDevice(g->h->device);
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->st) {
    (void)hipStreamSynchronize(g->st);
    (void)hipStreamDestroy(g->st);
  }
  g->w.release();
  delete g;
}

}  // extern "C"

// ---- saved HNSW indexes and the exact_below dispatch ---------------------------------------

namespace {

struct HnswHeader {
  uint64_t dim = 0, n = 0;
  uint32_t metric = 0;
};

// hnsw.cpp:369-389: 8-byte magic, then dim, n, metric, M, ef_construction, max_level,
// entry_point; the n x dim vector block starts at byte 52.
constexpr uint64_t kHnswHeaderBytes = 52;

int read_hnsw_header(std::ifstream& f, const char* path, HnswHeader& hd) {
  char magic[8];
  f.read(magic, 8);
  uint64_t dim = 0, n = 0, M = 0, efc = 0;
  uint32_t metric = 0, ep = 0;
  int32_t maxl = 0;
  f.read(reinterpret_cast<char*>(&dim), 8);
  f.read(reinterpret_cast<char*>(&n), 8);
  f.read(reinterpret_cast<char*>(&metric), 4);
  f.read(reinterpret_cast<char*>(&M), 8);
  f.read(reinterpret_cast<char*>(&efc), 8);
  f.read(reinterpret_cast<char*>(&maxl), 4);
  f.read(reinterpret_cast<char*>(&ep), 4);
  if (!f) return fail(VSG_ERR_INVALID_ARGUMENT, std::string("truncated header: ") + path);
  if (std::memcmp(magic, "VSHNSW01", 8) != 0)
    return fail(VSG_ERR_INVALID_ARGUMENT, std::string("not a vecsearch HNSW file: ") + path);
  if (dim == 0 || metric > 1) return fail(VSG_ERR_INVALID_ARGUMENT, std::string("bad header: ") + path);
  hd.dim = dim;
  hd.n = n;
  hd.metric = metric;
  return VSG_OK;
}

}  // namespace

extern "C" {

int vsg_hnsw_file_info(const char* path, size_t* dim, size_t* n, int* metric) {
  if (!path) return fail(VSG_ERR_INVALID_ARGUMENT, "null path");
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail(VSG_ERR_INVALID_ARGUMENT, std::string("cannot open ") + path);
  HnswHeader hd;
  int rc = read_hnsw_header(f, path, hd);
  if (rc) return rc;
  if (dim) *dim = size_t(hd.dim);
  if (n) *n = size_t(hd.n);
  if (metric) *metric = int(hd.metric);
  return VSG_OK;
}

int vsg_add_hnsw_file(vsg_index* h, const char* path, size_t* rows_added) {
  if (!h || !path) return fail(VSG_ERR_INVALID_ARGUMENT, "null argument");
  if (rows_added) *rows_added = 0;
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail(VSG_ERR_INVALID_ARGUMENT, std::string("cannot open ") + path);
  HnswHeader hd;
  int rc = read_hnsw_header(f, path, hd);
  if (rc) return rc;
  if (hd.dim != h->dim) return fail(VSG_ERR_INVALID_ARGUMENT, "the file's dim differs from the index's");
  if (int(hd.metric) != h->metric) return fail(VSG_ERR_INVALID_ARGUMENT, "the file's metric differs from the index's");
  f.seekg(0, std::ios::end);
  const uint64_t size = uint64_t(f.tellg());
  if (size < kHnswHeaderBytes + hd.n * hd.dim * sizeof(float))
    return fail(VSG_ERR_INVALID_ARGUMENT, std::string("truncated vector block: ") + path);
  f.seekg(std::streamoff(kHnswHeaderBytes), std::ios::beg);
  std::lock_guard<std::mutex> g(h->mu);
  std::unique_lock<std::shared_mutex> wl(h->rw);
  if (h->n + hd.n >= VSG_NONE) return fail(VSG_ERR_LENGTH, "an index holds at most 2^32 - 2 vectors");
  if (hd.n == 0) return VSG_OK;
  VSG_HIP(hipSetDevice(h->device));
  rc = grow(h, h->n + size_t(hd.n));  // one allocation for the whole block
  if (rc) return rc;
  // the block is read slice by slice straight into the pinned upload slices; a failed read
  // leaves the index as it was
  rc = upload_rows(h, size_t(hd.n), [&](float* dst, size_t, size_t c) {
    f.read(reinterpret_cast<char*>(dst), std::streamsize(c * h->dim * sizeof(float)));
    if (!f) return fail(VSG_ERR_INVALID_ARGUMENT, std::string("read failed: ") + path);
    return VSG_OK;
  });
  if (rc == VSG_OK && rows_added) *rows_added = size_t(hd.n);
  return rc;
}

int vsg_search_exact_below(vsg_index* h, const float* queries, size_t nq, size_t k,
                           const uint8_t* mask, size_t exact_below, uint32_t* out_ids,
                           float* out_dists, int* handled) {
  if (!h || !handled) return fail(VSG_ERR_INVALID_ARGUMENT, "null argument");
  *handled = 0;
  if (!mask) return VSG_OK;  // no filter: the graph search
  size_t n;
  {
    std::shared_lock<std::shared_mutex> rl(h->rw);
    n = h->n;
  }
  const size_t allowed = size_t(std::count_if(mask, mask + n, [](uint8_t v) { return v != 0; }));
  if (allowed != 0 && allowed > exact_below) return VSG_OK;
  *handled = 1;
  return vsg_search(h, queries, nq, k, VSG_METHOD_AUTO, mask, out_ids, out_dists);
}

}  // extern "C"
