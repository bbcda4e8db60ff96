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
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "vsg.h"
#include "vsg_internal.h"

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

struct vsg_index {
  size_t dim = 0, dp = 0;
  int metric = VSG_METRIC_L2;
  int device = 0;
  int num_cus = 256;
  size_t n = 0, cap = 0;
  float* base = nullptr;   // cap x dp, row-major, zero padded (gathers: filters, rescoring)
  float* tiles = nullptr;  // the same rows in 16-row step-major tiles (streaming scans)
  int scan_waves_per_cu = 8;
  hipStream_t stream = nullptr;
  std::mutex mu;

  DevBuf<float> q;
  DevBuf<uint64_t> part;
  DevBuf<uint32_t> ids;
  DevBuf<float> dists;
  DevBuf<uint8_t> mask;
  HostBuf<float> hq;
  HostBuf<uint32_t> hids;
  HostBuf<float> hdists;
  vsg::ListsWorkspace lists;

  // kernel timing (vsg_kernel_timing)
  struct Timed {
    hipEvent_t a, b;
    int kind;  // 0 scan, 1 merge
    uint32_t launches;
  };
  bool timing = false;
  std::vector<Timed> timed;

  ~vsg_index() {
    for (auto& t : timed) {
      (void)hipEventDestroy(t.a);
      (void)hipEventDestroy(t.b);
    }
    if (stream) (void)hipStreamSynchronize(stream);
    if (base) (void)hipFree(base);
    if (tiles) (void)hipFree(tiles);
    q.release();
    part.release();
    ids.release();
    dists.release();
    mask.release();
    hq.release();
    hids.release();
    hdists.release();
    lists.release();
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

// Plan of one streaming scan over the base: n_blocks workgroups of 4 waves, each wave owning
// rows_per_wave consecutive rows (a multiple of 64).
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
  p.W = uint32_t(blocks * 4);
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
    if (!h->timing) return;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess ||
        hipEventRecord(a, st) != hipSuccess) {
      h->timing = false;  // give up timing rather than fail the search
    }
  }
  ~TimedScope() {
    if (!a || !b) return;
    if (h->timing && hipEventRecord(b, st) == hipSuccess) {
      h->timed.push_back({a, b, kind, launches});
    } else {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  }
};

// Enqueue a full search of nq device-resident queries on stream st.
int enqueue_search(vsg_index* h, const float* d_q, size_t nq, size_t k, int method,
                   const uint8_t* d_mask, uint32_t* d_ids, float* d_dists, hipStream_t st) {
  const ScanPlan p = plan_scan(h);
  const int qb = queries_per_pass(method, nq, h->dp);
  const size_t per_query = size_t(p.W) * k * sizeof(uint64_t);
  size_t round = std::max<size_t>(qb, kPartBudgetBytes / per_query / qb * qb);
  round = std::min(round, (nq + qb - 1) / qb * qb);
  hipError_t e = h->part.reserve(round * size_t(p.W) * k);
  if (e != hipSuccess) return fail_hip("workspace", e);
  const bool l2 = h->metric == VSG_METRIC_L2;
  for (size_t r0 = 0; r0 < nq; r0 += round) {
    const size_t rq = std::min(round, nq - r0);
    const float* qr = d_q + r0 * h->dim;
    {
      TimedScope ts(h, st, 0);
      for (size_t q0 = 0; q0 < rq; q0 += qb) {
        const uint32_t nqp = uint32_t(std::min<size_t>(qb, rq - q0));
        e = vsg::launch_scan(qb, l2, uint32_t(k), p.n_blocks, st, h->tiles, uint32_t(h->n),
                             uint32_t(h->dp), uint32_t(h->dim), qr, uint32_t(q0), nqp,
                             p.rows_per_wave, d_mask, h->part.p);
        if (e != hipSuccess) return fail_hip("scan kernel", e);
        ts.launches++;
      }
    }
    {
      TimedScope ts(h, st, 1);
      e = vsg::launch_merge(uint32_t(rq), h->part.p, nullptr, p.W, uint32_t(k), nullptr,
                            uint32_t(k), nullptr, d_ids + r0 * k, d_dists + r0 * k, uint32_t(k), st);
      if (e != hipSuccess) return fail_hip("merge kernel", e);
      ts.launches++;
    }
  }
  return VSG_OK;
}

int check_search_args(const vsg_index* h, size_t k, int method) {
  if (!h) return fail(VSG_ERR_INVALID_ARGUMENT, "null index");
  if (k > VSG_MAX_K)
    return fail(VSG_ERR_INVALID_ARGUMENT, "k must be <= " + std::to_string(VSG_MAX_K));
  if (method < VSG_METHOD_AUTO || method > VSG_METHOD_TILED)
    return fail(VSG_ERR_INVALID_ARGUMENT, "unknown search method");
  return VSG_OK;
}

void fill_missing(uint32_t* ids, float* dists, size_t count) {
  std::fill(ids, ids + count, VSG_NONE);
  std::fill(dists, dists + count, INFINITY);
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
  if (n >= VSG_NONE) return fail(VSG_ERR_LENGTH, "an index holds at most 2^32 - 2 vectors");
  VSG_HIP(hipSetDevice(h->device));
  return grow(h, n);
}

int vsg_add(vsg_index* h, const float* data, size_t m) {
  if (!h) return fail(VSG_ERR_INVALID_ARGUMENT, "null index");
  std::lock_guard<std::mutex> g(h->mu);
  if (m == 0) return VSG_OK;
  if (!data) return fail(VSG_ERR_INVALID_ARGUMENT, "null data");
  if (h->n + m >= VSG_NONE) return fail(VSG_ERR_LENGTH, "an index holds at most 2^32 - 2 vectors");
  VSG_HIP(hipSetDevice(h->device));
  int rc = grow(h, h->n + m);
  if (rc) return rc;
  VSG_HIP(hipMemcpy2DAsync(h->base + h->n * h->dp, h->dp * sizeof(float), data, h->dim * sizeof(float),
                           h->dim * sizeof(float), m, hipMemcpyHostToDevice, h->stream));
  VSG_HIP(vsg::launch_tile_rows(h->stream, h->base, h->tiles, uint32_t(h->n), uint32_t(h->n + m),
                                uint32_t(h->dp)));
  VSG_HIP(hipStreamSynchronize(h->stream));
  h->n += m;
  return VSG_OK;
}

int vsg_search(vsg_index* h, const float* queries, size_t nq, size_t k, int method,
               const uint8_t* mask, uint32_t* out_ids, float* out_dists) {
  int rc = check_search_args(h, k, method);
  if (rc) return rc;
  if (nq == 0 || k == 0) return VSG_OK;
  if (!queries || !out_ids || !out_dists) return fail(VSG_ERR_INVALID_ARGUMENT, "null buffer");
  std::lock_guard<std::mutex> g(h->mu);
  if (h->n == 0) {
    fill_missing(out_ids, out_dists, nq * k);
    return VSG_OK;
  }
  VSG_HIP(hipSetDevice(h->device));
  if (mask) {
    const size_t allowed = size_t(std::count_if(mask, mask + h->n, [](uint8_t v) { return v != 0; }));
    if (allowed == 0) {
      fill_missing(out_ids, out_dists, nq * k);  // hnsw.cpp:326: no allowed row, no result
      return VSG_OK;
    }
    if (allowed * 4 < h->n) {
      // Sparse filter: gather the allowed rows (the exact_below regime) instead of streaming
      // the whole base; every query shares the one id list.
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
      return vsg::search_lists_host(h->lists, h->base, h->n, h->dim, h->dp, h->metric, h->num_cus,
                                    h->stream, queries, nq, ks.data(), k, range.data(), ids.data(),
                                    ids.size(), out_ids, out_dists, g_err);
    }
  }
  hipError_t e;
  if ((e = h->q.reserve(nq * h->dim)) != hipSuccess) return fail_hip("workspace", e);
  if ((e = h->ids.reserve(nq * k)) != hipSuccess) return fail_hip("workspace", e);
  if ((e = h->dists.reserve(nq * k)) != hipSuccess) return fail_hip("workspace", e);
  if ((e = h->hq.reserve(nq * h->dim)) != hipSuccess) return fail_hip("pinned staging", e);
  if ((e = h->hids.reserve(nq * k)) != hipSuccess) return fail_hip("pinned staging", e);
  if ((e = h->hdists.reserve(nq * k)) != hipSuccess) return fail_hip("pinned staging", e);
  const uint8_t* d_mask = nullptr;
  if (mask) {
    if ((e = h->mask.reserve(h->n)) != hipSuccess) return fail_hip("workspace", e);
    VSG_HIP(hipMemcpyAsync(h->mask.p, mask, h->n, hipMemcpyHostToDevice, h->stream));
    d_mask = h->mask.p;
  }
  std::memcpy(h->hq.p, queries, nq * h->dim * sizeof(float));
  VSG_HIP(hipMemcpyAsync(h->q.p, h->hq.p, nq * h->dim * sizeof(float), hipMemcpyHostToDevice, h->stream));
  rc = enqueue_search(h, h->q.p, nq, k, method, d_mask, h->ids.p, h->dists.p, h->stream);
  if (rc) return rc;
  VSG_HIP(hipMemcpyAsync(h->hids.p, h->ids.p, nq * k * sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream));
  VSG_HIP(hipMemcpyAsync(h->hdists.p, h->dists.p, nq * k * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  VSG_HIP(hipStreamSynchronize(h->stream));
  std::memcpy(out_ids, h->hids.p, nq * k * sizeof(uint32_t));
  std::memcpy(out_dists, h->hdists.p, nq * k * sizeof(float));
  return VSG_OK;
}

int vsg_search_device(vsg_index* h, const float* d_queries, size_t nq, size_t k, int method,
                      const uint8_t* d_mask, uint32_t* d_ids, float* d_dists, void* stream) {
  int rc = check_search_args(h, k, method);
  if (rc) return rc;
  if (nq == 0 || k == 0) return VSG_OK;
  if (!d_queries || !d_ids || !d_dists) return fail(VSG_ERR_INVALID_ARGUMENT, "null buffer");
  std::lock_guard<std::mutex> g(h->mu);
  VSG_HIP(hipSetDevice(h->device));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : h->stream;
  if (h->n == 0) {
    VSG_HIP(vsg::launch_fill_missing(st, d_ids, d_dists, nq * k));
    return VSG_OK;
  }
  return enqueue_search(h, d_queries, nq, k, method, d_mask, d_ids, d_dists, st);
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
  std::lock_guard<std::mutex> g(h->mu);
  VSG_HIP(hipSetDevice(h->device));
  return vsg::search_lists_host(h->lists, h->base, h->n, h->dim, h->dp, h->metric, h->num_cus,
                                h->stream, queries, nq, ks, kmax, range, allow_ids, n_ids,
                                out_ids, out_dists, g_err);
}

int vsg_set_param(vsg_index* h, const char* name, int64_t value) {
  if (!h || !name) return fail(VSG_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> g(h->mu);
  const std::string key(name);
  if (key == "scan_waves_per_cu") {
    if (value < 1 || value > 32) return fail(VSG_ERR_INVALID_ARGUMENT, "scan_waves_per_cu must be in 1..32");
    h->scan_waves_per_cu = int(value);
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

int vsg_kernel_times(vsg_index* h, double* scan_ms, uint64_t* scan_launches, double* merge_ms,
                     uint64_t* merge_launches) {
  if (!h) return fail(VSG_ERR_INVALID_ARGUMENT, "null index");
  std::lock_guard<std::mutex> g(h->mu);
  VSG_HIP(hipSetDevice(h->device));
  double ms[2] = {0, 0};
  uint64_t cnt[2] = {0, 0};
  hipError_t err = hipSuccess;
  for (auto& t : h->timed) {
    float v = 0;
    hipError_t e = hipEventSynchronize(t.b);
    if (e == hipSuccess) e = hipEventElapsedTime(&v, t.a, t.b);
    if (e != hipSuccess && err == hipSuccess) err = e;
    ms[t.kind] += v;
    cnt[t.kind] += t.launches;
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  h->timed.clear();
  if (err != hipSuccess) return fail_hip("vsg_kernel_times", err);
  if (scan_ms) *scan_ms = ms[0];
  if (scan_launches) *scan_launches = cnt[0];
  if (merge_ms) *merge_ms = ms[1];
  if (merge_launches) *merge_launches = cnt[1];
  return VSG_OK;
}

int vsg_synchronize(vsg_index* h) {
  if (!h) return fail(VSG_ERR_INVALID_ARGUMENT, "null index");
  VSG_HIP(hipSetDevice(h->device));
  VSG_HIP(hipStreamSynchronize(h->stream));
  return VSG_OK;
}

int vsg_base_info(const vsg_index* h, const float** d_base, size_t* padded_dim) {
  if (!h) return fail(VSG_ERR_INVALID_ARGUMENT, "null index");
  if (d_base) *d_base = h->base;
  if (padded_dim) *padded_dim = h->dp;
  return VSG_OK;
}

}  // extern "C"
