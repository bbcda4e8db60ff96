// vecsearch::GpuBruteForceIndex over the vsg C ABI (host C++, no HIP headers).
// Error mapping follows the reference (gpu_brute_force.cu:11-16, 297, 387, 408):
// invalid arguments -> std::invalid_argument, capacity -> std::length_error, anything from the
// runtime -> std::runtime_error.
#include "vecsearch/gpu_brute_force.h"

#include <stdexcept>
#include <string>

#include "vsg.h"

namespace vecsearch {

namespace {

void check(int rc) {
  if (rc == VSG_OK) return;
  const std::string msg = vsg_last_error();
  switch (rc) {
    case VSG_ERR_INVALID_ARGUMENT: throw std::invalid_argument(msg);
    case VSG_ERR_LENGTH: throw std::length_error(msg);
    default: throw std::runtime_error(msg);
  }
}

}  // namespace

bool hip_available() {
  int n = 0;
  vsg_device_count(&n);
  return n > 0;
}

bool cuda_available() { return hip_available(); }

GpuBruteForceIndex::GpuBruteForceIndex(size_t dim, Metric metric, int device)
    : dim_(dim), metric_(metric) {
  check(vsg_create(dim, static_cast<int>(metric), device, &h_));
}

GpuBruteForceIndex::~GpuBruteForceIndex() { vsg_destroy(h_); }

void GpuBruteForceIndex::add(const float* data, size_t n) { check(vsg_add(h_, data, n)); }

void GpuBruteForceIndex::search(const float* queries, size_t nq, size_t k, uint32_t* out_ids,
                                float* out_dists, GpuMethod method) {
  check(vsg_search(h_, queries, nq, k, static_cast<int>(method), nullptr, out_ids, out_dists));
}

void GpuBruteForceIndex::search(const float* queries, size_t nq, size_t k, uint32_t* out_ids,
                                float* out_dists, const uint8_t* allowed, GpuMethod method) {
  check(vsg_search(h_, queries, nq, k, static_cast<int>(method), allowed, out_ids, out_dists));
}

void GpuBruteForceIndex::search_lists(const float* queries, size_t nq, const uint32_t* ks,
                                      size_t kmax, const uint64_t* range,
                                      const uint32_t* allow_ids, size_t n_ids, uint32_t* out_ids,
                                      float* out_dists) {
  check(vsg_search_lists(h_, queries, nq, ks, kmax, range, allow_ids, n_ids, out_ids, out_dists));
}

size_t GpuBruteForceIndex::size() const {
  size_t n = 0;
  check(vsg_size(h_, &n));
  return n;
}

size_t GpuBruteForceIndex::dim() const { return dim_; }
Metric GpuBruteForceIndex::metric() const { return metric_; }

}  // namespace vecsearch
