// C entry points over the reference CPU index, compiled from the reference's own sources
// (src/brute_force.cpp, src/hnsw.cpp, include/vecsearch/*.h) into oracle/_ref/libvsref.so by
// oracle/Makefile. TEST INFRASTRUCTURE ONLY: used to pin the C oracle and, optionally, as the
// CPU baseline. This file is a driver of our own; no reference source is copied here.
#include <algorithm>
#include <cstdint>
#include <memory>

#include "vecsearch/brute_force.h"
#include "vecsearch/distance.h"
#include "vecsearch/hnsw.h"

using namespace vecsearch;

extern "C" {

float vsref_dot(const float* a, const float* b, size_t d) { return dot(a, b, d); }
float vsref_l2sq(const float* a, const float* b, size_t d) { return l2_sq(a, b, d); }
float vsref_distance(int metric, const float* a, const float* b, size_t d) {
  return distance(metric == 0 ? Metric::L2 : Metric::InnerProduct, a, b, d);
}

// BruteForceIndex (brute_force.h:8-32): create / add / search_batch / single search / destroy.
void* vsref_bf_create(size_t dim, int metric) {
  return new BruteForceIndex(dim, metric == 0 ? Metric::L2 : Metric::InnerProduct);
}
void vsref_bf_destroy(void* h) { delete static_cast<BruteForceIndex*>(h); }
void vsref_bf_add(void* h, const float* data, size_t n) {
  static_cast<BruteForceIndex*>(h)->add(data, n);
}
void vsref_bf_search_batch(void* h, const float* q, size_t nq, size_t k, uint32_t* ids,
                           float* dists, int threads) {
  static_cast<BruteForceIndex*>(h)->search_batch(q, nq, k, ids, dists, threads);
}
// BruteForceIndex::search (one query, priority_queue); returns the number of results.
size_t vsref_bf_search_one(void* h, const float* q, size_t k, uint32_t* ids, float* dists) {
  auto r = static_cast<BruteForceIndex*>(h)->search(q, k);
  for (size_t j = 0; j < r.size(); ++j) {
    ids[j] = r[j].id;
    dists[j] = r[j].dist;
  }
  return r.size();
}

// HNSWIndex::search_batch with SearchFilter{mask, count, exact_below} (hnsw.h:29-33): at or
// below exact_below allowed rows every query takes search_exact (hnsw.cpp:322-354).
void vsref_hnsw_filtered(const float* data, size_t n, size_t d, int metric, const float* q,
                         size_t nq, size_t k, const uint8_t* mask, size_t exact_below,
                         uint32_t* ids, float* dists, int threads) {
  HNSWIndex h(d, n, metric == 0 ? Metric::L2 : Metric::InnerProduct);
  h.add(data, n, threads);
  SearchFilter f;
  f.allowed = mask;
  f.n_allowed = size_t(std::count_if(mask, mask + n, [](uint8_t v) { return v != 0; }));
  f.exact_below = exact_below;
  h.search_batch(q, nq, k, 64, ids, dists, threads, f);
}

// HNSWIndex::save (hnsw.cpp:375-389): an index of the given rows written to path, so the
// engine's reader of that format is checked against the reference's own writer.
int vsref_hnsw_save(const float* data, size_t n, size_t d, int metric, const char* path) {
  try {
    HNSWIndex h(d, n, metric == 0 ? Metric::L2 : Metric::InnerProduct);
    h.add(data, n, 1);
    h.save(path);
    return 0;
  } catch (...) {
    return 1;
  }
}

}  // extern "C"
