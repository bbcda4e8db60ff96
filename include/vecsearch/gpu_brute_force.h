// Exact k-NN on one AMD Instinct GPU (MI355X, gfx950): the drop-in for the reference's
// vecsearch::GpuBruteForceIndex (reference include/vecsearch/gpu_brute_force.h:16-50).
//
// Same class name, constructor, add/search signatures, kNone/kMaxK, output layout (nq x k,
// nearest first, missing = kNone / +inf) and exception types. Distances are bit-identical to
// the reference CPU index (BruteForceIndex, brute_force.cpp) and ties go to the lower id.
// Additions: a search overload with an allowed-row mask (the SearchFilter convention,
// hnsw.h:29-33) and search_lists, a batched filtered scan with a per-query k and per-query
// allowed-id lists.
//
// The class is plain host C++ over the C ABI in vsg.h; it includes no HIP header.
#pragma once
#include <cstddef>
#include <cstdint>

struct vsg_index;

// Metric is the reference's own type (distance.h:15) wherever the reference headers are on the
// include path, so this header compiles beside brute_force.h / hnsw.h (as the reference's
// python/bindings.cpp:10-14 includes them); standalone it declares the same enum.
#if __has_include("vecsearch/distance.h")
#include "vecsearch/distance.h"
#else
namespace vecsearch {
enum class Metric : uint32_t { L2 = 0, InnerProduct = 1 };
}  // namespace vecsearch
#endif

namespace vecsearch {

// gpu_brute_force.h:16-21. Every method returns the same ids and distances; they differ in
// how the scan is run: Naive = one query per streaming pass, Skinny = up to 8 queries per
// streaming pass, Tiled / Bf16 = the certified BF16 MFMA shortlist, Fp8 = the certified FP8
// shortlist, Auto = by batch size (see VSG_METHOD_* in vsg.h).
enum class GpuMethod : uint32_t { Auto = 0, Naive = 1, Skinny = 2, Tiled = 3, Bf16 = 4, Fp8 = 5 };

// gpu_brute_force.h:23-24: true if at least one GPU is visible (HIP runtime). The name is kept
// for callers of the reference API; hip_available() is the same test.
bool cuda_available();
bool hip_available();

class GpuBruteForceIndex {
 public:
  static constexpr uint32_t kNone = 0xFFFFFFFFu;  // id of a missing result (k > size)
  static constexpr size_t kMaxK = 128;

  GpuBruteForceIndex(size_t dim, Metric metric = Metric::L2, int device = 0);
  ~GpuBruteForceIndex();
  GpuBruteForceIndex(const GpuBruteForceIndex&) = delete;
  GpuBruteForceIndex& operator=(const GpuBruteForceIndex&) = delete;

  void add(const float* data, size_t n);

  // Top-k for nq host queries into host arrays (nq x k, nearest first; missing = kNone, inf).
  // Includes copying the queries in and the results out. k <= kMaxK.
  void search(const float* queries, size_t nq, size_t k, uint32_t* out_ids, float* out_dists,
              GpuMethod method = GpuMethod::Auto);

  // Same, restricted to rows with allowed[row] != 0 (size() entries, shared by all queries):
  // the exact filtered scan of HNSWIndex::search_exact (hnsw.cpp:336-354).
  void search(const float* queries, size_t nq, size_t k, uint32_t* out_ids, float* out_dists,
              const uint8_t* allowed, GpuMethod method = GpuMethod::Auto);

  // Batched filtered scan: query i returns ks[i] <= kmax results among rows
  // allow_ids[range[2i] .. range[2i+1]) (strictly increasing ids); n_ids = len(allow_ids).
  // Outputs nq x kmax. Queries with identical ranges share their row reads.
  void search_lists(const float* queries, size_t nq, const uint32_t* ks, size_t kmax,
                    const uint64_t* range, const uint32_t* allow_ids, size_t n_ids,
                    uint32_t* out_ids, float* out_dists);

  size_t size() const;
  size_t dim() const;
  Metric metric() const;

  vsg_index* handle() const { return h_; }

 private:
  vsg_index* h_ = nullptr;
  size_t dim_;
  Metric metric_;
};

}  // namespace vecsearch
