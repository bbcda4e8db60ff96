/* CPU oracle for the exact k-NN path. TEST INFRASTRUCTURE ONLY: the tests, the smoke check and
 * the CPU-baseline leg of the benchmark link this; the product library never does.
 *
 * It restates, in plain C, the arithmetic and result semantics of the reference CPU index:
 *   - distance.h:18-27   hsum256 add tree
 *   - distance.h:30-52   l2_sq   (16 FMA accumulators over 16-dim steps, an 8-dim step, tree,
 *                                 scalar tail)
 *   - distance.h:54-69   dot     (same structure)
 *   - distance.h:71-73   distance = l2_sq, or 1 - dot for inner product
 *   - brute_force.cpp:116-169   search_batch: top-k per query, (dist, id) order, kNone / +inf pad
 *   - hnsw.cpp:322-354          search_exact: the same scan restricted to allowed rows
 * An FMA on one AVX2 lane is a single-rounding fused multiply-add, so scalar fmaf() in the same
 * order gives the same bits as the reference's vector code.
 */
#ifndef VS_ORACLE_H
#define VS_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VSO_METRIC_L2 0
#define VSO_METRIC_IP 1
#define VSO_NONE 0xFFFFFFFFu

float vso_dot(const float* a, const float* b, size_t d);
float vso_l2sq(const float* a, const float* b, size_t d);
float vso_distance(int metric, const float* a, const float* b, size_t d);

/* Exact top-k of nq queries against n rows (row-major, n x d). out_* are nq x k, nearest first,
 * ties to the lower id, missing = VSO_NONE / +inf. nthreads >= 1 splits the rows. */
void vso_search_batch(const float* data, size_t n, size_t d, int metric, const float* queries,
                      size_t nq, size_t k, uint32_t* out_ids, float* out_dists, int nthreads);

/* Same, restricted to rows with mask[i] != 0 (mask has n entries, shared by all queries). */
void vso_search_masked(const float* data, size_t n, size_t d, int metric, const float* queries,
                       size_t nq, size_t k, const uint8_t* mask, uint32_t* out_ids,
                       float* out_dists, int nthreads);

/* Per-query allowed-id lists and per-query k: query i scans rows ids[range[2i] .. range[2i+1])
 * and writes ks[i] results into row i of out_* (stride kmax; columns ks[i]..kmax-1 get the
 * pad). */
void vso_search_lists(const float* data, size_t n, size_t d, int metric, const float* queries,
                      size_t nq, const uint32_t* ks, size_t kmax, const uint64_t* range,
                      const uint32_t* ids, uint32_t* out_ids, float* out_dists, int nthreads);

/* Exact FP64 dot / squared L2 (used to report the FP32 error of returned scores). */
double vso_dot_f64(const float* a, const float* b, size_t d);
double vso_l2sq_f64(const float* a, const float* b, size_t d);

#ifdef __cplusplus
}
#endif
#endif
