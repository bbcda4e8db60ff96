/* CPU oracle for the exact k-NN path. TEST INFRASTRUCTURE ONLY (tests, smoke check, CPU
 * baseline); never linked into the product. Each function cites the reference lines it restates.
 * Build: see oracle/Makefile (gcc -O2 -march=x86-64-v3 -ffp-contract=off; every fused
 * multiply-add below is an explicit fmaf, so the compiler cannot add or remove one). */
#include "vs_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* distance.h:18-27 (hsum256 after _mm256_add_ps(acc0, acc1)). acc[0..7] is acc0's lanes,
 * acc[8..15] is acc1's lanes. v_j = acc0_j + acc1_j; the 128-bit halves are added
 * (w_j = v_j + v_{j+4}); then movehdup/movehl/add_ss give (w0 + w1) + (w2 + w3). */
static float tree16(const float acc[16]) {
  float v[8], w[4];
  for (int j = 0; j < 8; ++j) v[j] = acc[j] + acc[j + 8];
  for (int j = 0; j < 4; ++j) w[j] = v[j] + v[j + 4];
  return (w[0] + w[1]) + (w[2] + w[3]);
}

/* distance.h:54-69. Lane j of the two accumulators sums dims j, j+16, ... (acc0 for j < 8,
 * acc1 for j >= 8); an 8-dim step adds into acc0 only; then the tree; then the scalar tail of
 * d % 8 dims (distance.h:68). The tail's bits are set by the compiler, not the source: GCC 11.4
 * at the reference's release flags (-std=gnu++17 -O3 -march=x86-64-v3) turns
 * `r += a[i] * b[i]` into an in-order vector reduction over one 4-wide block (products rounded,
 * then added one by one, left to right) and contracts the remaining < 4 terms into fused
 * multiply-adds. The same holds for distances4 (brute_force.cpp:56-61). Pinned bit for bit
 * against oracle/_ref (the reference sources built by oracle/Makefile) for d % 8 = 1..7. */
float vso_dot(const float* a, const float* b, size_t d) {
  float acc[16];
  size_t i = 0;
  for (int j = 0; j < 16; ++j) acc[j] = 0.f;
  for (; i + 16 <= d; i += 16)
    for (int j = 0; j < 16; ++j) acc[j] = fmaf(a[i + j], b[i + j], acc[j]);
  for (; i + 8 <= d; i += 8)
    for (int j = 0; j < 8; ++j) acc[j] = fmaf(a[i + j], b[i + j], acc[j]);
  float r = tree16(acc);
  if (i + 4 <= d) /* scalar tail, part 1: rounded products added in order */
    for (int j = 0; j < 4; ++j, ++i) r = r + a[i] * b[i];
  for (; i < d; ++i) r = fmaf(a[i], b[i], r); /* part 2: fused */
  return r;
}

/* distance.h:30-52: the same structure on e = a - b with acc = fma(e, e, acc). */
float vso_l2sq(const float* a, const float* b, size_t d) {
  float acc[16];
  size_t i = 0;
  for (int j = 0; j < 16; ++j) acc[j] = 0.f;
  for (; i + 16 <= d; i += 16)
    for (int j = 0; j < 16; ++j) {
      const float e = a[i + j] - b[i + j];
      acc[j] = fmaf(e, e, acc[j]);
    }
  for (; i + 8 <= d; i += 8)
    for (int j = 0; j < 8; ++j) {
      const float e = a[i + j] - b[i + j];
      acc[j] = fmaf(e, e, acc[j]);
    }
  float r = tree16(acc);
  if (i + 4 <= d)
    for (int j = 0; j < 4; ++j, ++i) {
      const float e = a[i] - b[i];
      r = r + e * e;
    }
  for (; i < d; ++i) {
    const float e = a[i] - b[i];
    r = fmaf(e, e, r);
  }
  return r;
}

/* distance.h:71-73 */
float vso_distance(int metric, const float* a, const float* b, size_t d) {
  return metric == VSO_METRIC_L2 ? vso_l2sq(a, b, d) : 1.0f - vso_dot(a, b, d);
}

double vso_dot_f64(const float* a, const float* b, size_t d) {
  double r = 0;
  for (size_t i = 0; i < d; ++i) r += (double)a[i] * (double)b[i];
  return r;
}

double vso_l2sq_f64(const float* a, const float* b, size_t d) {
  double r = 0;
  for (size_t i = 0; i < d; ++i) {
    const double e = (double)a[i] - (double)b[i];
    r += e * e;
  }
  return r;
}

/* ---- top-k: bounded max-heap under the (dist, id) order ----------------------------------
 * brute_force.cpp:76-86 keeps the k smallest with a strict `<` on dist and merges thread lists
 * by (dist, id) (:152-163). Ordering the heap by (dist, id) gives the same result whenever no
 * two rows tie exactly at the k-th distance, and breaks exact ties by the lower id. */
typedef struct {
  float d;
  uint32_t id;
} nb_t;

static int nb_less(nb_t a, nb_t b) { return a.d < b.d || (a.d == b.d && a.id < b.id); }

static void heap_push(nb_t* h, size_t* n, size_t k, nb_t c) {
  if (*n < k) {
    size_t i = (*n)++;
    h[i] = c;
    while (i > 0) { /* sift up: the root holds the largest */
      size_t p = (i - 1) / 2;
      if (!nb_less(h[p], h[i])) break;
      nb_t t = h[p]; h[p] = h[i]; h[i] = t;
      i = p;
    }
  } else if (k > 0 && nb_less(c, h[0])) {
    size_t i = 0;
    h[0] = c;
    for (;;) { /* sift down */
      size_t l = 2 * i + 1, r = l + 1, m = i;
      if (l < k && nb_less(h[m], h[l])) m = l;
      if (r < k && nb_less(h[m], h[r])) m = r;
      if (m == i) break;
      nb_t t = h[m]; h[m] = h[i]; h[i] = t;
      i = m;
    }
  }
}

static int nb_cmp(const void* pa, const void* pb) {
  const nb_t a = *(const nb_t*)pa, b = *(const nb_t*)pb;
  return nb_less(a, b) ? -1 : nb_less(b, a) ? 1 : 0;
}

typedef struct {
  const float* data;
  size_t d;
  int metric;
  const float* queries;
  size_t nq, k, kstride;
  const uint8_t* mask;     /* optional */
  const uint32_t* ks;      /* optional per-query k (lists mode) */
  const uint64_t* off;     /* lists mode: [begin, end) per query */
  const uint32_t* ids;     /* lists mode */
  size_t n;                /* lists mode: rows of an unfiltered query */
  size_t lo, hi;           /* row range (batch / masked mode) or query range (lists mode) */
  nb_t* heaps;             /* nq x k for this thread (batch / masked mode) */
  size_t* sizes;           /* nq (batch / masked mode) */
  uint32_t* out_ids;       /* lists mode: nq x kstride */
  float* out_dists;
} job_t;

/* brute_force.cpp:127-143: a thread scans its own row range for every query. */
static void* scan_rows(void* arg) {
  job_t* j = (job_t*)arg;
  for (size_t i = j->lo; i < j->hi; ++i) {
    if (j->mask && !j->mask[i]) continue;
    const float* x = j->data + i * j->d;
    for (size_t q = 0; q < j->nq; ++q) {
      nb_t c = {vso_distance(j->metric, j->queries + q * j->d, x, j->d), (uint32_t)i};
      heap_push(j->heaps + q * j->k, &j->sizes[q], j->k, c);
    }
  }
  return NULL;
}

static void run_threads(job_t* jobs, int nt, void* (*fn)(void*)) {
  if (nt == 1) {
    fn(&jobs[0]);
    return;
  }
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nt);
  for (int t = 0; t < nt; ++t) pthread_create(&th[t], NULL, fn, &jobs[t]);
  for (int t = 0; t < nt; ++t) pthread_join(th[t], NULL);
  free(th);
}

/* brute_force.cpp:116-169 (and, with a mask, hnsw.cpp:336-354 per query). */
static void search_rows(const float* data, size_t n, size_t d, int metric, const float* queries,
                        size_t nq, size_t k, const uint8_t* mask, uint32_t* out_ids,
                        float* out_dists, int nthreads) {
  if (nq == 0 || k == 0) return;
  size_t nt = nthreads < 1 ? 1 : (size_t)nthreads;
  if (nt > n) nt = n ? n : 1;
  nb_t* heaps = (nb_t*)malloc(sizeof(nb_t) * nt * nq * k);
  size_t* sizes = (size_t*)calloc(nt * nq, sizeof(size_t));
  job_t* jobs = (job_t*)calloc(nt, sizeof(job_t));
  for (size_t t = 0; t < nt; ++t) {
    job_t* j = &jobs[t];
    j->data = data; j->d = d; j->metric = metric; j->queries = queries; j->nq = nq; j->k = k;
    j->mask = mask;
    j->lo = n * t / nt; j->hi = n * (t + 1) / nt;
    j->heaps = heaps + t * nq * k; j->sizes = sizes + t * nq;
  }
  run_threads(jobs, (int)nt, scan_rows);
  nb_t* all = (nb_t*)malloc(sizeof(nb_t) * nt * k);
  for (size_t q = 0; q < nq; ++q) {
    size_t m = 0;
    for (size_t t = 0; t < nt; ++t) {
      memcpy(all + m, heaps + (t * nq + q) * k, sizeof(nb_t) * sizes[t * nq + q]);
      m += sizes[t * nq + q];
    }
    qsort(all, m, sizeof(nb_t), nb_cmp);
    for (size_t c = 0; c < k; ++c) {
      out_ids[q * k + c] = c < m ? all[c].id : VSO_NONE;
      out_dists[q * k + c] = c < m ? all[c].d : INFINITY;
    }
  }
  free(all); free(jobs); free(sizes); free(heaps);
}

void vso_search_batch(const float* data, size_t n, size_t d, int metric, const float* queries,
                      size_t nq, size_t k, uint32_t* out_ids, float* out_dists, int nthreads) {
  search_rows(data, n, d, metric, queries, nq, k, NULL, out_ids, out_dists, nthreads);
}

void vso_search_masked(const float* data, size_t n, size_t d, int metric, const float* queries,
                       size_t nq, size_t k, const uint8_t* mask, uint32_t* out_ids,
                       float* out_dists, int nthreads) {
  search_rows(data, n, d, metric, queries, nq, k, mask, out_ids, out_dists, nthreads);
}

/* Lists mode: a thread takes a range of queries; each query scans its own allowed ids in
 * increasing order (hnsw.cpp:339-351 walks i = 0..n-1 and skips disallowed rows). */
static void* scan_lists(void* arg) {
  job_t* j = (job_t*)arg;
  nb_t* h = (nb_t*)malloc(sizeof(nb_t) * (j->k ? j->k : 1));
  for (size_t q = j->lo; q < j->hi; ++q) {
    const size_t kq = j->ks[q] < j->kstride ? j->ks[q] : j->kstride;
    size_t m = 0;
    const float* qv = j->queries + q * j->d;
    /* a range beginning with UINT64_MAX is unfiltered: every row, in id order */
    const int all = j->off[2 * q] == UINT64_MAX;
    const uint64_t pb = all ? 0 : j->off[2 * q], pe = all ? j->n : j->off[2 * q + 1];
    for (uint64_t p = pb; p < pe; ++p) {
      const uint32_t id = all ? (uint32_t)p : j->ids[p];
      nb_t c = {vso_distance(j->metric, qv, j->data + (size_t)id * j->d, j->d), id};
      heap_push(h, &m, kq, c);
    }
    qsort(h, m, sizeof(nb_t), nb_cmp);
    for (size_t c = 0; c < j->kstride; ++c) {
      const int have = c < m;
      j->out_ids[q * j->kstride + c] = have ? h[c].id : VSO_NONE;
      j->out_dists[q * j->kstride + c] = have ? h[c].d : INFINITY;
    }
  }
  free(h);
  return NULL;
}

void vso_search_lists(const float* data, size_t n, size_t d, int metric, const float* queries,
                      size_t nq, const uint32_t* ks, size_t kmax, const uint64_t* off,
                      const uint32_t* ids, uint32_t* out_ids, float* out_dists, int nthreads) {
  /* off: 2 * nq range bounds */
  if (nq == 0 || kmax == 0) return;
  size_t nt = nthreads < 1 ? 1 : (size_t)nthreads;
  if (nt > nq) nt = nq;
  job_t* jobs = (job_t*)calloc(nt, sizeof(job_t));
  for (size_t t = 0; t < nt; ++t) {
    job_t* j = &jobs[t];
    j->data = data; j->d = d; j->metric = metric; j->queries = queries; j->nq = nq;
    j->k = kmax; j->kstride = kmax; j->ks = ks; j->off = off; j->ids = ids; j->n = n;
    j->lo = nq * t / nt; j->hi = nq * (t + 1) / nt;
    j->out_ids = out_ids; j->out_dists = out_dists;
  }
  run_threads(jobs, (int)nt, scan_lists);
  free(jobs);
}
