// Certified shortlist for large batches: a low-precision MFMA scan nominates candidates, the
// exact reference-order FP32 distance ranks them. Ids and distances equal the exact scan's.
//
// Why it is exact. For a query q let d(x) be the reference CPU distance of row x (the value the
// exact path returns) and dh(x) the MFMA estimate, with |dh(x) - d(x)| <= E for every row (E is
// the stated bound below). Let Dk be the k-th smallest dh over all rows. The k rows with
// dh <= Dk have d <= Dk + E, so the k-th smallest d is <= Dk + E, and every row of the true
// top-k (by d, then id) has d <= Dk + E, hence dh <= Dk + 2E. So the candidate set
// C = {x : dh(x) <= Dk + 2E} contains the answer, and ranking C by exact (d, id) returns it.
//
// How C is found without a per-query top-k inside the GEMM:
//   pass 1  tile minima: for every query and every tile of kTileRows rows, the smallest dh;
//   select  U = the k-th smallest tile minimum (k distinct rows have dh <= U, so Dk <= U);
//   pass 2  emit every row with dh <= U + 2E into the query's candidate buffer;
//   rescore Dk is the k-th smallest dh among the emitted rows (they include every row with
//           dh <= Dk); keep dh <= Dk + 2E, compute the exact distance of each, rank (d, id).
// A query whose candidates overflow the buffer is reported back and answered by the exact
// scan instead.
//
// Error bound E (inner product; all norms computed in FP64 when the copies are built):
//   |sum q x - sum q^ x^| <= |q| |x - x^| + |q - q^| |x^|           (exact identity + C-S)
//   MFMA accumulation (products of bf16 / fp8 are exact in fp32)   <= 2 gamma_d |q^| |x^|
//   FP8 per-vector scales: two extra fp32 roundings                 <= 4 u |q^| |x^|
//   reference CPU order vs the exact sum                            <= gamma_m |q| |x|
//   the final 1 - s roundings on both sides                         <= 4 u (1 + |q| |x|)
// with gamma_n = n u / (1 - n u), u = 2^-24, m = ceil(d/16) + 12, and the row terms taken as
// maxima over the base (|x|max, |x^|max, |x - x^|max). E is padded by 1 ulp-scale cushion.
//
// Data layouts (built from the row-major base when rows are added):
//   BF16 copy: for row block rb (16 rows) and k-step s (32 dims) a 1 KB chunk; lane l of the
//     chunk holds rows rb*16 + l%16, dims 32s + 8*(l/16) .. +7 -- exactly the B operand of
//     v_mfma_f32_16x16x32_bf16, so a wave loads one chunk with one coalesced 16-byte load.
//   FP8 copy: the same with e4m3 values (8 bytes per lane, 512 B per chunk) of x / s_x, where
//     s_x = max|x_i| / 448 is the row's fp32 scale (the B operand of ..._16x16x32_fp8_fp8).
//   Queries of a batch are converted into the same layouts (A operands).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "vsg_internal.h"
#include "vsg_kernels.h"

namespace vsg {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8_t;
typedef __attribute__((ext_vector_type(4))) float f32x4m_t;

constexpr int kTileRows = 256;  // rows per tile minimum (pass 1)

// ---- conversion into fragment layouts ----------------------------------------------------

__device__ __forceinline__ uint8_t to_e4m3(float v) {
  return uint8_t(__builtin_amdgcn_cvt_pk_fp8_f32(v, 0.f, 0, false) & 0xFF);
}
__device__ __forceinline__ float from_e4m3(uint8_t b) {
  return __builtin_amdgcn_cvt_f32_fp8(int(b), 0);
}

// One wave per row (or query): writes the row's fragment slots and returns, in lane 0,
// |x|^2, |x^|^2 and |x - x^|^2 (FP64) and for FP8 the scale.
// src: row-major fp32, stride ld, d valid dims; rows [r0, r1); KS k-steps of 32 dims.
// prec 1 = bf16 -> out16, prec 2 = fp8 -> out8 + scale.
__global__ __launch_bounds__(256) void quantize_kernel(
    const float* __restrict__ src, uint32_t ld, uint32_t d, uint32_t r0, uint32_t r1, uint32_t KS,
    int prec, __bf16* __restrict__ out16, uint8_t* __restrict__ out8, float* __restrict__ scale,
    double* __restrict__ norms /* per row: |x|^2, |x^|^2, |x - x^|^2 */) {
  const int lane = threadIdx.x & 63;
  const uint32_t r = r0 + blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= r1) return;
  const float* x = src + size_t(r) * ld;
  float sx = 1.f;
  if (prec == 2) {
    float mx = 0.f;
    for (uint32_t j = lane; j < d; j += 64) mx = fmaxf(mx, fabsf(x[j]));
    for (int off = 32; off >= 1; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
    sx = mx > 0.f ? mx / 448.f : 1.f;
  }
  double s_x = 0, s_h = 0, s_e = 0;
  const uint32_t groups = KS * 4;  // 8-dim groups
  for (uint32_t gi = lane; gi < groups; gi += 64) {
    const uint32_t s = gi >> 2, g = gi & 3;
    const size_t chunk = size_t(r >> 4) * KS + s;
    const uint32_t slot = g * 16 + (r & 15);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const uint32_t dim = 32 * s + 8 * g + j;
      const float v = dim < d ? x[dim] : 0.f;
      float h;
      if (prec == 1) {
        const __bf16 b = (__bf16)v;
        out16[chunk * 512 + slot * 8 + j] = b;
        h = float(b);
      } else {
        const uint8_t b = to_e4m3(v / sx);
        out8[chunk * 512 + slot * 8 + j] = b;
        h = from_e4m3(b) * sx;
      }
      s_x += double(v) * v;
      s_h += double(h) * h;
      s_e += (double(v) - h) * (double(v) - h);
    }
  }
  for (int off = 32; off >= 1; off >>= 1) {
    s_x += __shfl_xor(s_x, off, 64);
    s_h += __shfl_xor(s_h, off, 64);
    s_e += __shfl_xor(s_e, off, 64);
  }
  if (lane == 0) {
    norms[size_t(r) * 3 + 0] = s_x;
    norms[size_t(r) * 3 + 1] = s_h;
    norms[size_t(r) * 3 + 2] = s_e;
    if (prec == 2) scale[r] = sx;
  }
}

// Max of each norm column over rows [r0, r1) into maxes[0..2] (as sqrt, FP64 bit patterns of
// non-negative values order as unsigned integers).
__global__ void norm_max_kernel(const double* __restrict__ norms, uint32_t r0, uint32_t r1,
                                unsigned long long* __restrict__ maxes) {
  double m[3] = {0, 0, 0};
  for (uint32_t r = r0 + blockIdx.x * blockDim.x + threadIdx.x; r < r1; r += gridDim.x * blockDim.x)
    for (int c = 0; c < 3; ++c) m[c] = fmax(m[c], sqrt(norms[size_t(r) * 3 + c]));
  for (int c = 0; c < 3; ++c) {
    for (int off = 32; off >= 1; off >>= 1) m[c] = fmax(m[c], __shfl_xor(m[c], off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(maxes + c, (unsigned long long)__double_as_longlong(m[c]));
  }
}

// Per-query error bound E (see the header) from the query norms and the base maxima.
__global__ void query_bound_kernel(const double* __restrict__ qnorms, uint32_t nq, int prec,
                                   uint32_t d, const unsigned long long* __restrict__ maxes,
                                   float* __restrict__ E) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  const double X = __longlong_as_double((long long)maxes[0]);
  const double Xh = __longlong_as_double((long long)maxes[1]);
  const double dX = __longlong_as_double((long long)maxes[2]);
  const double nq_ = sqrt(qnorms[size_t(q) * 3 + 0]);
  const double nqh = sqrt(qnorms[size_t(q) * 3 + 1]);
  const double dq = sqrt(qnorms[size_t(q) * 3 + 2]);
  const double u = 1.0 / 16777216.0;
  const double dd = double(d), m = ceil(dd / 16.0) + 12.0;  // d: the padded width 32 * KS
  const double gd = dd * u / (1 - dd * u), gm = m * u / (1 - m * u);
  double e = nq_ * dX + dq * Xh;
  e += 2.0 * gd * nqh * Xh;
  if (prec == 2) e += 4.0 * u * nqh * Xh;
  e += gm * nq_ * X;
  e += 4.0 * u * (1.0 + nq_ * X);
  E[q] = float(e * (1.0 + 1e-6) + 1e-12);
}

// ---- the MFMA scan -----------------------------------------------------------------------
// A wave owns a task (64 queries = 4 query fragments of 16, one row segment). PASS 1 writes
// tile minima; PASS 2 emits candidates. KS: 32-dim k-steps (template: fully unrolled, the A
// operands of all k-steps live in registers).
template <int PREC>
struct Frag;
template <>
struct Frag<1> {
  typedef bf16x8_t T;
  static constexpr int kChunkBytes = 1024;
};
template <>
struct Frag<2> {
  typedef long T;
  static constexpr int kChunkBytes = 512;
};

template <int PREC>
__device__ __forceinline__ f32x4m_t mfma(typename Frag<PREC>::T a, typename Frag<PREC>::T b, f32x4m_t c);
template <>
__device__ __forceinline__ f32x4m_t mfma<1>(bf16x8_t a, bf16x8_t b, f32x4m_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
template <>
__device__ __forceinline__ f32x4m_t mfma<2>(long a, long b, f32x4m_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a, b, c, 0, 0, 0);
}

constexpr int kQF = 4;  // query fragments per wave (64 queries)

template <int PREC, int KS, int PASS>
__global__ __launch_bounds__(256) void mfma_scan_kernel(
    const void* __restrict__ xfrag, const float* __restrict__ xscale, uint32_t n,
    const void* __restrict__ qfrag, const float* __restrict__ qscale, uint32_t nq,
    uint32_t n_qblocks, uint32_t seg_rows, uint32_t n_tasks, uint32_t n_tiles,
    uint32_t* __restrict__ tilemin, const float* __restrict__ thr, uint64_t* __restrict__ cand,
    uint32_t cap, uint32_t* __restrict__ count) {
  typedef typename Frag<PREC>::T FT;
  const int lane = threadIdx.x & 63;
  const uint32_t task = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (task >= n_tasks) return;
  const uint32_t qb = task % n_qblocks, seg = task / n_qblocks;
  const uint32_t row0 = seg * seg_rows, row1 = min(n, row0 + seg_rows);
  if (row0 >= row1) return;
  const FT* xf = reinterpret_cast<const FT*>(xfrag);
  const FT* qf = reinterpret_cast<const FT*>(qfrag);

  // A operands: 4 query fragments (queries qb*64 + f*16 + ..) for every k-step
  FT a[kQF][KS];
#pragma unroll
  for (int f = 0; f < kQF; ++f)
#pragma unroll
    for (int s = 0; s < KS; ++s) a[f][s] = qf[(size_t(qb * kQF + f) * KS + s) * 64 + lane];

  // per-lane query indices: value i of fragment f is query qb*64 + f*16 + 4*(lane/16) + i
  const uint32_t qbase = qb * 64 + 4 * (lane >> 4);
  float qs[kQF][4], th[kQF][4], run[kQF][4];
#pragma unroll
  for (int f = 0; f < kQF; ++f)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t q = qbase + f * 16 + i;
      qs[f][i] = (PREC == 2 && q < nq) ? qscale[q] : 1.f;
      th[f][i] = (PASS == 2 && q < nq) ? thr[q] : -__builtin_inff();
      run[f][i] = __builtin_inff();
    }

  for (uint32_t r = row0; r < row1; r += 16) {
    const uint32_t rb = r >> 4;
    FT b[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) b[s] = xf[(size_t(rb) * KS + s) * 64 + lane];
    f32x4m_t acc[kQF];
#pragma unroll
    for (int f = 0; f < kQF; ++f) acc[f] = f32x4m_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
      for (int f = 0; f < kQF; ++f) acc[f] = mfma<PREC>(a[f][s], b[s], acc[f]);

    const uint32_t row = r + (lane & 15);
    const bool ok = row < row1;
    const float xs = PREC == 2 ? (ok ? xscale[row] : 1.f) : 1.f;
#pragma unroll
    for (int f = 0; f < kQF; ++f)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float s = PREC == 2 ? acc[f][i] * qs[f][i] * xs : acc[f][i];
        const float dh = ok ? 1.0f - s : __builtin_inff();
        if (PASS == 1) {
          run[f][i] = fminf(run[f][i], dh);
        } else if (dh <= th[f][i]) {
          const uint32_t q = qbase + f * 16 + i;
          const uint32_t slot = atomicAdd(count + q, 1u);
          if (slot < cap) cand[size_t(q) * cap + slot] = make_key(dh, row);
        }
      }

    if (PASS == 1 && ((r + 16 - row0) % kTileRows == 0 || r + 16 >= row1)) {
      // tile done: minimum over the 16 row lanes of each lane group, one writer per group
      const uint32_t tile = r / kTileRows;
#pragma unroll
      for (int f = 0; f < kQF; ++f)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          float v = run[f][i];
          v = fminf(v, __shfl_xor(v, 1, 64));
          v = fminf(v, __shfl_xor(v, 2, 64));
          v = fminf(v, __shfl_xor(v, 4, 64));
          v = fminf(v, __shfl_xor(v, 8, 64));
          const uint32_t q = qbase + f * 16 + i;
          if ((lane & 15) == 0 && q < nq) tilemin[size_t(q) * n_tiles + tile] = ord_bits(v);
          run[f][i] = __builtin_inff();
        }
    }
  }
}

// ---- threshold select --------------------------------------------------------------------
// Per query (one 256-thread block): U = the kq-th smallest tile minimum (radix select on the
// order-preserving bits), threshold = U + 2E. With fewer tiles than kq the threshold is +inf.
__global__ __launch_bounds__(256) void select_threshold_kernel(
    const uint32_t* __restrict__ tilemin, uint32_t n_tiles, const uint32_t* __restrict__ ks,
    uint32_t k, const float* __restrict__ E, float* __restrict__ thr) {
  __shared__ uint32_t hist[256];
  __shared__ uint32_t sel[2];
  const uint32_t q = blockIdx.x, tid = threadIdx.x;
  const uint32_t kq = ks ? ks[q] : k;
  const uint32_t* v = tilemin + size_t(q) * n_tiles;
  if (kq == 0 || kq > n_tiles) {
    if (tid == 0) thr[q] = __builtin_inff();
    return;
  }
  uint32_t prefix = 0, pmask = 0, rank = kq;  // 1-based rank within the current prefix
  for (int shift = 24; shift >= 0; shift -= 8) {
    hist[tid] = 0;
    __syncthreads();
    for (uint32_t i = tid; i < n_tiles; i += 256) {
      const uint32_t x = v[i];
      if ((x & pmask) == prefix) atomicAdd(&hist[(x >> shift) & 255], 1u);
    }
    __syncthreads();
    if (tid == 0) {
      uint32_t acc = 0, b = 0;
      for (; b < 256; ++b) {
        if (acc + hist[b] >= rank) break;
        acc += hist[b];
      }
      sel[0] = b;
      sel[1] = rank - acc;
    }
    __syncthreads();
    prefix |= sel[0] << shift;
    pmask |= 255u << shift;
    rank = sel[1];
    __syncthreads();
  }
  if (tid == 0) {
    const float U = unord_bits(prefix);
    // absolute cushion 2^-20 covers the rounding of this sum
    thr[q] = U + 2.0f * E[q] * (1.0f + 1e-6f) + 9.5367431640625e-07f;
  }
}

// ---- rescore -----------------------------------------------------------------------------
constexpr int kCandCap = 2048;

template <int D16, bool L2>
__global__ __launch_bounds__(256) void rescore_kernel(
    const float* __restrict__ base, uint32_t dp, uint32_t d, const float* __restrict__ queries,
    const uint32_t* __restrict__ ks, uint32_t k, const float* __restrict__ E,
    const uint64_t* __restrict__ cand, const uint32_t* __restrict__ count,
    uint32_t* __restrict__ out_ids, float* __restrict__ out_dists, uint32_t ldo,
    uint32_t* __restrict__ overflow_flags, uint32_t* __restrict__ n_overflow,
    uint32_t* __restrict__ rescored) {
  __shared__ uint64_t cbuf[kCandCap];
  __shared__ uint64_t ebuf[kCandCap];
  __shared__ uint32_t nsel;
  extern __shared__ float q_lds[];  // dp floats
  const uint32_t q = blockIdx.x, tid = threadIdx.x, lane = tid & 63, t = lane & 3, g = lane >> 2;
  const uint32_t kq = min(ks ? ks[q] : k, ldo);
  uint32_t* oi = out_ids + size_t(q) * ldo;
  float* od = out_dists + size_t(q) * ldo;
  const uint32_t c = count[q];
  if (c > uint32_t(kCandCap)) {
    if (tid == 0) {
      overflow_flags[q] = 1;
      rescored[q] = 0;
      atomicAdd(n_overflow, 1u);
    }
    return;
  }
  if (tid == 0) overflow_flags[q] = 0;
  for (uint32_t i = kq + tid; i < ldo; i += 256) {
    oi[i] = kNoneId;
    od[i] = __builtin_inff();
  }
  for (uint32_t i = tid; i < dp; i += 256) q_lds[i] = i < d ? queries[size_t(q) * d + i] : 0.f;
  // candidates sorted by (dh, id)
  uint32_t n2 = 256;
  while (n2 < c) n2 <<= 1;
  for (uint32_t i = tid; i < n2; i += 256) cbuf[i] = i < c ? cand[size_t(q) * kCandCap + i] : kPadKey;
  __syncthreads();
  for (uint32_t size = 2; size <= n2; size <<= 1)
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      for (uint32_t i = tid; i < n2 / 2; i += 256) {
        const uint32_t lo = 2 * i - (i & (stride - 1)), hi = lo + stride;
        const bool up = (lo & size) == 0;
        const uint64_t a = cbuf[lo], b = cbuf[hi];
        if ((a > b) == up) {
          cbuf[lo] = b;
          cbuf[hi] = a;
        }
      }
      __syncthreads();
    }
  // keep dh <= Dk + 2E
  const uint32_t kk = min(kq, c);
  if (tid == 0) nsel = 0;
  __syncthreads();
  if (kk > 0) {
    const float lim = key_dist(cbuf[kk - 1]) + 2.0f * E[q] * (1.0f + 1e-6f) + 9.5367431640625e-07f;
    for (uint32_t i = tid; i < c; i += 256)
      if (key_dist(cbuf[i]) <= lim) atomicMax(&nsel, i + 1);
  }
  __syncthreads();
  const uint32_t r = nsel;  // sorted, so the kept candidates are a prefix
  if (tid == 0) rescored[q] = r;
  // exact distances, 16 rows per wave instruction group, 4 waves
  for (uint32_t base_i = (tid >> 6) * 16; base_i < r; base_i += 64) {
    const uint32_t i = base_i + g;
    bool ok[1] = {i < r};
    const uint32_t row = key_id(cbuf[ok[0] ? i : 0]);
    const float* xr[1] = {base + size_t(row) * dp};
    float dist[1][1];
    quad_dists<1, 1, D16, L2, 16, kLoadPlain>(xr, ok, q_lds, dp, d, int(t), dist);
    if (ok[0] && t == 0) ebuf[i] = make_key(dist[0][0], row);
  }
  __syncthreads();
  // rank the exact keys (unique) and write the first kq
  for (uint32_t i = tid; i < r; i += 256) {
    const uint64_t v = ebuf[i];
    uint32_t rank = 0;
    for (uint32_t j = 0; j < r; ++j) rank += ebuf[j] < v;
    if (rank < kq) {
      oi[rank] = key_id(v);
      od[rank] = key_dist(v);
    }
  }
  for (uint32_t i = r + tid; i < kq; i += 256) {
    oi[i] = kNoneId;
    od[i] = __builtin_inff();
  }
}

__global__ void zero_u32_kernel(uint32_t* __restrict__ p, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    p[i] = 0;
}

// ---- host launchers ----------------------------------------------------------------------

hipError_t launch_quantize(hipStream_t st, const float* src, uint32_t ld, uint32_t d, uint32_t r0,
                           uint32_t r1, uint32_t KS, int prec, void* out, float* scale,
                           double* norms) {
  if (r1 <= r0) return hipSuccess;
  const unsigned blocks = (r1 - r0 + 3) / 4;
  hipLaunchKernelGGL(quantize_kernel, dim3(blocks), dim3(256), 0, st, src, ld, d, r0, r1, KS, prec,
                     prec == 1 ? static_cast<__bf16*>(out) : nullptr,
                     prec == 2 ? static_cast<uint8_t*>(out) : nullptr, scale, norms);
  return hipGetLastError();
}

hipError_t launch_norm_max(hipStream_t st, const double* norms, uint32_t r0, uint32_t r1,
                           unsigned long long* maxes) {
  if (r1 <= r0) return hipSuccess;
  const unsigned blocks = std::min<unsigned>(1024, (r1 - r0 + 255) / 256);
  hipLaunchKernelGGL(norm_max_kernel, dim3(blocks), dim3(256), 0, st, norms, r0, r1, maxes);
  return hipGetLastError();
}

hipError_t launch_query_bound(hipStream_t st, const double* qnorms, uint32_t nq, int prec, uint32_t d,
                              const unsigned long long* maxes, float* E) {
  hipLaunchKernelGGL(query_bound_kernel, dim3((nq + 255) / 256), dim3(256), 0, st, qnorms, nq, prec, d,
                     maxes, E);
  return hipGetLastError();
}

hipError_t launch_zero_u32(hipStream_t st, uint32_t* p, size_t n) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(zero_u32_kernel, dim3(unsigned(std::min<size_t>(1024, (n + 255) / 256))), dim3(256), 0,
                     st, p, n);
  return hipGetLastError();
}

namespace {
template <int PREC, int KS>
hipError_t launch_mfma_ks(int pass, const ShortScanArgs& a) {
  const dim3 grid((a.n_tasks + 3) / 4);
  if (pass == 1)
    hipLaunchKernelGGL((mfma_scan_kernel<PREC, KS, 1>), grid, dim3(256), 0, a.st, a.xfrag, a.xscale, a.n,
                       a.qfrag, a.qscale, a.nq, a.n_qblocks, a.seg_rows, a.n_tasks, a.n_tiles, a.tilemin,
                       a.thr, a.cand, a.cap, a.count);
  else
    hipLaunchKernelGGL((mfma_scan_kernel<PREC, KS, 2>), grid, dim3(256), 0, a.st, a.xfrag, a.xscale, a.n,
                       a.qfrag, a.qscale, a.nq, a.n_qblocks, a.seg_rows, a.n_tasks, a.n_tiles, a.tilemin,
                       a.thr, a.cand, a.cap, a.count);
  return hipGetLastError();
}

template <int PREC>
hipError_t launch_mfma_p(int pass, const ShortScanArgs& a) {
  switch (a.KS) {
    case 2: return launch_mfma_ks<PREC, 2>(pass, a);
    case 4: return launch_mfma_ks<PREC, 4>(pass, a);
    case 8: return launch_mfma_ks<PREC, 8>(pass, a);
    case 12: return launch_mfma_ks<PREC, 12>(pass, a);
    default: return hipErrorInvalidValue;
  }
}
}  // namespace

int short_ks(uint32_t d) {
  const uint32_t ks = (d + 31) / 32;
  if (ks <= 2) return 2;
  if (ks <= 4) return 4;
  if (ks <= 8) return 8;
  if (ks <= 12) return 12;
  return 0;  // too wide for the register-resident query operands
}

hipError_t launch_mfma_scan(int prec, int pass, const ShortScanArgs& a) {
  return prec == 1 ? launch_mfma_p<1>(pass, a) : launch_mfma_p<2>(pass, a);
}

hipError_t launch_select_threshold(hipStream_t st, const uint32_t* tilemin, uint32_t n_tiles, uint32_t nq,
                                   const uint32_t* ks, uint32_t k, const float* E, float* thr) {
  hipLaunchKernelGGL(select_threshold_kernel, dim3(nq), dim3(256), 0, st, tilemin, n_tiles, ks, k, E, thr);
  return hipGetLastError();
}

hipError_t launch_rescore(hipStream_t st, bool l2, const float* base, uint32_t dp, uint32_t d,
                          const float* queries, uint32_t nq, const uint32_t* ks, uint32_t k,
                          const float* E, const uint64_t* cand, const uint32_t* count, uint32_t* ids,
                          float* dists, uint32_t ldo, uint32_t* flags, uint32_t* n_overflow,
                          uint32_t* rescored) {
  const size_t lds = size_t(dp) * sizeof(float);
#define VSG_RESCORE(D16, L2V)                                                                        \
  hipLaunchKernelGGL((rescore_kernel<D16, L2V>), dim3(nq), dim3(256), lds, st, base, dp, d, queries, ks, \
                     k, E, cand, count, ids, dists, ldo, flags, n_overflow, rescored)
  if (l2) {
    if (d == 384) VSG_RESCORE(24, true);
    else if (d == 128) VSG_RESCORE(8, true);
    else VSG_RESCORE(0, true);
  } else {
    if (d == 384) VSG_RESCORE(24, false);
    else if (d == 128) VSG_RESCORE(8, false);
    else VSG_RESCORE(0, false);
  }
#undef VSG_RESCORE
  return hipGetLastError();
}

int short_cand_cap() { return kCandCap; }
int short_tile_rows() { return kTileRows; }

}  // namespace vsg
