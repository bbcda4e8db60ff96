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
//   pass 1  tile minima: for every query and every tile (512 or 256 rows), the smallest dh;
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
//   FP8 only: the FP8 instruction's own sum of its products        <= 2^-10 |q^| |x^|
//     is measured at up to about 2^-12.3 of their absolute sum, both for the 16x16x32
//     instruction (widths of 64) and for the 16x16x128 one (every other width): not the fp32
//     accuracy the term above assumes; 2^-10 of sum |q^_i x^_i| <= 2^-10 |q^| |x^| covers both
//     (tests/gpu_probes/mfma_accuracy.hip, run by the GPU tests, holds each within 2^-11)
//   FP8 per-vector scales: two extra fp32 roundings                 <= 4 u |q^| |x^|
//   reference CPU order vs the exact sum                            <= gamma_m |q| |x|
//   the final 1 - s roundings on both sides                         <= 4 u (1 + |q| |x|)
// with gamma_n = n u / (1 - n u), u = 2^-24, m = ceil(d/16) + 12, and the row terms taken as
// maxima over the base (|x|max, |x^|max, |x - x^|max). E is padded by 1 ulp-scale cushion.
//
// Squared L2 runs through the same passes. With hn(x) = |x|^2 / 2 (per row, rounded to fp32),
// the estimate becomes dh(x) = 1 - (q^.x^ - hn(x)), and the reference distance d(x) maps to
// t(x) = (d(x) - |q|^2 + 2) / 2 = 1 - q.x + |x|^2 / 2 + (d(x) - |q - x|^2) / 2, an increasing
// function of d(x) for the query. The argument above then holds for t with the same dh:
// ranking by t is ranking by (d, id), and |dh - t| <= E_L2, where E_L2 is the inner-product
// bound's dot-product terms plus
//   hn rounded to fp32                                              <= u |x|max^2 / 2
//   the reference's l2_sq order vs the exact sum (halved)           <= gamma_m (|q| + |x|max)^2 / 2
//   the subtraction of hn and the final 1 - s roundings             <= 4 u (1 + |q| |x| + |x|max^2)
// The rows-stationary pass's L2 instance subtracts hn from every score; L2 batches use that pass
// for pass 2 too.
//
// Data layouts (built from the row-major base when rows are added):
//   BF16 copy: for row block rb (16 rows) and k-step s (32 dims) a 1 KB chunk; lane l of the
//     chunk holds rows rb*16 + l%16, dims 32s + 8*(l/16) .. +7 -- exactly the B operand of
//     v_mfma_f32_16x16x32_bf16, so a wave loads one chunk with one coalesced 16-byte load.
//   FP8 copy: the same with e4m3 values (8 bytes per lane, 512 B per chunk) of x / s_x, where
//     s_x = max|x_i| / 448 is the row's fp32 scale; four consecutive k-steps' pieces of a lane form
//     its operand of v_mfma_f32_16x16x128_f8f6f4.
//   Queries of a batch are converted into the same layouts (A operands).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "vsg_internal.h"
#include "vsg_kernels.h"

namespace vsg {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8_t;
typedef __attribute__((ext_vector_type(4))) float f32x4m_t;

// rows per tile of the MFMA scan: 8 waves x RF row fragments x 16 rows (RF = 4 or 2)

// ---- conversion into fragment layouts ----------------------------------------------------

__device__ __forceinline__ uint8_t to_e4m3(float v) {
  return uint8_t(__builtin_amdgcn_cvt_pk_fp8_f32(v, 0.f, 0, false) & 0xFF);
}
__device__ __forceinline__ float from_e4m3(uint8_t b) {
  return __builtin_amdgcn_cvt_f32_fp8(int(b), 0);
}

// One wave per row (or query): writes the row's fragment slots and returns, in every lane,
// |x|^2, |x^|^2 and |x - x^|^2 (FP64) and for FP8 the scale sx (1 for BF16).
// x: the row's d valid fp32 dims; r: its fragment row; KS k-steps of 32 dims.
// prec 1 = bf16 -> out16, prec 2 = fp8 -> out8.
__device__ __forceinline__ void quantize_row(const float* __restrict__ x, uint32_t d, uint32_t r, uint32_t KS,
                                             int prec, __bf16* __restrict__ out16, uint8_t* __restrict__ out8,
                                             double& s_x, double& s_h, double& s_e, float& sx) {
  const int lane = threadIdx.x & 63;
  sx = 1.f;
  if (prec == 2) {
    float mx = 0.f;
    for (uint32_t j = lane; j < d; j += 64) mx = fmaxf(mx, fabsf(x[j]));
    for (int off = 32; off >= 1; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
    sx = mx > 0.f ? mx / 448.f : 1.f;
  }
  s_x = 0, s_h = 0, s_e = 0;
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
}

// Base rows [r0, r1) of a row-major fp32 matrix (stride ld): fragments, per-row norms
// (|x|^2, |x^|^2, |x - x^|^2) and, for FP8, per-row scales.
__global__ __launch_bounds__(256) void quantize_kernel(
    const float* __restrict__ src, uint32_t ld, uint32_t d, uint32_t r0, uint32_t r1, uint32_t KS,
    int prec, __bf16* __restrict__ out16, uint8_t* __restrict__ out8, float* __restrict__ scale,
    double* __restrict__ norms) {
  const uint32_t r = r0 + blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= r1) return;
  double s_x, s_h, s_e;
  float sx;
  quantize_row(src + size_t(r) * ld, d, r, KS, prec, out16, out8, s_x, s_h, s_e, sx);
  if ((threadIdx.x & 63) == 0) {
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

// The aux block of a query fragment (16 queries), kAux floats: [0, 16) the pass-2 score cuts,
// [16, 32) the FP8 query scales (1 for BF16), [32, 48) the folded FP8 cuts of pass 2 (below).
constexpr int kAux = 48;

// Per-query error bound E (see the header) from the query norms and the base maxima.
__device__ __forceinline__ float query_error_bound(double qx2, double qh2, double qe2, int prec, uint32_t d,
                                                   const unsigned long long* __restrict__ maxes, bool l2) {
  const double X = __longlong_as_double((long long)maxes[0]);
  const double Xh = __longlong_as_double((long long)maxes[1]);
  const double dX = __longlong_as_double((long long)maxes[2]);
  const double nq_ = sqrt(qx2);
  const double nqh = sqrt(qh2);
  const double dq = sqrt(qe2);
  const double u = 1.0 / 16777216.0;
  const double dd = double(d), m = ceil(dd / 16.0) + 12.0;  // d: the padded width 32 * KS
  const double gd = dd * u / (1 - dd * u), gm = m * u / (1 - m * u);
  double e = nq_ * dX + dq * Xh;
  e += 2.0 * gd * nqh * Xh;
  if (prec == 2) e += (4.0 * u + 0x1p-10) * nqh * Xh;  // scales; the FP8 instruction's own sum
  if (l2) {
    e += 0.5 * u * X * X;                       // hn in fp32
    e += 0.5 * gm * (nq_ + X) * (nq_ + X);      // the reference's l2_sq order
    e += 4.0 * u * (1.0 + nq_ * X + X * X);     // s - hn and 1 - s roundings
  } else {
    e += gm * nq_ * X;
    e += 4.0 * u * (1.0 + nq_ * X);
  }
  return float(e * (1.0 + 1e-6) + 1e-12);
}

// The first launch of a shortlist round, one wave per query: query fragments, E[q], the FP8
// scale, the scale slots of the pass-1 aux block (pass 1 reads no cuts), and flags[0..1] = 0.
__global__ __launch_bounds__(256) void query_prep_kernel(
    const float* __restrict__ queries, uint32_t d, uint32_t nq, uint32_t KS, int prec,
    __bf16* __restrict__ out16, uint8_t* __restrict__ out8, float* __restrict__ qscale,
    const unsigned long long* __restrict__ maxes, float* __restrict__ E, float* __restrict__ aux,
    uint32_t* __restrict__ flags, int l2) {
  if (blockIdx.x == 0 && threadIdx.x < 2) flags[threadIdx.x] = 0;
  const uint32_t q = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (q >= nq) return;
  double s_x, s_h, s_e;
  float sx;
  quantize_row(queries + size_t(q) * d, d, q, KS, prec, out16, out8, s_x, s_h, s_e, sx);
  if ((threadIdx.x & 63) == 0) {
    E[q] = query_error_bound(s_x, s_h, s_e, prec, 32 * KS, maxes, l2 != 0);
    qscale[q] = sx;
    aux[(q >> 4) * kAux + 16 + (q & 15)] = sx;
  }
}

// ---- the MFMA scan (rows stationary, queries streaming) ---------------------------------
// A block of 8 waves owns one tile of 8 * RF * 16 base rows: wave w keeps rows w*RF*16 .. +RF*16
// of the tile as B operands of every k-step in registers (loaded once from HBM). The batch's
// query fragments (16 queries x all k-steps, small and L2-resident) stream past them through
// a double-buffered LDS ring, kGroup fragments per stage. PASS 1 writes, per query, the tile's
// smallest dh (tilemin[q * n_space + j], tile j of the visited tiles); PASS 2 emits
// candidates with dh <= thr[q]. Pass 1 visits every tstride-th tile.
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

// FP8 over 128 dims in one instruction (v_mfma_f32_16x16x128_f8f6f4, twice the rate of four
// 16x16x32 steps): a lane's 32-byte operand is its 8-byte pieces of four consecutive 32-dim
// k-steps side by side, for A and B alike, so byte j of lane l pairs query and row dim
// 32 (4t + j / 8) + 8 (l / 16) + j % 8 of both; the fragment copies keep their 32-dim layout.
// Formats 0, 0: both operands e4m3; scale operands 0: the unscaled instruction.
typedef __attribute__((ext_vector_type(8))) int i32x8m_t;
typedef __attribute__((ext_vector_type(4))) long i64x4m_t;
__device__ __forceinline__ f32x4m_t mfma_fp8x128(i64x4m_t a, i64x4m_t b, f32x4m_t c) {
  return __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(__builtin_bit_cast(i32x8m_t, a),
                                                          __builtin_bit_cast(i32x8m_t, b), c, 0, 0, 0, 0, 0, 0);
}
// k-steps in fours where the width allows (every FP8 width but 64 dims)
__host__ __device__ constexpr bool fp8_x128(int prec, int ks) { return prec == 2 && ks % 4 == 0; }

constexpr int kWaves = 8;   // waves per block
constexpr int kNBuf = 4;    // LDS stages in the ring
constexpr int kAhead = kNBuf - 1;  // stages in flight ahead of the one being consumed

// Query fragments per LDS stage: the stage's 1 KB pieces must split evenly over the 8 waves
// (every wave then issues the same number of loads per stage, which the vmcnt wait counts).
__host__ __device__ constexpr int stage_group(int prec, int ks) {
  return (ks * (prec == 1 ? 1024 : 512) / 1024) % 8 == 0 ? 1
         : (ks * (prec == 1 ? 1024 : 512) / 1024) % 4 == 0 ? 2
         : (ks * (prec == 1 ? 1024 : 512) / 1024) % 2 == 0 ? 4
                                                            : 8;
}

// row fragments (16 rows each) per wave: the B operands of all k-steps stay in registers
__host__ __device__ constexpr int rows_frags(int prec, int ks) { return ks * (prec == 1 ? 4 : 3) <= 48 ? 4 : 2; }

// Minimum over each group of 16 lanes (a DPP row), left in every lane of the row: two
// quad_perm swaps, then rotations by 4 and 8 within the row.
__device__ __forceinline__ float row16_min(float v) {
  v = fminf(v, dpp_f<kDppXor1>(v));
  v = fminf(v, dpp_f<kDppXor2>(v));
  v = fminf(v, dpp_f<0x124>(v));  // row_ror:4
  v = fminf(v, dpp_f<0x128>(v));  // row_ror:8
  return v;
}

template <int PREC, int KS>
struct StageShape {
  static constexpr int RF = rows_frags(PREC, KS);
  static constexpr int G = stage_group(PREC, KS);  // query fragments per stage
  static constexpr int CB = Frag<PREC>::kChunkBytes;
  static constexpr int FRAG_BYTES = KS * CB;       // one query fragment, all k-steps
  static constexpr int STAGE_Q = G * FRAG_BYTES;   // query operands of a stage
  // the stage's aux blocks (G x kAux floats), rounded up to whole 256-byte wave loads (64 lanes
  // x 4 bytes); the caller pads qaux by short_aux_pad() floats past the last stage
  static constexpr int AUX_BYTES = (G * kAux * 4 + 255) / 256 * 256;
  static constexpr int STAGE_BYTES = STAGE_Q + AUX_BYTES;
  static constexpr int PIECES = STAGE_Q / 1024;    // 1 KB pieces (one wave load each)
  static constexpr int PER_WAVE = PIECES / kWaves; // loads per wave per stage (exact)
  static constexpr int AUX_LOADS = AUX_BYTES / 256;
  static constexpr int LOADS = PER_WAVE + AUX_LOADS;  // vector-memory loads per thread per stage
  static_assert(PIECES % kWaves == 0, "stage pieces must split evenly over the waves");
};

// PASS 2 appends candidates to the wave's own region of `emit` (wave_cap entries, region
// blockIdx.x * 8 + wave), each (ord(dh) << 32) | (query << 6) | row within the wave's 64 rows.
// Slots come from a ballot and a wave-uniform count, and entries leave with plain stores:
// no atomic sits in the MFMA loop (a returning atomic would make the wave wait for every
// outstanding load, the LDS-DMA prefetch included). emit_count[region] receives the number of
// entries wanted (more than wave_cap: the region overflowed). scatter_candidates_kernel then
// files them per query.
template <int PREC, int KS, int PASS, bool L2>
__global__ __launch_bounds__(64 * kWaves) void mfma_scan_kernel(
    const void* __restrict__ xfrag, const float* __restrict__ xscale, const float* __restrict__ xhn, uint32_t n,
    const void* __restrict__ qfrag, const float* __restrict__ qaux, uint32_t nq, uint32_t tstride,
    uint32_t n_space, uint32_t qsplit, uint32_t* __restrict__ tilemin, uint64_t* __restrict__ emit,
    uint32_t wave_cap, uint32_t* __restrict__ emit_count) {
  typedef typename Frag<PREC>::T FT;
  typedef StageShape<PREC, KS> SH;
  constexpr int RF = SH::RF, G = SH::G, LOADS = SH::LOADS;
  // qsplit blocks share a visited tile (pass 1 of a large batch), each taking a contiguous share
  // of the query stages [st0, nstages); pass 2 runs with qsplit = 1
  const uint32_t nst_all = (((nq + 15) >> 4) + G - 1) / G;
  const uint32_t per_part = (nst_all + qsplit - 1) / qsplit;
  const uint32_t st0 = min(nst_all, (blockIdx.x % qsplit) * per_part);
  const uint32_t nstages = min(nst_all, st0 + per_part);
  if (st0 == nstages) return;  // no stage left for this share (block-uniform, before any barrier)
  // Four ring slots as distinct LDS objects: the compiler can then tell the slot being read
  // from the slots still being filled by LDS-DMA loads, and waits only on the counted vmcnt.
  __shared__ __attribute__((aligned(16))) uint8_t slot0[SH::STAGE_BYTES];
  __shared__ __attribute__((aligned(16))) uint8_t slot1[SH::STAGE_BYTES];
  __shared__ __attribute__((aligned(16))) uint8_t slot2[SH::STAGE_BYTES];
  __shared__ __attribute__((aligned(16))) uint8_t slot3[SH::STAGE_BYTES];
  __shared__ float red[2][kWaves * G * 16];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t j = blockIdx.x / qsplit;  // visited tile
  const uint32_t tile = j * tstride;
  const uint32_t nrb = (n + 15) >> 4;
  const uint32_t rb0 = tile * (kWaves * RF) + wave * RF;  // first row block of this wave
  uint64_t* wemit = emit + (size_t(j) * kWaves + wave) * wave_cap;
  uint32_t n_emit = 0;  // wave-uniform

  // this wave's rows, all k-steps (zero for row blocks past the base)
  FT b[RF][KS];
  const FT* xf = reinterpret_cast<const FT*>(xfrag);
#pragma unroll
  for (int f = 0; f < RF; ++f)
#pragma unroll
    for (int s = 0; s < KS; ++s) b[f][s] = (rb0 + f < nrb) ? xf[(size_t(rb0 + f) * KS + s) * 64 + lane] : FT{};
  float xs[RF], hn[RF];  // FP8 row scale; L2: |x|^2 / 2 (xhn), subtracted from every score
  bool ok[RF];
  uint64_t okm[RF];  // wave masks of the rows inside the base
#pragma unroll
  for (int f = 0; f < RF; ++f) {
    const uint32_t row = (rb0 + f) * 16 + (lane & 15);
    ok[f] = row < n;
    okm[f] = ballot(ok[f]);
    xs[f] = (PREC == 2 && ok[f]) ? xscale[row] : 1.f;
    hn[f] = (L2 && ok[f]) ? xhn[row] : 0.f;
  }
  // the row's score from its MFMA sum (FP8: times both scales; L2: less |x|^2 / 2)
  auto score = [&](float v, float qs, int f) -> float {
    const float s = PREC == 2 ? v * qs * xs[f] : v;
    if constexpr (L2) return s - hn[f];
    return s;
  };

  // The caller pads the query fragments (and qaux) to whole stages, so every stage load is
  // in bounds and every wave issues exactly LOADS loads per stage.
  const uint8_t* qsrc = static_cast<const uint8_t*>(qfrag);
  auto slot = [&](auto I) -> uint8_t* {
    constexpr int v = decltype(I)::value & 3;
    return v == 0 ? slot0 : v == 1 ? slot1 : v == 2 ? slot2 : slot3;
  };
  auto fetch = [&](uint8_t* dst, uint32_t st) {
#pragma unroll
    for (int r = 0; r < SH::PER_WAVE; ++r) {
      const int piece = wave + r * kWaves;
      load_lds16(qsrc + size_t(st) * SH::STAGE_Q + size_t(piece) * 1024 + lane * 16, dst + piece * 1024);
    }
#pragma unroll
    for (int r = 0; r < SH::AUX_LOADS; ++r)
      load_lds4(qaux + size_t(st) * G * kAux + r * 64 + lane, dst + SH::STAGE_Q + r * 256);
  };
  auto flush_min = [&](uint32_t st) {  // tile minima of stage st (pass 1), after a barrier
    if (threadIdx.x < G * 16) {
      const float* rp = red[st & 1];
      const uint32_t q = st * G * 16 + threadIdx.x;
      float m = rp[threadIdx.x];
#pragma unroll
      for (int w = 1; w < kWaves; ++w) m = fminf(m, rp[w * (G * 16) + threadIdx.x]);
      if (q < nq) tilemin[size_t(q) * n_space + j] = ord_bits(m);
    }
  };

  // Consume stage st from slot I; refill slot I + 3 (consumed at st - 1) with stage st + 3.
  // Every step issues the same loads (past the last stage it fetches the last stage again, into
  // a slot nobody reads), so the counted wait is the same everywhere and the compiler's own load
  // tracking stays exact.
  auto step = [&](auto I, uint32_t st) {
    wait_vm<(kAhead - 1) * LOADS>();  // the two later stages may stay in flight
    ring_barrier();  // every wave's part of stage st is in LDS; stage st - 1 is consumed
    if (PASS == 1 && st > st0) flush_min(st - 1);
    fetch(slot(Slot<decltype(I)::value + kAhead>()), min(st + kAhead, nstages - 1));
    const uint8_t* stage = slot(I);
    const FT* as = reinterpret_cast<const FT*>(stage);
    const float* aux = reinterpret_cast<const float*>(stage + SH::STAGE_Q);
    float* rd = red[st & 1];
    // The MFMA products of fragment g (query fragment st * G + g) against the wave's rows.
    auto products = [&](int g, f32x4m_t (&acc)[RF]) __attribute__((always_inline)) {
#pragma unroll
      for (int f = 0; f < RF; ++f) acc[f] = f32x4m_t{0.f, 0.f, 0.f, 0.f};
      if constexpr (fp8_x128(PREC, KS)) {
        // as below: operands read kAD instructions ahead (four 8-byte reads each, which the
        // compiler pairs into two ds_read2st64_b64)
        constexpr int T = KS / 4, kAD = T < 2 ? T : 2;
        i64x4m_t av[T];
        auto read_a = [&](int t) {
          const FT* ap = as + (g * KS + 4 * t) * 64 + lane;
          av[t] = i64x4m_t{ap[0], ap[64], ap[128], ap[192]};
        };
#pragma unroll
        for (int t = 0; t < kAD; ++t) read_a(t);
#pragma unroll
        for (int t = 0; t < T; ++t) {
          if (t + kAD < T) read_a(t + kAD);
#pragma unroll
          for (int f = 0; f < RF; ++f)
            acc[f] = mfma_fp8x128(av[t], i64x4m_t{b[f][4 * t], b[f][4 * t + 1], b[f][4 * t + 2], b[f][4 * t + 3]},
                                  acc[f]);
        }
        __builtin_amdgcn_sched_group_barrier(0x100, 2 * kAD, 0);
#pragma unroll
        for (int t = 0; t + kAD < T; ++t) {
          __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
          __builtin_amdgcn_sched_group_barrier(0x008, RF, 0);
        }
        __builtin_amdgcn_sched_group_barrier(0x008, RF * kAD, 0);
      } else {
        // A operands read kAD k-steps ahead of their MFMAs: the scheduler is held to that order
        // (kAD reads, then one read per k-step's RF MFMAs), so each read has kAD steps of MFMA
        // issue to land in instead of being waited for right away
        constexpr int kAD = KS < 3 ? KS : 3;
        FT av[KS];
#pragma unroll
        for (int s = 0; s < kAD; ++s) av[s] = as[(g * KS + s) * 64 + lane];
#pragma unroll
        for (int s = 0; s < KS; ++s) {
          if (s + kAD < KS) av[s + kAD] = as[(g * KS + s + kAD) * 64 + lane];
#pragma unroll
          for (int f = 0; f < RF; ++f) acc[f] = mfma<PREC>(av[s], b[f][s], acc[f]);
        }
        __builtin_amdgcn_sched_group_barrier(0x100, kAD, 0);
#pragma unroll
        for (int s = 0; s + kAD < KS; ++s) {
          __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
          __builtin_amdgcn_sched_group_barrier(0x008, RF, 0);
        }
        __builtin_amdgcn_sched_group_barrier(0x008, RF * kAD, 0);
      }
    };
    // FP8 pass 2 (registers to spare) issues fragment g + 1's MFMAs before fragment g's
    // epilogue, so the matrix pipe works through them while the wave runs the compares.
    constexpr bool kPipe = PASS == 2 && PREC == 2;
    f32x4m_t accs[2][RF];
    if constexpr (kPipe) products(0, accs[0]);
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const uint32_t qf = st * G + g;
      f32x4m_t(&acc)[RF] = accs[g & 1];
      if constexpr (kPipe) {
        if (g + 1 < G) products(g + 1, accs[(g + 1) & 1]);
      } else {
        products(g, acc);
      }
      // value i of acc[f] belongs to query qf*16 + 4*(lane/16) + i and row (rb0+f)*16 + lane%16
      const int q4 = 4 * (lane >> 4);
      if (PASS == 1) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float qsc = PREC == 2 ? aux[g * kAux + 16 + q4 + i] : 1.f;
          float m = __builtin_inff();
#pragma unroll
          for (int f = 0; f < RF; ++f) {
            const float sc = score(acc[f][i], qsc, f);
            m = fminf(m, ok[f] ? 1.0f - sc : __builtin_inff());
          }
          m = row16_min(m);
          if ((lane & 15) == 0) rd[wave * (G * 16) + g * 16 + q4 + i] = m;
        }
      } else {
        // aux holds score cuts: every row with 1 - score <= threshold has score >= cut (padding
        // queries carry +inf and never emit). Each value's test leaves a lane mask in scalar
        // registers; the emission path runs only when some mask is not empty, and only for the
        // values whose mask is not empty.
        // FP8 inner product tests fl(acc * xs) >= cq, the query's folded cut (one multiply
        // instead of two); the emission path applies the exact test to the values that pass.
        constexpr bool kFold = PREC == 2 && !L2;
        float cut[4], qsc[4], cq[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          cut[i] = aux[g * kAux + q4 + i];
          qsc[i] = PREC == 2 ? aux[g * kAux + 16 + q4 + i] : 1.f;
          cq[i] = kFold ? aux[g * kAux + 32 + q4 + i] : 0.f;
        }
        uint64_t hm[RF][4];
        uint64_t any = 0;
#pragma unroll
        for (int f = 0; f < RF; ++f)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const bool pass = kFold ? acc[f][i] * xs[f] >= cq[i] : score(acc[f][i], qsc[i], f) >= cut[i];
            hm[f][i] = ballot(pass) & okm[f];
            any |= hm[f][i];
          }
        if (any) {
#pragma unroll
          for (int f = 0; f < RF; ++f)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              uint64_t m = hm[f][i];
              if (m) {
                const float sc = score(acc[f][i], qsc[i], f);
                bool hit = (m >> lane) & 1;
                if constexpr (kFold) {
                  hit = hit && sc >= cut[i];
                  m = ballot(hit);
                }
                const uint32_t pre =
                    __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
                const uint32_t at = n_emit + pre;
                const uint32_t q = qf * 16 + q4 + i;
                if (hit && at < wave_cap)
                  wemit[at] = make_key(1.0f - sc, (q << 6) | uint32_t(f * 16 + (lane & 15)));
                n_emit += uint32_t(__builtin_popcountll(m));
              }
            }
        }
      }
    }
  };

  // nstages > st0 (checked above)
  fetch(slot0, st0);
  fetch(slot1, min(st0 + 1, nstages - 1));
  fetch(slot2, min(st0 + 2, nstages - 1));
  for (uint32_t st = st0;;) {  // leaves the loop right after the last stage
    step(Slot<0>(), st);
    if (++st == nstages) break;
    step(Slot<1>(), st);
    if (++st == nstages) break;
    step(Slot<2>(), st);
    if (++st == nstages) break;
    step(Slot<3>(), st);
    if (++st == nstages) break;
  }
  wait_vm<0>();  // the refetches past the last stage land before the block's LDS is released
  if (PASS == 1) {  // tile minima of the last stage
    __syncthreads();
    flush_min(nstages - 1);
  }
  if (PASS == 2 && lane == 0) emit_count[size_t(j) * kWaves + wave] = n_emit;
}

// ---- pass 2 for small batches (queries stationary, rows streaming) ------------------------
// When a batch's query fragments fit the registers of one block's four waves, the rows need
// not stay in registers: wave w keeps fragments w * QW .. + QW of its query group (A operands
// of every k-step, with their score cuts and scales), and the block streams its range of row
// fragments through an LDS ring (NBUF slots of SF fragments, filled by LDS-DMA NBUF - 1 stages
// ahead). Every wave runs its fragments against every row
// fragment, so HBM reads each row once per query group. The MFMA operands and their order are
// those of mfma_scan_kernel, so the estimates are the same bits. Emission: wave w of block b,
// query group y appends to its own region ((y * nbx + b) * 4 + w) entries
// (ord(dh) << 32) | (query << kQsShift) | (row - first row of the block).
constexpr int kQsWaves = 4;
constexpr uint32_t kQsShift = 20;          // rows of a block < 2^20
constexpr uint32_t kQsMaxRows = 1u << 16;  // rows per block (entries carry 2^kQsShift row offsets)
// Query groups per launch (each reads the rows once): up to where the queries-stationary pass
// beats the rows-stationary one, measured at 1M x 384 (pass 2 at 512 queries: BF16 0.354 ms
// against 0.436, FP8 0.304 against 0.383), given one wave of blocks (rows per block up to
// kQsMaxRows) and, for FP8, the row scales carried by the stages.
constexpr uint32_t kQsMaxGroups = 4;

// row fragments per stage: the largest of 8, 4, 2, 1 whose stage fits 12 KB and splits into
// whole 1 KB pieces, the same number for each wave; 1 if none fits (d = 768 in BF16: 24 KB).
// A block's four-slot ring is then at most 48 KB (the compiler tracks LDS-DMA targets for only
// a few distinct LDS objects, so the ring has four slots, not more), and up to 3 blocks share a
// CU (qs_bpc: as many as the LDS holds), each with up to 3 stages in flight.
__host__ __device__ constexpr bool qs_sf_ok(int prec, int ks, int sf) {
  return sf * ks * (prec == 1 ? 1024 : 512) <= (12 << 10) && (sf * ks * (prec == 1 ? 1024 : 512)) % (1024 * kQsWaves) == 0;
}
__host__ __device__ constexpr int qs_sf(int prec, int ks) {
  return qs_sf_ok(prec, ks, 8) ? 8 : qs_sf_ok(prec, ks, 4) ? 4 : qs_sf_ok(prec, ks, 2) ? 2 : 1;
}
// stages per row fragment: 2 where one fragment alone exceeds 12 KB (d = 768 in BF16: its k-steps
// in two halves, the accumulators carried from the first to the second)
__host__ __device__ constexpr int qs_hs(int prec, int ks) { return ks * (prec == 1 ? 1024 : 512) > (12 << 10) ? 2 : 1; }
__host__ __device__ constexpr int qs_stage(int prec, int ks) {
  return qs_sf(prec, ks) * ks * (prec == 1 ? 1024 : 512) / qs_hs(prec, ks);
}
__host__ __device__ constexpr int qs_lds(int prec, int ks) {  // the ring (FP8: with its scales)
  return 4 * (qs_stage(prec, ks) + (prec == 2 ? kQsWaves * 256 * ((qs_sf(prec, ks) * 16 + 63) / 64) : 0));
}
__host__ __device__ constexpr int qs_bpc(int prec, int ks) {  // blocks per CU
  // (FP8 at d = 768 keeps 2: its 172 registers a lane exceed the 168 of three blocks a CU)
  return prec == 2 && ks > 12 ? 2
         : (160 << 10) / qs_lds(prec, ks) > 3 ? 3
         : (160 << 10) / qs_lds(prec, ks) < 1 ? 1
                                               : (160 << 10) / qs_lds(prec, ks);
}
// query fragments per wave: A operands of all k-steps in at most 96 registers
__host__ __device__ constexpr int qs_qw(int prec, int ks) { return 2 * ks * (prec == 1 ? 4 : 2) <= 96 ? 2 : 1; }

template <int PREC, int KS>
struct QsShape {
  static constexpr int CB = Frag<PREC>::kChunkBytes;
  static constexpr int SF = qs_sf(PREC, KS);
  static constexpr int HS = qs_hs(PREC, KS);
  static constexpr int STAGE = qs_stage(PREC, KS);
  static constexpr int KH = KS / HS;                      // k-steps per stage
  static constexpr int NBUF = 4;
  static constexpr int LOADS = STAGE / 1024 / kQsWaves;  // DMA loads per wave per stage
  static constexpr int QW = qs_qw(PREC, KS);
  static_assert(STAGE % (1024 * kQsWaves) == 0 && (KS * CB) % 1024 == 0, "stage shape");
  static_assert(HS == 1 || (SF == 1 && KS % HS == 0), "half stages hold one fragment");
};

// NT: the rows stream through non-temporal loads (a single query group reads them once;
// several groups re-read them from L2)
template <int PREC, int KS, bool NT>
__global__ __launch_bounds__(64 * kQsWaves, qs_bpc(PREC, KS)) void mfma_qs_kernel(
    const void* __restrict__ xfrag, const float* __restrict__ xscale, uint32_t n,
    const void* __restrict__ qfrag, const float* __restrict__ qaux, uint32_t nq, uint32_t rows_per_block,
    uint32_t nbx, uint32_t ngroups, uint64_t* __restrict__ emit, uint32_t wave_cap,
    uint32_t* __restrict__ emit_count) {
  typedef typename Frag<PREC>::T FT;
  typedef QsShape<PREC, KS> SH;
  constexpr int SF = SH::SF, HS = SH::HS, KH = SH::KH, NBUF = SH::NBUF, QW = SH::QW, LOADS = SH::LOADS,
                ST = SH::STAGE;
  constexpr int FB = KS * SH::CB;  // bytes of one row fragment (all k-steps)
  // distinct LDS objects per slot: the compiler then waits only on the counted vmcnt
  static_assert(NBUF == 4, "four ring slots");
  // FP8: each slot also holds the row scales of its stage (XN 64-row wave loads), one copy per wave
  constexpr int XN = PREC == 2 ? (SF * 16 + 63) / 64 : 0, XB = kQsWaves * 256 * XN;
  static_assert(PREC != 2 || HS == 1, "FP8 stages hold whole fragments");
  __shared__ __attribute__((aligned(16))) uint8_t s0[ST + XB];
  __shared__ __attribute__((aligned(16))) uint8_t s1[ST + XB];
  __shared__ __attribute__((aligned(16))) uint8_t s2[ST + XB];
  __shared__ __attribute__((aligned(16))) uint8_t s3[ST + XB];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // Block (bx, qg) of nbx x ngroups: the query groups of one row range are 8 workgroups apart,
  // so they run on the same XCD (workgroups go round-robin over the 8 XCDs) at about the same
  // time and the later ones read the rows from that XCD's L2. nbx is a multiple of 8.
  const uint32_t bx = blockIdx.x / (8 * ngroups) * 8 + blockIdx.x % 8, qg = blockIdx.x / 8 % ngroups;
  const uint32_t row0 = bx * rows_per_block;
  const uint32_t region = (qg * nbx + bx) * kQsWaves + wave;
  uint64_t* wemit = emit + size_t(region) * wave_cap;
  if (row0 >= n) {  // the whole block: no rows
    if (lane == 0) emit_count[region] = 0;
    return;
  }
  const uint32_t nrf = (n + 15) >> 4;
  const uint32_t rf0 = row0 >> 4, rf1 = min(nrf, rf0 + rows_per_block / 16);
  const uint32_t nst = (rf1 - rf0 + SF - 1) / SF * HS;  // stage st: fragments (st / HS) * SF .., part st % HS

  // this wave's query fragments: A operands, cuts, scales (fragments past the batch never emit)
  const uint32_t nqf = (nq + 15) >> 4, q4 = 4 * (lane >> 4);
  const FT* qsrc = reinterpret_cast<const FT*>(qfrag);
  FT a[QW][KS];
  float cut[QW][4], qsc[QW][4], cq[QW][4];
#pragma unroll
  for (int j = 0; j < QW; ++j) {
    const uint32_t qf = (qg * kQsWaves + wave) * QW + j;
    const bool qok = qf < nqf;
#pragma unroll
    for (int s = 0; s < KS; ++s) a[j][s] = qok ? qsrc[(size_t(qf) * KS + s) * 64 + lane] : FT{};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      cut[j][i] = qok ? qaux[qf * kAux + q4 + i] : __builtin_inff();
      qsc[j][i] = (PREC == 2 && qok) ? qaux[qf * kAux + 16 + q4 + i] : 1.f;
      cq[j][i] = (PREC == 2 && qok) ? qaux[qf * kAux + 32 + q4 + i] : __builtin_inff();
    }
  }
  wait_vm<0>();
  __syncthreads();

  const uint8_t* xsrc = static_cast<const uint8_t*>(xfrag);
  // FP8: the row scales of a stage's SF fragments travel with the stage by LDS-DMA (lane l:
  // row 16 rf + l of the stage's first fragment rf), each wave into its own copy, so the
  // stage's counted wait covers them and no register load crosses the loop; VL = vector-memory
  // loads per thread per stage
  constexpr int VL = LOADS + XN;
  auto slot = [&](auto I) -> uint8_t* {
    constexpr int v = decltype(I)::value % NBUF;
    return v == 0 ? s0 : v == 1 ? s1 : v == 2 ? s2 : s3;
  };
  auto fetch = [&](auto I, uint32_t st) __attribute__((always_inline)) {
    uint8_t* dst = slot(I);
    const uint32_t fg = st / HS, part = st - fg * HS;
#pragma unroll
    for (int r = 0; r < LOADS; ++r) {
      const uint32_t piece = uint32_t(wave + r * kQsWaves);
      const uint32_t at = part * ST + piece * 1024, fi = at / FB;  // byte of the stage's fragments
      uint32_t rf = rf0 + fg * SF + fi;
      if (rf >= rf1) rf = rf0;  // past the block's rows: a fragment that is never read
      if constexpr (NT)
        load_lds16_nt(xsrc + size_t(rf) * FB + (at - fi * FB) + lane * 16, dst + piece * 1024);
      else
        load_lds16(xsrc + size_t(rf) * FB + (at - fi * FB) + lane * 16, dst + piece * 1024);
    }
#pragma unroll
    for (int r = 0; r < XN; ++r) {  // rows past the stage's SF * 16: loaded, never read
      const uint32_t row = (rf0 + fg * SF) * 16 + uint32_t(r * 64 + lane);
      if constexpr (NT)
        load_lds4_nt(xscale + (row < n ? row : n - 1), dst + ST + (wave * XN + r) * 256);
      else
        load_lds4(xscale + (row < n ? row : n - 1), dst + ST + (wave * XN + r) * 256);
    }
  };
  uint32_t n_emit = 0;  // wave-uniform
  f32x4m_t acc[QW];     // carried from part 0 to part 1 of a fragment when HS = 2
  // part: st % HS, known at compile time (the loop runs 8 steps a trip from a multiple of 8)
  auto compute = [&](auto I, uint32_t st, auto part_c) __attribute__((always_inline)) {
    constexpr int part = decltype(part_c)::value;
    const FT* bs = reinterpret_cast<const FT*>(slot(I));
#pragma unroll
    for (int fi = 0; fi < SF; ++fi) {
      const uint32_t rf = rf0 + st / HS * SF + fi;  // (past the block's rows: masked, not skipped)
      if (part == 0) {
#pragma unroll
        for (int j = 0; j < QW; ++j) acc[j] = f32x4m_t{0.f, 0.f, 0.f, 0.f};
      }
      if constexpr (fp8_x128(PREC, KS)) {  // FP8 stages hold whole fragments (KH = KS)
#pragma unroll
        for (int t = 0; t < KS / 4; ++t) {
          const FT* bp = bs + (fi * KS + 4 * t) * 64 + lane;
          const i64x4m_t b = {bp[0], bp[64], bp[128], bp[192]};
#pragma unroll
          for (int j = 0; j < QW; ++j)
            acc[j] = mfma_fp8x128(i64x4m_t{a[j][4 * t], a[j][4 * t + 1], a[j][4 * t + 2], a[j][4 * t + 3]}, b,
                                  acc[j]);
        }
      } else {
#pragma unroll
        for (int s = 0; s < KH; ++s) {
          const FT b = bs[(fi * KH + s) * 64 + lane];
#pragma unroll
          for (int j = 0; j < QW; ++j) acc[j] = mfma<PREC>(a[j][part * KH + s], b, acc[j]);
        }
      }
      if (part + 1 < HS) continue;  // the fragment's second half comes with the next stage
      // value i of acc[j] belongs to query (fragment of j) * 16 + q4 + i and row rf * 16 + lane % 16
      const uint32_t row = rf * 16 + (lane & 15);
      const bool ok = rf < rf1 && row < n;
      float xs = 1.f;
      if constexpr (PREC == 2)
        xs = ok ? reinterpret_cast<const float*>(slot(I) + ST + wave * XN * 256)[fi * 16 + (lane & 15)] : 1.f;
      // one lane mask per value (FP8: the folded test, exact test at emission; see
      // folded_cut), the emission path only for values whose mask is not empty
      const uint64_t okm = ballot(ok);
      uint64_t hm[QW][4];
      uint64_t any = 0;
#pragma unroll
      for (int j = 0; j < QW; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          hm[j][i] = ballot(PREC == 2 ? acc[j][i] * xs >= cq[j][i] : acc[j][i] >= cut[j][i]) & okm;
          any |= hm[j][i];
        }
      if (any) {
#pragma unroll
        for (int j = 0; j < QW; ++j)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            uint64_t m = hm[j][i];
            if (m) {
              const float sc = PREC == 2 ? acc[j][i] * qsc[j][i] * xs : acc[j][i];
              bool hit = (m >> lane) & 1;
              if constexpr (PREC == 2) {
                hit = hit && sc >= cut[j][i];
                m = ballot(hit);
              }
              const uint32_t pre =
                  __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
              const uint32_t at = n_emit + pre;
              const uint32_t q = ((qg * kQsWaves + wave) * QW + j) * 16 + q4 + i;
              if (hit && at < wave_cap) wemit[at] = make_key(1.0f - sc, (q << kQsShift) | (row - row0));
              n_emit += uint32_t(__builtin_popcountll(m));
            }
          }
      }
    }
  };
  // consume stage st from slot I; refill slot I + NBUF - 1 (consumed at st - 1) with stage
  // st + NBUF - 1 (past the last stage: the last stage again, into a slot nobody reads)
  auto step = [&](auto I, uint32_t st) __attribute__((always_inline)) {
    wait_vm<(NBUF - 2) * VL>();  // the later stages may stay in flight
    ring_barrier();
    fetch(Slot<decltype(I)::value + NBUF - 1>(), min(st + NBUF - 1, nst - 1));
    compute(I, st, Slot<decltype(I)::value % HS>());
  };
  fetch(Slot<0>(), 0);
  fetch(Slot<1>(), min(1u, nst - 1));
  fetch(Slot<2>(), min(2u, nst - 1));
  // The loop runs eight steps a trip: at the loop head the compiler waits for every load in
  // flight (it does not carry its per-slot tracking of the LDS-DMA loads around the loop).
  for (uint32_t st = 0;;) {  // leaves the loop right after the last stage
    step(Slot<0>(), st);
    if (++st == nst) break;
    step(Slot<1>(), st);
    if (++st == nst) break;
    step(Slot<2>(), st);
    if (++st == nst) break;
    step(Slot<3>(), st);
    if (++st == nst) break;
    step(Slot<4>(), st);
    if (++st == nst) break;
    step(Slot<5>(), st);
    if (++st == nst) break;
    step(Slot<6>(), st);
    if (++st == nst) break;
    step(Slot<7>(), st);
    if (++st == nst) break;
  }
  wait_vm<0>();  // the refetches land before the block's LDS is released
  if (lane == 0) emit_count[region] = n_emit;
}

// ---- pass 2 for small batches, BF16: queries stationary, k-split, rows through registers ----
// The LDS-DMA ring of mfma_qs_kernel holds at most three 12 KB stages in flight per block and
// three blocks per CU, and streams at about 0.76 of HBM even without its products. Here a block
// of 4 waves serves one query group (8 fragments, 128 queries) over a row range with no staging:
// wave w holds the A operands of all 8 fragments for k-steps [w KQ, (w + 1) KQ) (KQ = KS / 4) and
// loads its k-steps of every row fragment straight into registers (one coalesced 1 KB chunk per
// k-step), kQkDepth fragments ahead, so the block reads each row fragment once and two blocks a
// CU keep 2 x 4 x kQkDepth x KQ KB in flight. A wave's MFMAs give partial sums over its k-steps
// for the group's 128 queries against the fragment's 16 rows; the partials go to LDS (two
// buffers), and after one barrier wave w adds the four waves' partials of fragments 2w and
// 2w + 1 (in wave order) and runs mfma_qs_kernel's emission for them: the same fragments, the
// same regions and the same entries as that kernel. The products are summed in another order
// than in the other passes, so an estimate may differ from theirs in its last bits; each is
// within the bound E (any fp32 order of the d products' sum is within 2 gamma_d |q^| |x^|).
// fragments in flight a wave: pass 2 at 1M x 384, batch 8 / 128: depth 3 0.118 / 0.130 ms,
// 4 0.118 / 0.127, 5 0.119 / 0.128, 6 0.121 / 0.130, 8 0.164 / 0.180 (registers spill)
constexpr int kQkDepth = 4;
// The queries-stationary passes stream their rows with non-temporal loads when one query group
// reads them once (at 1M x 384, batch 128: BF16 0.151 -> 0.127 ms, FP8 0.086 -> 0.071 ms); with
// two to four groups the later groups re-read the rows from L2 and plain loads are as fast or
// faster (BF16 at 512 queries: 0.332 against 0.360 ms).
constexpr uint32_t kQsNtGroups = 1;
__host__ __device__ constexpr bool qk_supported(int prec, int ks) {
  return prec == 1 && (ks == 4 || ks == 8 || ks == 12);
}

template <int KS>
__global__ __launch_bounds__(64 * kQsWaves, 2) void mfma_qk_kernel(
    const void* __restrict__ xfrag, uint32_t n, const void* __restrict__ qfrag, const float* __restrict__ qaux,
    uint32_t nq, uint32_t rows_per_block, uint32_t nbx, uint32_t ngroups, uint64_t* __restrict__ emit,
    uint32_t wave_cap, uint32_t* __restrict__ emit_count) {
  static_assert(KS % kQsWaves == 0, "k-steps split over the waves");
  constexpr int KQ = KS / kQsWaves;  // k-steps per wave
  constexpr int QF = kQsWaves * 2;   // query fragments per group, two emitted by each wave
  constexpr int D = kQkDepth;
  __shared__ __attribute__((aligned(16))) f32x4m_t part[2][kQsWaves][QF][64];  // 64 KB
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // block (bx, qg) as in mfma_qs_kernel: the query groups of one row range on one XCD
  const uint32_t bx = blockIdx.x / (8 * ngroups) * 8 + blockIdx.x % 8, qg = blockIdx.x / 8 % ngroups;
  const uint32_t row0 = bx * rows_per_block;
  const uint32_t region = (qg * nbx + bx) * kQsWaves + wave;
  uint64_t* wemit = emit + size_t(region) * wave_cap;
  if (row0 >= n) {  // the whole block: no rows
    if (lane == 0) emit_count[region] = 0;
    return;
  }
  const uint32_t nrf = (n + 15) >> 4;
  const uint32_t rf0 = row0 >> 4, rf1 = min(nrf, rf0 + rows_per_block / 16);
  const uint32_t nf = rf1 - rf0;  // row fragments of the block (>= 1)

  // A operands: the group's fragments, this wave's k-steps (fragments past the batch: zero)
  const uint32_t nqf = (nq + 15) >> 4, q4 = 4 * (lane >> 4);
  const bf16x8_t* qsrc = reinterpret_cast<const bf16x8_t*>(qfrag);
  bf16x8_t a[QF][KQ];
#pragma unroll
  for (int j = 0; j < QF; ++j) {
    const uint32_t qf = qg * QF + j;
#pragma unroll
    for (int s = 0; s < KQ; ++s)
      a[j][s] = qf < nqf ? qsrc[(size_t(qf) * KS + wave * KQ + s) * 64 + lane] : bf16x8_t{};
  }
  // score cuts of the two fragments this wave emits (padding queries: +inf, never emit)
  float cut[2][4];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const uint32_t qf = qg * QF + 2 * wave + j;
#pragma unroll
    for (int i = 0; i < 4; ++i) cut[j][i] = qf < nqf ? qaux[qf * kAux + q4 + i] : __builtin_inff();
  }

  // this wave's k-steps of row fragment f (past the block's last fragment: the last one again,
  // loaded and never used)
  const bf16x8_t* xsrc = reinterpret_cast<const bf16x8_t*>(xfrag);
  bf16x8_t b[D][KQ];
  auto load = [&](int slot, uint32_t f) __attribute__((always_inline)) {
    const uint32_t rf = rf0 + (f < nf ? f : nf - 1);
#pragma unroll
    for (int s = 0; s < KQ; ++s) b[slot][s] = __builtin_nontemporal_load(&xsrc[(size_t(rf) * KS + wave * KQ + s) * 64 + lane]);
  };
#pragma unroll
  for (int u = 0; u < D; ++u) load(u, uint32_t(u));

  uint32_t n_emit = 0;  // wave-uniform
  for (uint32_t f0 = 0; f0 < nf; f0 += D) {
#pragma unroll
    for (int u = 0; u < D; ++u) {
      const uint32_t f = f0 + uint32_t(u);
      if (f >= nf) break;  // block-uniform
      f32x4m_t acc[QF];
#pragma unroll
      for (int j = 0; j < QF; ++j) acc[j] = f32x4m_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < KQ; ++s)
#pragma unroll
        for (int j = 0; j < QF; ++j) acc[j] = mfma<1>(a[j][s], b[u][s], acc[j]);
      load(u, f + D);  // the slot's next fragment
      f32x4m_t* pw = &part[f & 1][wave][0][0];
#pragma unroll
      for (int j = 0; j < QF; ++j) pw[j * 64 + lane] = acc[j];
      // every wave's partials of fragment f are in LDS; the buffer written next (f + 1) was
      // last read for fragment f - 1, which every wave finished before reaching this barrier
      __syncthreads();
      // value i of t[j] belongs to query (qg * QF + 2 wave + j) * 16 + q4 + i and row rf * 16 + lane % 16
      const uint32_t rf = rf0 + f;
      const uint32_t row = rf * 16 + (lane & 15);
      const uint64_t okm = ballot(row < n);
      f32x4m_t t[2];
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        t[j] = part[f & 1][0][2 * wave + j][lane];
#pragma unroll
        for (int v = 1; v < kQsWaves; ++v) t[j] += part[f & 1][v][2 * wave + j][lane];
      }
      uint64_t hm[2][4];
      uint64_t any = 0;
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          hm[j][i] = ballot(t[j][i] >= cut[j][i]) & okm;
          any |= hm[j][i];
        }
      if (any) {
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const uint64_t m = hm[j][i];
            if (m) {
              const bool hit = (m >> lane) & 1;
              const uint32_t pre =
                  __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
              const uint32_t at = n_emit + pre;
              const uint32_t q = (qg * QF + 2 * wave + j) * 16 + q4 + i;
              if (hit && at < wave_cap) wemit[at] = make_key(1.0f - t[j][i], (q << kQsShift) | (row - row0));
              n_emit += uint32_t(__builtin_popcountll(m));
            }
          }
      }
    }
  }
  wait_vm<0>();  // the loads past the last fragment land before the wave ends
  if (lane == 0) emit_count[region] = n_emit;
}

// Where the rows of an emission region start, and how an entry's low word splits into query
// and row within the region: lo = (query << qshift) | row offset. Region r starts at row
// ((r / waves) % blocks) * rows (rows-stationary pass 2: waves 1, one region per wave of
// rows_frags * 16 rows, qshift 6; queries-stationary: 4 regions per block, one per wave, each
// over the block's row range, qshift 20).
struct RegionGeometry {
  uint32_t waves, blocks, rows, qshift;
  __device__ __forceinline__ uint32_t row(uint32_t r, uint32_t lo) const {
    return ((r / waves) % blocks) * rows + (lo & ((1u << qshift) - 1u));
  }
};

// Files the pass-2 entries of every wave region into per-query candidate lists (count[q] keys
// (ord(dh) << 32) | row at cand[q * cap ..]) without global atomics on the per-query counters
// (device-scope atomics on a few hundred hot addresses serialise). Three launches over S
// slices of `per` consecutive regions each:
//   count   per slice, entries per query in LDS -> slice_cnt[q * kScatterSlices + s]; a region that
//           overflowed sets *overflow (the round is then answered by the exact scan);
//   offsets per query (one wave), exclusive scan over the slices in place, total -> count[q];
//   place   per slice, each entry at cand[q * cap + offset + its LDS rank] if below cap.
// A batch of at most 128 queries skips the offsets launch: each place block sums the counts of
// the slices before its own.
// The order of a query's candidates is not fixed; the rescore's result does not depend on it.
// A slice's block has 16 waves, each walking its own regions, so that a CU keeps 16 regions'
// loads in flight: each wave's loop is a chain of a load and an LDS atomic.
constexpr uint32_t kScatterSlices = 1024;
constexpr uint32_t kScatterThreads = 1024;
constexpr uint32_t kScatterWaves = kScatterThreads / 64;

__global__ __launch_bounds__(kScatterThreads) void scatter_count_kernel(
    const uint64_t* __restrict__ emit, const uint32_t* __restrict__ emit_count, uint32_t wave_cap,
    RegionGeometry geo, uint32_t n_regions, uint32_t per, uint32_t nq, uint32_t S,
    uint32_t* __restrict__ slice_cnt, uint32_t* __restrict__ overflow) {
  extern __shared__ uint32_t hist[];  // nq
  __shared__ uint32_t rc[64];
  const uint32_t s = blockIdx.x, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const uint32_t r0 = s * per, r1 = min(n_regions, r0 + per);
  for (uint32_t q = tid; q < nq; q += kScatterThreads) hist[q] = 0;
  for (uint32_t rb = r0; rb < r1; rb += 64) {  // region counts staged 64 at a time
    __syncthreads();
    if (tid < min(64u, r1 - rb)) rc[tid] = emit_count[rb + tid];
    __syncthreads();
    for (uint32_t r = rb + wave; r < min(r1, rb + 64); r += kScatterWaves) {
      const uint32_t c = rc[r - rb];
      if (c > wave_cap) {
        if (lane == 0) atomicAdd(overflow, 1u);
        continue;
      }
      for (uint32_t i = lane; i < c; i += 64) atomicAdd(&hist[uint32_t(emit[size_t(r) * wave_cap + i]) >> geo.qshift], 1u);
    }
  }
  __syncthreads();
  for (uint32_t q = tid; q < nq; q += kScatterThreads) slice_cnt[size_t(q) * kScatterSlices + s] = hist[q];
}

// slice_cnt rows have kScatterSlices entries (16-byte aligned); each lane scans 16 of them
__global__ __launch_bounds__(256) void scatter_offsets_kernel(uint32_t* __restrict__ slice_cnt, uint32_t S,
                                                               uint32_t nq, uint32_t* __restrict__ count) {
  static_assert(kScatterSlices == 64 * 16, "one wave of 16 slices a lane");
  const uint32_t q = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (q >= nq) return;
  uint4* v = reinterpret_cast<uint4*>(slice_cnt + size_t(q) * kScatterSlices + lane * 16);
  uint32_t x[16];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint4 t = v[j];
    x[4 * j] = t.x, x[4 * j + 1] = t.y, x[4 * j + 2] = t.z, x[4 * j + 3] = t.w;
  }
  uint32_t sum = 0;
#pragma unroll
  for (int j = 0; j < 16; ++j) {  // exclusive offsets within the lane; slices >= S hold nothing
    const uint32_t c = lane * 16 + j < S ? x[j] : 0u;
    x[j] = sum;
    sum += c;
  }
  uint32_t incl = sum;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(incl, off, 64);
    if (lane >= uint32_t(off)) incl += y;
  }
  const uint32_t base = incl - sum;
#pragma unroll
  for (int j = 0; j < 4; ++j)
    v[j] = make_uint4(base + x[4 * j], base + x[4 * j + 1], base + x[4 * j + 2], base + x[4 * j + 3]);
  if (lane == 63) count[q] = incl;
}

// With own_offsets set (a batch of at most kPlaceOwnQ queries), slice_off holds the raw counts
// of scatter_count_kernel and each slice's block sums the counts of the slices before its own
// (8 lanes a query), and the last slice's block writes the totals to count: no offsets launch.
constexpr uint32_t kPlaceOwnQ = kScatterThreads / 8;
__global__ __launch_bounds__(kScatterThreads) void scatter_place_kernel(
    const uint64_t* __restrict__ emit, const uint32_t* __restrict__ emit_count, uint32_t wave_cap,
    RegionGeometry geo, uint32_t n_regions, uint32_t per, uint32_t nq, uint32_t S,
    const uint32_t* __restrict__ slice_off, uint64_t* __restrict__ cand, uint32_t cap, int own_offsets,
    uint32_t* __restrict__ count) {
  extern __shared__ uint32_t base[];  // nq
  __shared__ uint32_t rc[64];
  const uint32_t s = blockIdx.x, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const uint32_t r0 = s * per, r1 = min(n_regions, r0 + per);
  if (own_offsets) {
    const uint32_t q = tid >> 3, part = tid & 7;  // nq <= kPlaceOwnQ: every query has its 8 lanes
    uint32_t sum = 0;
    if (q < nq) {
      const uint32_t* row = slice_off + size_t(q) * kScatterSlices;
      for (uint32_t j0 = 0; j0 < s; j0 += 128) {  // 16 loads in flight a lane
        uint32_t v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          const uint32_t j = j0 + part + 8 * u;
          v[u] = j < s ? row[j] : 0u;
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) sum += v[u];
      }
    }
    sum += __shfl_xor(sum, 1, 64);
    sum += __shfl_xor(sum, 2, 64);
    sum += __shfl_xor(sum, 4, 64);
    if (q < nq && part == 0) {
      base[q] = sum;
      if (s == S - 1) count[q] = sum + slice_off[size_t(q) * kScatterSlices + s];
    }
  } else {
    for (uint32_t q = tid; q < nq; q += kScatterThreads) base[q] = slice_off[size_t(q) * kScatterSlices + s];
  }
  for (uint32_t rb = r0; rb < r1; rb += 64) {
    __syncthreads();
    if (tid < min(64u, r1 - rb)) rc[tid] = emit_count[rb + tid];
    __syncthreads();
    for (uint32_t r = rb + wave; r < min(r1, rb + 64); r += kScatterWaves) {
      const uint32_t c = rc[r - rb];
      if (c > wave_cap) continue;
      for (uint32_t i = lane; i < c; i += 64) {
        const uint64_t e = emit[size_t(r) * wave_cap + i];
        const uint32_t lo = uint32_t(e), q = lo >> geo.qshift;
        const uint32_t slot = atomicAdd(&base[q], 1u);
        if (slot < cap) cand[size_t(q) * cap + slot] = (e & 0xFFFFFFFF00000000ull) | geo.row(r, lo);
      }
    }
  }
}

// Smallest float above or equal to the real number s (s finite or infinite, not NaN).
__device__ __forceinline__ float round_up_f(double s) {
  float r = float(s);
  if (double(r) < s) r = __uint_as_float(r >= 0.f ? __float_as_uint(r) + 1u : __float_as_uint(r) - 1u);
  return r;
}
__device__ __forceinline__ float round_down_f(double s) { return -round_up_f(-s); }

// Score cut of a distance threshold T: every float score s with fl(1 - s) <= T has s >= cut.
// fl(1 - s) = (1 - s)(1 + d), |d| <= 2^-24, so fl(1 - s) <= T gives 1 - s <= T + |T| 2^-23
// and s >= 1 - T - |T| 2^-23; the cut takes twice that margin and rounds down. T = -inf (no
// candidates) gives +inf; T = +inf or NaN gives -inf (every row is a candidate).
__device__ __forceinline__ float score_cut(float T) {
  if (T == -__builtin_inff()) return __builtin_inff();
  if (!(T < __builtin_inff())) return -__builtin_inff();
  const double t = T;
  return round_down_f(1.0 - t - fabs(t) * 0x1p-22 - 0x1p-60);
}

// Folded FP8 cut: a value passes fl(acc * xs) >= folded_cut(cut, qsc) whenever its score
// fl(fl(acc * qsc) * xs) >= cut (qsc, xs > 0). The three roundings between the two tests
// (two in the score, one in acc * xs) and the one of t = fl(cut / qsc) move the compared values
// by at most 4u of their size, so lowering t by 2^-18 of its size (64u) keeps every passing
// value (the sign of a product survives its rounding). Infinite cuts stay infinite (+inf:
// nothing passes; -inf: everything does).
__device__ __forceinline__ float folded_cut(float cut, float qsc) {
  const float t = cut / qsc;
  return __builtin_isinf(t) || t != t ? t : t - fabsf(t) * 0x1p-18f;
}

// Pass-2 aux block of the query stages: for query q = 16 f + i, aux[f * kAux + i] = its score
// cut (score_cut of its threshold; +inf for padding queries, which then never emit),
// aux[f * kAux + 16 + i] = its scale (1 for padding queries) and aux[f * kAux + 32 + i] = the
// folded cut of the two.
__device__ __forceinline__ void set_pass2_aux(float* aux, uint32_t q, float cut, float sc) {
  aux[(q >> 4) * kAux + (q & 15)] = cut;
  aux[(q >> 4) * kAux + 16 + (q & 15)] = sc;
  aux[(q >> 4) * kAux + 32 + (q & 15)] = folded_cut(cut, sc);
}

// ---- threshold select --------------------------------------------------------------------
// Per query (one 256-thread block): U = the kq-th smallest tile minimum (radix select on the
// order-preserving bits), threshold = U + 2E. With fewer tiles than kq the threshold is +inf.
// The same selection by one wave per query when the tile count is at most 64 * NJ (NJ <= 16):
// each lane holds NJ minima and the kq-th smallest is found bit by bit with ballot counts.
template <int NJ>
__global__ __launch_bounds__(256) void select_threshold_wave_kernel(
    const uint32_t* __restrict__ tilemin, uint32_t n_tiles, uint32_t nq, uint32_t nq_pad,
    const uint32_t* __restrict__ ks, uint32_t k, const float* __restrict__ E, const float* __restrict__ qscale,
    float* __restrict__ thr, float* __restrict__ aux) {
  const uint32_t q = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (q >= nq_pad) return;
  if (q >= nq) {
    if (lane == 0) set_pass2_aux(aux, q, __builtin_inff(), 1.f);
    return;
  }
  const uint32_t kq = ks ? ks[q] : k;
  if (kq == 0 || kq > n_tiles) {
    if (lane == 0) {
      thr[q] = __builtin_inff();
      set_pass2_aux(aux, q, score_cut(__builtin_inff()), qscale[q]);
    }
    return;
  }
  const uint32_t* v = tilemin + size_t(q) * n_tiles;
  uint32_t key[NJ];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const uint32_t i = lane + 64 * j;
    key[j] = i < n_tiles ? v[i] : 0xFFFFFFFFu;
  }
  uint32_t prefix = 0, rank = kq;  // find the rank-th smallest (1-based) among the keys
  for (int b = 31; b >= 0; --b) {
    const uint32_t hi_mask = b == 31 ? 0u : (0xFFFFFFFFu << (b + 1));
    uint32_t cnt = 0;  // keys matching the prefix above bit b with bit b clear
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const uint32_t i = lane + 64 * j;
      const bool m = i < n_tiles && (key[j] & hi_mask) == (prefix & hi_mask) && !((key[j] >> b) & 1u);
      cnt += uint32_t(__builtin_popcountll(ballot(m)));
    }
    if (cnt < rank) {
      rank -= cnt;
      prefix |= 1u << b;
    }
  }
  if (lane == 0) {
    const float U = unord_bits(prefix);
    const float t = round_up_f(double(U) + 2.0 * double(E[q]) * (1.0 + 1e-6) + 0x1p-20);
    thr[q] = t;
    set_pass2_aux(aux, q, score_cut(t), qscale[q]);
  }
}

__global__ __launch_bounds__(256) void select_threshold_kernel(
    const uint32_t* __restrict__ tilemin, uint32_t n_tiles, uint32_t nq, const uint32_t* __restrict__ ks,
    uint32_t k, const float* __restrict__ E, const float* __restrict__ qscale, float* __restrict__ thr,
    float* __restrict__ aux) {
  __shared__ uint32_t hist[256];
  __shared__ uint32_t sel[2];
  const uint32_t q = blockIdx.x, tid = threadIdx.x;
  if (q >= nq) {
    if (tid == 0) set_pass2_aux(aux, q, __builtin_inff(), 1.f);
    return;
  }
  const uint32_t kq = ks ? ks[q] : k;
  const uint32_t* v = tilemin + size_t(q) * n_tiles;
  if (kq == 0 || kq > n_tiles) {
    if (tid == 0) {
      thr[q] = __builtin_inff();
      set_pass2_aux(aux, q, score_cut(__builtin_inff()), qscale[q]);
    }
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
    // U + 2E (with a relative cushion on E and an absolute one of 2^-20), rounded up
    const float t = round_up_f(double(U) + 2.0 * double(E[q]) * (1.0 + 1e-6) + 0x1p-20);
    thr[q] = t;
    set_pass2_aux(aux, q, score_cut(t), qscale[q]);
  }
}

// ---- rescore -----------------------------------------------------------------------------
// Up to kCandCap candidates per query; a query with more is flagged and answered by the exact
// scan. Candidates, then the exact keys, live in one dynamic LDS buffer of CAP keys followed by
// the query (dp floats). Two launches: CAP = kCandSmall for every query with at most that many
// candidates (small LDS, high occupancy), then CAP = kCandCap for the rest; a batch of at most
// 256 queries (one 16-wave block each, occupancy is not the limit) takes one CAP = kCandCap launch.
constexpr int kCandCap = 8192;
constexpr int kCandSmall = 2048;
constexpr int kTopCap = 256;  // keys ranked after the rescore's select

// Ascending bitonic sort of buf[0, n2) (n2 a power of two) by an NT-thread block.
template <int NT>
__device__ void bitonic_sort_lds(uint64_t* buf, uint32_t n2, uint32_t tid) {
  for (uint32_t size = 2; size <= n2; size <<= 1)
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      for (uint32_t i = tid; i < n2 / 2; i += NT) {
        const uint32_t lo = 2 * i - (i & (stride - 1)), hi = lo + stride;
        const bool up = (lo & size) == 0;
        const uint64_t a = buf[lo], b = buf[hi];
        if ((a > b) == up) {
          buf[lo] = b;
          buf[hi] = a;
        }
      }
      __syncthreads();
    }
}

// Writes the first kq of the r unique keys in buf[0, r) (r <= CAP) in ascending order through
// put(rank, key): by counting for up to 256 keys; else through a radix select of the kq-th
// smallest distance bits T, which leaves the keys with distance bits <= T (every key of rank < kq,
// all others being larger) to be ranked among themselves; a bitonic sort only where ties put more
// than kTopCap keys at or below T. Every thread of the block calls it.
template <int NT, class Put>
__device__ __forceinline__ void rank_first(uint64_t* buf, uint32_t r, uint32_t kq, uint32_t* hist, uint32_t* sel, uint64_t* top,
                           uint32_t* ntop, uint32_t tid, Put put) {
  bool sorted = false;
  if (r > 256 && kq > 0) {
    const uint32_t th = block_select_hi32<NT>(buf, r, min(kq, r), hist, sel, tid);
    if (tid == 0) *ntop = 0;
    __syncthreads();
    for (uint32_t i = tid; i < r; i += NT) {
      const uint64_t v = buf[i];
      if (uint32_t(v >> 32) <= th) {
        const uint32_t p = atomicAdd(ntop, 1u);
        if (p < uint32_t(kTopCap)) top[p] = v;
      }
    }
    __syncthreads();
    const uint32_t nt = *ntop;
    if (nt <= uint32_t(kTopCap)) {
      for (uint32_t i = tid; i < nt; i += NT) {
        const uint64_t v = top[i];
        uint32_t rank = 0;
        for (uint32_t j = 0; j < nt; ++j) rank += top[j] < v;
        if (rank < kq) put(rank, v);
      }
      sorted = true;
    }
  }
  if (!sorted && r <= uint32_t(NT)) {
    for (uint32_t i = tid; i < r; i += NT) {
      const uint64_t v = buf[i];
      uint32_t rank = 0;
      for (uint32_t j = 0; j < r; ++j) rank += buf[j] < v;
      if (rank < kq) put(rank, v);
    }
  } else if (!sorted) {
    uint32_t m2 = 2 * NT;  // <= CAP: more than NT keys came from at most CAP
    while (m2 < r) m2 <<= 1;
    for (uint32_t i = r + tid; i < m2; i += NT) buf[i] = kPadKey;
    __syncthreads();
    bitonic_sort_lds<NT>(buf, m2, tid);
    for (uint32_t i = tid; i < kq && i < r; i += NT) put(i, buf[i]);
  }
}

// Handles the queries with cmin <= count <= CAP; with final_pass set, flags those above CAP.
// NT threads per block: 1024 where the batch has few queries (one block each), so that a
// query's candidate rows are read by 16 waves at once instead of 4.
template <int D16, bool L2, int CAP, int NT>
__global__ __launch_bounds__(NT) void rescore_kernel(
    const float* __restrict__ base, uint32_t dp, uint32_t d, const float* __restrict__ queries,
    const uint32_t* __restrict__ ks, uint32_t k, const float* __restrict__ E,
    const uint64_t* __restrict__ cand, const uint32_t* __restrict__ count,
    uint32_t* __restrict__ out_ids, float* __restrict__ out_dists, uint32_t ldo,
    uint32_t* __restrict__ overflow_flags, uint32_t* __restrict__ n_overflow,
    uint32_t* __restrict__ rescored, uint32_t cmin, int final_pass, const uint32_t* __restrict__ qmap) {
  extern __shared__ uint64_t dyn[];
  uint64_t* buf = dyn;                                // CAP keys
  float* q_lds = reinterpret_cast<float*>(dyn + CAP);  // dp floats
  __shared__ uint32_t nsel, ntop;
  __shared__ uint32_t hist[256], sel[2];
  __shared__ uint64_t top[kTopCap];
  // block b answers query qmap[b] (rescore_order_kernel; kNoneId: none), else query b
  const uint32_t q = qmap ? qmap[blockIdx.x] : blockIdx.x;
  if (q == kNoneId) return;
  const uint32_t tid = threadIdx.x, lane = tid & 63, t = lane & 3, g = lane >> 2;
  const uint32_t kq = min(ks ? ks[q] : k, ldo);
  uint32_t* oi = out_ids + size_t(q) * ldo;
  float* od = out_dists + size_t(q) * ldo;
  const uint32_t c = count[q];
  if (c < cmin) return;  // done by the previous launch
  if (c > uint32_t(CAP)) {
    if (final_pass && tid == 0) {
      overflow_flags[q] = 1;
      rescored[q] = 0;
      atomicAdd(n_overflow, 1u);
    }
    return;
  }
  if (tid == 0) overflow_flags[q] = 0;
  for (uint32_t i = kq + tid; i < ldo; i += NT) {
    oi[i] = kNoneId;
    od[i] = __builtin_inff();
  }
  for (uint32_t i = tid; i < dp; i += NT) q_lds[i] = i < d ? queries[size_t(q) * d + i] : 0.f;
  // Dk = the kk-th smallest estimate; keep dh <= Dk + 2E (compacted, in no particular order)
  const uint64_t* cq = cand + size_t(q) * kCandCap;
  for (uint32_t i = tid; i < c; i += NT) buf[i] = cq[i];
  if (tid == 0) nsel = 0;
  __syncthreads();
  const uint32_t kk = min(kq, c);
  if (kk > 0) {
    const float Dk = unord_bits(block_select_hi32<NT>(buf, c, kk, hist, sel, tid));
    const float lim = round_up_f(double(Dk) + 2.0 * double(E[q]) * (1.0 + 1e-6) + 0x1p-20);
    __syncthreads();  // every read of buf by the select is done
    for (uint32_t i = tid; i < c; i += NT) {
      const uint64_t v = cq[i];
      if (key_dist(v) <= lim) buf[atomicAdd(&nsel, 1u)] = v;
    }
  }
  __syncthreads();
  const uint32_t r = nsel;
  if (tid == 0) rescored[q] = r;
  // exact distances, 16 rows per wave instruction group, NT / 64 waves; a quad reads its
  // candidate's row id and later overwrites that slot with the exact key
  for (uint32_t base_i = (tid >> 6) * 16; base_i < r; base_i += 16 * (NT / 64)) {
    const uint32_t i = base_i + g;
    bool ok[1] = {i < r};
    const uint32_t row = key_id(buf[ok[0] ? i : 0]);
    const float* xr[1] = {base + size_t(row) * dp};
    float dist[1][1];
    quad_dists<1, 1, D16, L2, 16, kLoadPlain>(xr, ok, q_lds, dp, d, int(t), dist);
    if (ok[0] && t == 0) buf[i] = make_key(dist[0][0], row);
  }
  __syncthreads();
  // rank the exact keys (unique) and write the first kq: by counting for up to 256 keys; else
  // through a radix select of the kq-th smallest distance bits T, which leaves the keys with
  // distance bits <= T (every key of rank < kq, all others being larger) to be ranked among
  // themselves; a bitonic sort only where ties put more than kTopCap keys at or below T
  bool sorted = false;
  if (r > 256 && kq > 0) {
    const uint32_t th = block_select_hi32<NT>(buf, r, min(kq, r), hist, sel, tid);
    if (tid == 0) ntop = 0;
    __syncthreads();
    for (uint32_t i = tid; i < r; i += NT) {
      const uint64_t v = buf[i];
      if (uint32_t(v >> 32) <= th) {
        const uint32_t p = atomicAdd(&ntop, 1u);
        if (p < uint32_t(kTopCap)) top[p] = v;
      }
    }
    __syncthreads();
    const uint32_t nt = ntop;
    if (nt <= uint32_t(kTopCap)) {
      for (uint32_t i = tid; i < nt; i += NT) {
        const uint64_t v = top[i];
        uint32_t rank = 0;
        for (uint32_t j = 0; j < nt; ++j) rank += top[j] < v;
        if (rank < kq) {
          oi[rank] = key_id(v);
          od[rank] = key_dist(v);
        }
      }
      sorted = true;
    }
  }
  if (!sorted && r <= uint32_t(NT)) {
    for (uint32_t i = tid; i < r; i += NT) {
      const uint64_t v = buf[i];
      uint32_t rank = 0;
      for (uint32_t j = 0; j < r; ++j) rank += buf[j] < v;
      if (rank < kq) {
        oi[rank] = key_id(v);
        od[rank] = key_dist(v);
      }
    }
  } else if (!sorted) {
    uint32_t m2 = 2 * NT;  // <= CAP: r > NT candidates came from at most CAP
    while (m2 < r) m2 <<= 1;
    for (uint32_t i = r + tid; i < m2; i += NT) buf[i] = kPadKey;
    __syncthreads();
    bitonic_sort_lds<NT>(buf, m2, tid);
    for (uint32_t i = tid; i < kq && i < r; i += NT) {
      oi[i] = key_id(buf[i]);
      od[i] = key_dist(buf[i]);
    }
  }
  for (uint32_t i = r + tid; i < kq; i += NT) {
    oi[i] = kNoneId;
    od[i] = __builtin_inff();
  }
}

// The rescore of a batch of at most half as many queries as CUs, with `split` blocks a
// query: block s rescores the candidates i with i % split == s (after the same select and cut as
// rescore_kernel, over all of the query's candidates) and writes its first kq exact keys to
// rpart (kPadKey past its own); rescore_merge_kernel then ranks the split x kq keys into the
// output. The rows of one query are then read by split CUs instead of one (at 1M x 384, batch 8:
// about 1,000 FP8 candidates a query, 1.5 MB of rows, on one CU each; 32 shares: FP8 50.7k ->
// 59.9k QPS). The merge is a launch of its own: a last-block hand-off inside this kernel needs
// device-scope fences, which write back and invalidate the L2 caches of the XCDs (the kernel
// took 99 instead of 21 us at batch 128).
template <int D16, bool L2, int CAP, int NT>
__global__ __launch_bounds__(NT) void rescore_share_kernel(
    const float* __restrict__ base, uint32_t dp, uint32_t d, const float* __restrict__ queries,
    const uint32_t* __restrict__ ks, uint32_t k, const float* __restrict__ E,
    const uint64_t* __restrict__ cand, const uint32_t* __restrict__ count, uint32_t ldo,
    uint32_t* __restrict__ overflow_flags, uint32_t* __restrict__ n_overflow,
    uint32_t* __restrict__ rescored, uint32_t cmin, int final_pass, const uint32_t* __restrict__ qmap,
    uint32_t split, uint64_t* __restrict__ rpart) {
  extern __shared__ uint64_t dyn[];
  uint64_t* buf = dyn;                                // CAP keys
  float* q_lds = reinterpret_cast<float*>(dyn + CAP);  // dp floats
  __shared__ uint32_t nsel, nall, ntop;
  __shared__ uint32_t hist[256], sel[2];
  __shared__ uint64_t top[kTopCap];
  // block b: share b % split of query b / split (qmap: of query qmap[b / split])
  const uint32_t bq = blockIdx.x / split, share = blockIdx.x % split;
  const uint32_t q = qmap ? qmap[bq] : bq;
  if (q == kNoneId) return;
  const uint32_t tid = threadIdx.x, lane = tid & 63, t = lane & 3, g = lane >> 2;
  const uint32_t kq = min(ks ? ks[q] : k, ldo);
  const uint32_t c = count[q];
  if (c < cmin) return;  // done by the previous launch
  if (c > uint32_t(CAP)) {
    if (final_pass && tid == 0 && share == 0) {
      overflow_flags[q] = 1;
      rescored[q] = 0;
      atomicAdd(n_overflow, 1u);
    }
    return;
  }
  if (tid == 0 && share == 0) overflow_flags[q] = 0;
  for (uint32_t i = tid; i < dp; i += NT) q_lds[i] = i < d ? queries[size_t(q) * d + i] : 0.f;
  // Dk = the kk-th smallest estimate; keep dh <= Dk + 2E (compacted, in no particular order; a
  // share keeps its own candidates and counts them all)
  const uint64_t* cq = cand + size_t(q) * kCandCap;
  for (uint32_t i = tid; i < c; i += NT) buf[i] = cq[i];
  if (tid == 0) nsel = nall = 0;
  __syncthreads();
  const uint32_t kk = min(kq, c);
  if (kk > 0) {
    const float Dk = unord_bits(block_select_hi32<NT>(buf, c, kk, hist, sel, tid));
    const float lim = round_up_f(double(Dk) + 2.0 * double(E[q]) * (1.0 + 1e-6) + 0x1p-20);
    __syncthreads();  // every read of buf by the select is done
    for (uint32_t i = tid; i < c; i += NT) {
      const uint64_t v = cq[i];
      if (key_dist(v) <= lim) {
        atomicAdd(&nall, 1u);
        if (i % split == share) buf[atomicAdd(&nsel, 1u)] = v;
      }
    }
  }
  __syncthreads();
  const uint32_t r = nsel;
  if (tid == 0 && share == 0) rescored[q] = nall;
  // exact distances, 16 rows per wave instruction group, NT / 64 waves; a quad reads its
  // candidate's row id and later overwrites that slot with the exact key
  for (uint32_t base_i = (tid >> 6) * 16; base_i < r; base_i += 16 * (NT / 64)) {
    const uint32_t i = base_i + g;
    bool ok[1] = {i < r};
    const uint32_t row = key_id(buf[ok[0] ? i : 0]);
    const float* xr[1] = {base + size_t(row) * dp};
    float dist[1][1];
    quad_dists<1, 1, D16, L2, 16, kLoadPlain>(xr, ok, q_lds, dp, d, int(t), dist);
    if (ok[0] && t == 0) buf[i] = make_key(dist[0][0], row);
  }
  __syncthreads();
  // this share's first kq keys (kPadKey past its r) to its part
  uint64_t* rp = rpart + (size_t(q) * split + share) * ldo;
  rank_first<NT>(buf, r, kq, hist, sel, top, &ntop, tid, [&](uint32_t rank, uint64_t v) { rp[rank] = v; });
  for (uint32_t i = r + tid; i < kq; i += NT) rp[i] = kPadKey;
}

// The output of the queries whose rescore was split: query q's split parts (share s at
// rpart + (q * split + s) * ldo, kq keys each, kPadKey where a share had fewer) ranked together.
// One 256-thread block per query; queries above the candidate capacity are left to the exact
// fallback, as by rescore_kernel.
__global__ __launch_bounds__(256) void rescore_merge_kernel(const uint32_t* __restrict__ ks, uint32_t k,
                                                            const uint32_t* __restrict__ count, uint32_t split,
                                                            const uint64_t* __restrict__ rpart,
                                                            uint32_t* __restrict__ out_ids,
                                                            float* __restrict__ out_dists, uint32_t ldo) {
  extern __shared__ uint64_t dyn[];  // the next power of two >= max(512, split * ldo) keys
  __shared__ uint32_t nsel, ntop;
  __shared__ uint32_t hist[256], sel[2];
  __shared__ uint64_t top[kTopCap];
  const uint32_t q = blockIdx.x, tid = threadIdx.x;
  if (count[q] > uint32_t(kCandCap)) return;
  const uint32_t kq = min(ks ? ks[q] : k, ldo);
  uint32_t* oi = out_ids + size_t(q) * ldo;
  float* od = out_dists + size_t(q) * ldo;
  if (tid == 0) nsel = 0;
  __syncthreads();
  const uint64_t* parts = rpart + size_t(q) * split * ldo;
  for (uint32_t i = tid; i < split * kq; i += 256) {
    const uint32_t s = i / kq;
    const uint64_t v = parts[size_t(s) * ldo + (i - s * kq)];
    if (v != kPadKey) dyn[atomicAdd(&nsel, 1u)] = v;
  }
  __syncthreads();
  const uint32_t r = nsel;
  rank_first<256>(dyn, r, kq, hist, sel, top, &ntop, tid, [&](uint32_t rank, uint64_t v) {
    oi[rank] = key_id(v);
    od[rank] = key_dist(v);
  });
  for (uint32_t i = min(r, kq) + tid; i < ldo; i += 256) {
    oi[i] = kNoneId;
    od[i] = __builtin_inff();
  }
}

// ---- rescore order ------------------------------------------------------------------------
// Queries whose candidate sets overlap read the same base rows. The rescore of a large batch
// runs them one after another on one XCD, so that the rows the first one reads are in that
// XCD's L2 (or the Infinity Cache) when the others read them: queries are ordered by the
// smallest row id among their candidates (a min-hash: two queries with the same candidates get
// the same value, largely overlapping sets mostly do), and the sorted order is dealt to the XCDs
// in eight contiguous runs (workgroup b runs on XCD b % 8, so block b takes sorted position
// (b % 8) * per + b / 8). The order changes no result: each query's rescore is independent.
// The FP8 shortlist takes it (at 1M x 384, batch 4,096: rescore 0.94 -> 0.73 ms): its error
// window spans a query's whole neighbourhood, so queries of one neighbourhood share most of their
// candidates. The BF16 window is narrower; the same order measured no gain there.

// sig[q] = the smallest candidate row id of query q (kNoneId without candidates or above the
// buffer); one wave per query.
__global__ __launch_bounds__(256) void cand_signature_kernel(const uint64_t* __restrict__ cand,
                                                             const uint32_t* __restrict__ count, uint32_t nq,
                                                             uint32_t* __restrict__ sig) {
  const uint32_t q = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (q >= nq) return;
  const uint32_t c = count[q];
  uint32_t m = kNoneId;
  if (c <= uint32_t(kCandCap))
    for (uint32_t i = lane; i < c; i += 64) m = min(m, key_id(cand[size_t(q) * kCandCap + i]));
  for (int off = 32; off >= 1; off >>= 1) m = min(m, uint32_t(__shfl_xor(int(m), off, 64)));
  if (lane == 0) sig[q] = m;
}

// One block: orders nq <= kOrderMax queries by signature and writes qmap[0, 8 * per) (per =
// ceil(nq / 8)): block b of the rescore takes ordered position p = (b % 8) * per + b / 8, or no
// query when p >= nq. The order is a bucket sort on a multiplicative hash of the signature
// (kOrderBuckets buckets, one more for queries without candidates): queries with the same
// signature land in one bucket and so side by side, and different signatures, which crowd the
// small row ids (a signature is a minimum), spread over the buckets instead of sharing a few.
constexpr uint32_t kOrderMax = 4096;
constexpr uint32_t kOrderBuckets = 4096;
__global__ __launch_bounds__(1024) void rescore_order_kernel(const uint32_t* __restrict__ sig, uint32_t nq,
                                                             uint32_t* __restrict__ qmap) {
  static_assert(kOrderBuckets == 4096, "12-bit hash");
  __shared__ uint32_t cnt[kOrderBuckets + 1], sorted[kOrderMax], wsum[16];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (uint32_t i = tid; i <= kOrderBuckets; i += 1024) cnt[i] = 0;
  __syncthreads();
  uint32_t bk[kOrderMax / 1024], rk[kOrderMax / 1024];  // bucket and rank within it, per query
#pragma unroll
  for (uint32_t j = 0; j < kOrderMax / 1024; ++j) {
    const uint32_t q = tid + j * 1024;
    if (q < nq) {
      const uint32_t v = sig[q];
      bk[j] = v != kNoneId ? (v * 0x9E3779B1u) >> 20 : kOrderBuckets;
      rk[j] = atomicAdd(&cnt[bk[j]], 1u);
    }
  }
  __syncthreads();
  // exclusive scan of cnt[0 .. kOrderBuckets] (4 buckets a thread, the last one by thread 0)
  uint32_t c[4], t = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    c[i] = cnt[tid * 4 + i];
    t += c[i];
  }
  uint32_t incl = t;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(incl, off, 64);
    if (lane >= uint32_t(off)) incl += y;
  }
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  uint32_t before = 0;
  for (uint32_t w = 0; w < wave; ++w) before += wsum[w];
  uint32_t run = before + incl - t;
  __syncthreads();  // every cnt[] is read before it is overwritten with its offset
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    cnt[tid * 4 + i] = run;
    run += c[i];
  }
  if (tid == 1023) cnt[kOrderBuckets] = run;  // after every bucket: the queries without candidates
  __syncthreads();
#pragma unroll
  for (uint32_t j = 0; j < kOrderMax / 1024; ++j) {
    const uint32_t q = tid + j * 1024;
    if (q < nq) sorted[cnt[bk[j]] + rk[j]] = q;
  }
  __syncthreads();
  const uint32_t per = (nq + 7) / 8;
  for (uint32_t b = tid; b < 8 * per; b += 1024) {
    const uint32_t p = (b % 8) * per + b / 8;
    qmap[b] = p < nq ? sorted[p] : kNoneId;
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

hipError_t launch_query_prep(hipStream_t st, const float* queries, uint32_t d, uint32_t nq, uint32_t KS, int prec,
                             void* out, float* qscale, const unsigned long long* maxes, float* E, float* aux,
                             uint32_t* flags, bool l2) {
  hipLaunchKernelGGL(query_prep_kernel, dim3((std::max(nq, 1u) + 3) / 4), dim3(256), 0, st, queries, d, nq, KS, prec,
                     prec == 1 ? static_cast<__bf16*>(out) : nullptr,
                     prec == 2 ? static_cast<uint8_t*>(out) : nullptr, qscale, maxes, E, aux, flags, l2 ? 1 : 0);
  return hipGetLastError();
}

// hn[r] = |x_r|^2 / 2 in fp32 (from the FP64 norms) for rows [r0, r1): the L2 term of the estimate
__global__ void half_norm_kernel(const double* __restrict__ norms, uint32_t r0, uint32_t r1, float* __restrict__ hn) {
  const uint32_t r = r0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (r < r1) hn[r] = float(0.5 * norms[size_t(r) * 3]);
}

hipError_t launch_half_norms(hipStream_t st, const double* norms, uint32_t r0, uint32_t r1, float* hn) {
  if (r1 <= r0) return hipSuccess;
  hipLaunchKernelGGL(half_norm_kernel, dim3((r1 - r0 + 255) / 256), dim3(256), 0, st, norms, r0, r1, hn);
  return hipGetLastError();
}

// query fragments are staged G at a time: callers pad fragment and aux buffers to whole stages,
// and the aux buffer by short_aux_pad() more floats (a stage reads whole 256-byte loads)
int short_group(int prec, int ks) { return stage_group(prec, ks); }
int short_aux_pad() { return 64; }
int short_aux_floats() { return kAux; }

int short_waves() { return kWaves; }

hipError_t launch_scatter(int prec, const ShortScanArgs& a) {
  RegionGeometry geo;
  uint32_t n_regions;
  if (a.qs_blocks) {  // queries-stationary pass 2
    n_regions = a.qs_blocks * a.qs_groups * uint32_t(kQsWaves);
    geo = RegionGeometry{uint32_t(kQsWaves), a.qs_blocks, a.qs_rows, kQsShift};
  } else {
    n_regions = a.n_blocks * uint32_t(kWaves);
    geo = RegionGeometry{1u, n_regions, uint32_t(rows_frags(prec, a.KS) * 16), 6u};
  }
  if (n_regions == 0 || a.nq == 0) return hipErrorInvalidValue;  // count[] is written only below
  // about 16 regions a slice (4 a wave), at most kScatterSlices slices, and at most 2^20 / nq so
  // that the per-slice counters of a large batch (nq words each) stay well below its entries
  // (at 4,096 queries: 256 slices, 4 MB of counters for about 18 MB of entries)
  const uint32_t S = std::min({kScatterSlices, std::max(1u, (n_regions + 15) / 16),
                               std::max(64u, (1u << 20) / a.nq)});
  const uint32_t per = (n_regions + S - 1) / S;
  const size_t lds = size_t(a.nq) * 4;
  hipLaunchKernelGGL(scatter_count_kernel, dim3(S), dim3(kScatterThreads), lds, a.st, a.emit, a.emit_count,
                     a.wave_cap, geo, n_regions, per, a.nq, S, a.slice, a.overflow);
  // a small batch's place blocks sum their own offsets (one launch less, about 5 us)
  const int own = a.nq <= kPlaceOwnQ ? 1 : 0;
  if (!own)
    hipLaunchKernelGGL(scatter_offsets_kernel, dim3((a.nq + 3) / 4), dim3(256), 0, a.st, a.slice, S, a.nq,
                       a.count);
  hipLaunchKernelGGL(scatter_place_kernel, dim3(S), dim3(kScatterThreads), lds, a.st, a.emit, a.emit_count,
                     a.wave_cap, geo, n_regions, per, a.nq, S, a.slice, a.cand, a.cap, own, a.count);
  return hipGetLastError();
}

size_t short_scatter_slices() { return kScatterSlices; }

namespace {
template <int PREC, int KS>
hipError_t launch_qs_t(const ShortScanArgs& a) {
  if (a.qs_groups <= kQsNtGroups) {
    hipLaunchKernelGGL((mfma_qs_kernel<PREC, KS, true>), dim3(a.qs_blocks * a.qs_groups), dim3(64 * kQsWaves), 0,
                       a.st, a.xfrag, a.xscale, a.n, a.qfrag, a.qaux, a.nq, a.qs_rows, a.qs_blocks, a.qs_groups,
                       a.emit, a.wave_cap, a.emit_count);
    return hipGetLastError();
  }
  hipLaunchKernelGGL((mfma_qs_kernel<PREC, KS, false>), dim3(a.qs_blocks * a.qs_groups), dim3(64 * kQsWaves), 0, a.st,
                     a.xfrag, a.xscale, a.n, a.qfrag, a.qaux, a.nq, a.qs_rows, a.qs_blocks, a.qs_groups, a.emit,
                     a.wave_cap, a.emit_count);
  return hipGetLastError();
}
template <int KS>
hipError_t launch_qk_t(const ShortScanArgs& a) {
  hipLaunchKernelGGL((mfma_qk_kernel<KS>), dim3(a.qs_blocks * a.qs_groups), dim3(64 * kQsWaves), 0, a.st, a.xfrag,
                     a.n, a.qfrag, a.qaux, a.nq, a.qs_rows, a.qs_blocks, a.qs_groups, a.emit, a.wave_cap,
                     a.emit_count);
  return hipGetLastError();
}
template <int PREC>
hipError_t launch_qs_p(const ShortScanArgs& a) {
  if constexpr (PREC == 1) {  // BF16, one query group, widths whose k-steps split over 4 waves
    if (a.qs_groups == 1) {
      if (a.KS == 4) return launch_qk_t<4>(a);
      if (a.KS == 8) return launch_qk_t<8>(a);
      if (a.KS == 12) return launch_qk_t<12>(a);
    }
  }
  switch (a.KS) {
    case 2: return launch_qs_t<PREC, 2>(a);
    case 4: return launch_qs_t<PREC, 4>(a);
    case 8: return launch_qs_t<PREC, 8>(a);
    case 12: return launch_qs_t<PREC, 12>(a);
    case 24: return launch_qs_t<PREC, 24>(a);
    default: return hipErrorInvalidValue;
  }
}
}  // namespace

bool short_qs_plan(int prec, int ks, uint32_t n, uint32_t nq, int num_cus, uint32_t* blocks, uint32_t* groups,
                   uint32_t* rows, uint32_t* queries_per_region) {
  if (ks != 2 && ks != 4 && ks != 8 && ks != 12 && ks != 24) return false;
  const uint32_t qpg = uint32_t(kQsWaves * qs_qw(prec, ks) * 16);  // queries per query group
  const uint32_t g = (nq + qpg - 1) / qpg;
  if (nq == 0 || g > kQsMaxGroups || n == 0 || num_cus <= 0) return false;
  // blocks for all groups: about qs_bpc per CU (one pass over the grid; the k-split kernel,
  // which takes a single BF16 group, runs two), rows a multiple of the stage's rows, row ranges
  // a multiple of 8 (the XCD interleave of the groups)
  const uint32_t unit = 16u * uint32_t(qs_sf(prec, ks));
  const uint32_t bpc = qk_supported(prec, ks) && g == 1 ? 2u : uint32_t(qs_bpc(prec, ks));
  const uint32_t nb = std::max(8u, uint32_t(num_cus) * bpc / g);
  uint32_t r = (n + nb - 1) / nb;
  r = std::min(kQsMaxRows, (r + unit - 1) / unit * unit);
  *rows = r;
  *blocks = ((n + r - 1) / r + 7) / 8 * 8;
  *groups = g;
  *queries_per_region = uint32_t(qs_qw(prec, ks)) * 16;
  return true;
}

hipError_t launch_mfma_qs(int prec, const ShortScanArgs& a) {
  if (a.nq == 0 || a.qs_blocks == 0 || a.qs_groups == 0 || a.qs_rows == 0 || a.qs_rows > kQsMaxRows ||
      a.qs_rows % (16u * uint32_t(qs_sf(prec, a.KS))) != 0)
    return hipErrorInvalidValue;
  return prec == 1 ? launch_qs_p<1>(a) : launch_qs_p<2>(a);
}

hipError_t launch_zero_u32(hipStream_t st, uint32_t* p, size_t n) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(zero_u32_kernel, dim3(unsigned(std::min<size_t>(1024, (n + 255) / 256))), dim3(256), 0,
                     st, p, n);
  return hipGetLastError();
}

namespace {
template <int PREC, int KS, bool L2>
hipError_t launch_mfma_ksm(int pass, const ShortScanArgs& a) {
  const size_t lds = 0;  // static: four ring slots and the reduction buffer
  const dim3 grid(a.n_blocks), block(64 * kWaves);
  if (pass == 1)
    hipLaunchKernelGGL((mfma_scan_kernel<PREC, KS, 1, L2>), grid, block, lds, a.st, a.xfrag, a.xscale, a.xhn,
                       a.n, a.qfrag, a.qaux, a.nq, a.tstride, a.n_space, a.qsplit, a.tilemin, a.emit, a.wave_cap,
                       a.emit_count);
  else
    hipLaunchKernelGGL((mfma_scan_kernel<PREC, KS, 2, L2>), grid, block, lds, a.st, a.xfrag, a.xscale, a.xhn,
                       a.n, a.qfrag, a.qaux, a.nq, a.tstride, a.n_space, 1u, a.tilemin, a.emit, a.wave_cap,
                       a.emit_count);
  return hipGetLastError();
}
// L2 (the rows' |x|^2 / 2 given) or inner product
template <int PREC, int KS>
hipError_t launch_mfma_ks(int pass, const ShortScanArgs& a) {
  return a.xhn ? launch_mfma_ksm<PREC, KS, true>(pass, a) : launch_mfma_ksm<PREC, KS, false>(pass, a);
}

template <int PREC>
hipError_t launch_mfma_p(int pass, const ShortScanArgs& a) {
  switch (a.KS) {
    case 2: return launch_mfma_ks<PREC, 2>(pass, a);
    case 4: return launch_mfma_ks<PREC, 4>(pass, a);
    case 8: return launch_mfma_ks<PREC, 8>(pass, a);
    case 12: return launch_mfma_ks<PREC, 12>(pass, a);
    case 24: return launch_mfma_ks<PREC, 24>(pass, a);
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
  if (ks <= 24) return 24;
  return 0;  // too wide for the register-resident row operands
}

// rows per tile of the MFMA scan for this precision and width (8 waves x row fragments x 16)
int short_tile_rows_for(int prec, int ks) { return kWaves * rows_frags(prec, ks) * 16; }

hipError_t launch_mfma_scan(int prec, int pass, const ShortScanArgs& a) {
  if (a.nq == 0 || a.n_blocks == 0) return hipErrorInvalidValue;  // a block streams at least one stage
  if (pass == 1 && (a.qsplit == 0 || a.n_blocks != a.n_space * a.qsplit)) return hipErrorInvalidValue;
  if (pass == 2 && a.n_blocks != a.n_space) return hipErrorInvalidValue;
  return prec == 1 ? launch_mfma_p<1>(pass, a) : launch_mfma_p<2>(pass, a);
}

hipError_t launch_select_threshold(hipStream_t st, const uint32_t* tilemin, uint32_t n_tiles, uint32_t nq,
                                   uint32_t nq_pad, const uint32_t* ks, uint32_t k, const float* E,
                                   const float* qscale, float* thr, float* aux) {
  const dim3 wg((nq_pad + 3) / 4), b(256);
  if (n_tiles <= 128)
    hipLaunchKernelGGL(select_threshold_wave_kernel<2>, wg, b, 0, st, tilemin, n_tiles, nq, nq_pad, ks, k, E, qscale,
                       thr, aux);
  else if (n_tiles <= 256)
    hipLaunchKernelGGL(select_threshold_wave_kernel<4>, wg, b, 0, st, tilemin, n_tiles, nq, nq_pad, ks, k, E, qscale,
                       thr, aux);
  else if (n_tiles <= 512)
    hipLaunchKernelGGL(select_threshold_wave_kernel<8>, wg, b, 0, st, tilemin, n_tiles, nq, nq_pad, ks, k, E, qscale,
                       thr, aux);
  else if (n_tiles <= 1024)
    hipLaunchKernelGGL(select_threshold_wave_kernel<16>, wg, b, 0, st, tilemin, n_tiles, nq, nq_pad, ks, k, E,
                       qscale, thr, aux);
  else
    hipLaunchKernelGGL(select_threshold_kernel, dim3(nq_pad), b, 0, st, tilemin, n_tiles, nq, ks, k, E, qscale, thr,
                       aux);
  return hipGetLastError();
}

hipError_t launch_rescore(hipStream_t st, bool l2, const float* base, uint32_t dp, uint32_t d,
                          const float* queries, uint32_t nq, const uint32_t* ks, uint32_t k,
                          const float* E, const uint64_t* cand, const uint32_t* count, uint32_t* ids,
                          float* dists, uint32_t ldo, uint32_t* flags, uint32_t* n_overflow,
                          uint32_t* rescored, uint32_t* sig, uint32_t* qmap, uint32_t split, uint64_t* rpart,
                          size_t rpart_keys) {
  // sig / qmap given: order the queries of a large batch (rescore_order_kernel)
  // every query's flag is set or cleared by the launch that handles it
  const size_t lds_small = size_t(kCandSmall) * sizeof(uint64_t) + size_t(dp) * sizeof(float);
  const size_t lds = size_t(kCandCap) * sizeof(uint64_t) + size_t(dp) * sizeof(float);
  const bool wide = nq <= 256;  // fewer queries (blocks) than CUs: 16 waves per query, one launch
  // shares per query (wide launches only; the caller sizes rpart for them)
  if (!wide || !rpart || split == 0 || size_t(split) * ldo > size_t(kCandCap) ||
      size_t(nq) * split * ldo > rpart_keys)
    split = 1;
  // larger batches: queries in min-hash order, XCD by XCD (rescore_order_kernel)
  const uint32_t* order = nullptr;
  uint32_t nb = nq;
  if (!wide && sig && qmap && nq <= kOrderMax) {
    hipLaunchKernelGGL(cand_signature_kernel, dim3((nq + 3) / 4), dim3(256), 0, st, cand, count, nq, sig);
    hipLaunchKernelGGL(rescore_order_kernel, dim3(1), dim3(1024), 0, st, sig, nq, qmap);
    order = qmap;
    nb = 8 * ((nq + 7) / 8);
  }
#define VSG_RESCORE(D16, L2V)                                                                         \
  do {                                                                                                \
    if (wide) { /* one block per query either way: one launch with the full buffer */               \
      const void* f_ = split > 1 ? reinterpret_cast<const void*>(&rescore_share_kernel<D16, L2V, kCandCap, 1024>) \
                                 : reinterpret_cast<const void*>(&rescore_kernel<D16, L2V, kCandCap, 1024>); \
      hipError_t w_ = hipFuncSetAttribute(f_, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));  \
      if (w_ != hipSuccess) return w_;                                                                \
      if (split == 1)                                                                                 \
        hipLaunchKernelGGL((rescore_kernel<D16, L2V, kCandCap, 1024>), dim3(nq), dim3(1024), lds, st, \
                           base, dp, d, queries, ks, k, E, cand, count, ids, dists, ldo, flags,       \
                           n_overflow, rescored, 0u, 1, nullptr);                                     \
      else                                                                                            \
        hipLaunchKernelGGL((rescore_share_kernel<D16, L2V, kCandCap, 1024>), dim3(nq * split),        \
                           dim3(1024), lds, st, base, dp, d, queries, ks, k, E, cand, count, ldo,     \
                           flags, n_overflow, rescored, 0u, 1, nullptr, split, rpart);                \
      if (split > 1) { /* split * ldo <= kCandCap keys; the sort of rank_first pads to 2^j >= 512 */ \
        size_t m2_ = 512;                                                                             \
        while (m2_ < size_t(split) * ldo) m2_ <<= 1;                                                  \
        const size_t lds_m = m2_ * sizeof(uint64_t);                                                  \
        hipError_t m_ = hipFuncSetAttribute(reinterpret_cast<const void*>(&rescore_merge_kernel),      \
                                            hipFuncAttributeMaxDynamicSharedMemorySize, int(lds_m));   \
        if (m_ != hipSuccess) return m_;                                                              \
        hipLaunchKernelGGL(rescore_merge_kernel, dim3(nq), dim3(256), lds_m, st, ks, k, count, split,  \
                           rpart, ids, dists, ldo);                                                   \
      }                                                                                               \
      break;                                                                                          \
    }                                                                                                 \
    hipLaunchKernelGGL((rescore_kernel<D16, L2V, kCandSmall, 256>), dim3(nb), dim3(256), lds_small,    \
                       st, base, dp, d, queries, ks, k, E, cand, count, ids, dists, ldo, flags,       \
                       n_overflow, rescored, 0u, 0, order);                     \
    hipError_t e_ = hipFuncSetAttribute(                                                              \
        reinterpret_cast<const void*>(&rescore_kernel<D16, L2V, kCandCap, 256>),                      \
        hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));                                        \
    if (e_ != hipSuccess) return e_;                                                                  \
    hipLaunchKernelGGL((rescore_kernel<D16, L2V, kCandCap, 256>), dim3(nb), dim3(256), lds, st, base, dp, \
                       d, queries, ks, k, E, cand, count, ids, dists, ldo, flags, n_overflow,          \
                       rescored, uint32_t(kCandSmall) + 1, 1, order);           \
  } while (0)
  if (l2) {
    if (d == 384) VSG_RESCORE(24, true);
    else if (d == 768) VSG_RESCORE(48, true);
    else if (d == 128) VSG_RESCORE(8, true);
    else VSG_RESCORE(0, true);
  } else {
    if (d == 384) VSG_RESCORE(24, false);
    else if (d == 768) VSG_RESCORE(48, false);
    else if (d == 128) VSG_RESCORE(8, false);
    else VSG_RESCORE(0, false);
  }
#undef VSG_RESCORE
  return hipGetLastError();
}

int short_cand_cap() { return kCandCap; }

// Blocks per query of the rescore of a small batch: about one block per CU in all, at most 32
// shares, and the shares' first-k keys within one candidate buffer for the final rank (at
// 1M x 384: FP8 batch 8 and 32 +15 %, batch 128 with 2 shares +4 %).
uint32_t short_rescore_split(uint32_t nq, uint32_t k, int num_cus) {
  if (nq == 0 || nq > 256 || num_cus <= 0) return 1;
  uint32_t s = std::max(1u, uint32_t(num_cus) / nq);
  s = std::min(s, 32u);
  s = std::min(s, std::max(1u, uint32_t(kCandCap) / std::max(1u, k)));
  return s;
}


}  // namespace vsg
