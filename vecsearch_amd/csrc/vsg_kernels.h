// Device helpers shared by the vsg kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vsg {

constexpr uint64_t kPadKey = ~0ull;  // sorts after every real (dist, id); decodes to (inf, kNone)

// Monotone map of a float onto uint32 (total order; -0 and +0 are canonicalised to +0 so the
// order matches IEEE `<` on every non-NaN distance).
__device__ __forceinline__ uint32_t ord_bits(float f) {
  uint32_t u = __float_as_uint(f == 0.f ? 0.f : f);
  return u ^ ((u & 0x80000000u) ? 0xFFFFFFFFu : 0x80000000u);
}
__device__ __forceinline__ float unord_bits(uint32_t o) {
  uint32_t u = o ^ ((o & 0x80000000u) ? 0x80000000u : 0xFFFFFFFFu);
  return __uint_as_float(u);
}
// A candidate as one 64-bit key: (dist, id) lexicographic order == unsigned order of the key.
// This is the tie rule of the reference merge (brute_force.cpp:159-161).
__device__ __forceinline__ uint64_t make_key(float d, uint32_t id) {
  return (uint64_t(ord_bits(d)) << 32) | id;
}
__device__ __forceinline__ uint32_t key_id(uint64_t k) { return uint32_t(k); }
__device__ __forceinline__ float key_dist(uint64_t k) {
  return k == kPadKey ? __builtin_inff() : unord_bits(uint32_t(k >> 32));
}

// quad_perm DPP controls: lane t of each group of four reads lane perm[t].
constexpr int kDppXor2 = 0x4E;  // [2,3,0,1]
constexpr int kDppXor1 = 0xB1;  // [1,0,3,2]

template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

__device__ __forceinline__ uint64_t readlane64(uint64_t v, int lane) {
  const uint32_t lo = __builtin_amdgcn_readlane(uint32_t(v), lane);
  const uint32_t hi = __builtin_amdgcn_readlane(uint32_t(v >> 32), lane);
  return (uint64_t(hi) << 32) | lo;
}

__device__ __forceinline__ uint64_t shfl_up1_64(uint64_t v, int lane) {
  // value of lane-1 (lane 0 keeps its own value; callers overwrite it): DPP wave_shr:1 (0x138),
  // a VALU move, not an LDS round trip
  (void)lane;
  const uint32_t lo = uint32_t(__builtin_amdgcn_update_dpp(int(uint32_t(v)), int(uint32_t(v)), 0x138, 0xF, 0xF, false));
  const uint32_t hi = uint32_t(__builtin_amdgcn_update_dpp(int(uint32_t(v >> 32)), int(uint32_t(v >> 32)), 0x138, 0xF, 0xF, false));
  return (uint64_t(hi) << 32) | lo;
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// Sorted k-list of one query, spread over the wave: slot s lives in lane s & 63 of register
// s >> 6 (KP = 64 or 128 slots). insert() keeps the list sorted ascending and returns the new
// k-th key (the admission threshold). The key must be smaller than the current k-th key.
template <int KP>
struct WaveList {
  uint64_t e[KP / 64];

  __device__ __forceinline__ void init() {
#pragma unroll
    for (int r = 0; r < KP / 64; ++r) e[r] = kPadKey;
  }

  __device__ __forceinline__ uint64_t insert(uint64_t key, int lane, int k) {
    int pos = 0;
#pragma unroll
    for (int r = 0; r < KP / 64; ++r) pos += __popcll(ballot(e[r] < key));
    uint64_t prev[KP / 64];
#pragma unroll
    for (int r = 0; r < KP / 64; ++r) {
      prev[r] = shfl_up1_64(e[r], lane);
      if (lane == 0) prev[r] = r == 0 ? kPadKey : readlane64(e[r - 1], 63);
    }
#pragma unroll
    for (int r = 0; r < KP / 64; ++r) {
      const int s = r * 64 + lane;
      e[r] = s < pos ? e[r] : (s == pos ? key : prev[r]);
    }
    const int kk = k - 1;
    uint64_t thr = kPadKey;
#pragma unroll
    for (int r = 0; r < KP / 64; ++r)
      if ((kk >> 6) == r) thr = readlane64(e[r], kk & 63);
    return thr;
  }
};

template <bool L2>
__device__ __forceinline__ float fma_term(float q, float x, float acc) {
  if (L2) {
    const float e = q - x;  // brute_force.cpp:28-31 (q0 = q - x; fma(q0, q0, acc))
    return __builtin_fmaf(e, e, acc);
  }
  return __builtin_fmaf(q, x, acc);  // brute_force.cpp:33-34
}

// Row addressing. A row's 16-dim step s starts at xr + s * STEP floats and element j sits at
// xr + (j >> 4) * STEP + (j & 15):
//   STEP = 16   row-major base (rows of dp floats);
//   STEP = 256  16-row tiles, step-major: the 64-byte step-s blocks of a tile's 16 rows are
//               contiguous, so a wave instruction over 16 rows reads 1 KB in one piece.
// Load modes: kLoadIfOk skips rows with ok == false (masked scans), kLoadPlain and
// kLoadStream always load (the caller clamps xr to a readable row); kLoadStream marks the
// loads non-temporal (the base is streamed once per pass and should not displace L2 data).
enum LoadMode { kLoadIfOk = 0, kLoadPlain = 1, kLoadStream = 2 };

typedef float f32x4_t __attribute__((ext_vector_type(4)));

template <int MODE>
__device__ __forceinline__ float4 load_step(const float* p, bool ok) {
  if (MODE == kLoadIfOk) return ok ? *reinterpret_cast<const float4*>(p) : make_float4(0.f, 0.f, 0.f, 0.f);
  if (MODE == kLoadStream) {
    const f32x4_t v = __builtin_nontemporal_load(reinterpret_cast<const f32x4_t*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
  }
  return *reinterpret_cast<const float4*>(p);
}

// Scalar tail of d % 8 dims after the tree (distance.h:50, 68), as the reference's release
// build computes it: one block of 4 rounded products added in order, then fused terms.
template <bool L2, int STEP>
__device__ __forceinline__ float scalar_tail(const float* q, const float* __restrict__ x, int i,
                                             int d, float r) {
  if (i + 4 <= d) {
#pragma unroll
    for (int j = 0; j < 4; ++j, ++i) {
      const float a = q[i], b = x[(i >> 4) * STEP + (i & 15)];
      const float p = L2 ? __fmul_rn(a - b, a - b) : __fmul_rn(a, b);
      r = __fadd_rn(r, p);
    }
  }
  for (; i < d; ++i) r = fma_term<L2>(q[i], x[(i >> 4) * STEP + (i & 15)], r);
  return r;
}

// Reference-order distances of R rows x QB queries for the calling quad lane t.
// xr[j]: start of row j (see STEP); ok[j] marks rows that carry a result.
// q: QB queries in LDS, query b at q + b * dp (zero beyond d). Every lane of a quad returns the
// same dist[j][b]; all 64 lanes of the wave must call this together (DPP across the quad).
//
// Accumulator layout: lane t keeps the reference's accumulators 4t..4t+3, i.e. dims 4t..4t+3
// mod 16 (acc0 = accumulators 0..7, acc1 = 8..15 in distance.h:57-61). After the 16-dim steps
// (and the 8-dim step into acc0), the tree of distance.h:18-27 runs across the quad with DPP:
// v = acc + acc[lane ^ 2], w = v + v[lane ^ 1], r = (w0 + w1) + (w2 + w3).
template <int QB, int R, int D16, bool L2, int STEP, int MODE>
__device__ __forceinline__ void quad_dists(const float* const (&xr)[R], const bool (&ok)[R],
                                           const float* q, uint32_t dp, uint32_t d, int t,
                                           float (&dist)[R][QB]) {
  float acc[R][QB][4];
#pragma unroll
  for (int j = 0; j < R; ++j)
#pragma unroll
    for (int b = 0; b < QB; ++b)
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[j][b][c] = 0.f;

  const uint32_t nb16 = D16 > 0 ? uint32_t(D16) : d >> 4;
  auto fma_step = [&](const float4 (&xv)[R], uint32_t s) __attribute__((always_inline)) {
#pragma unroll
    for (int b = 0; b < QB; ++b) {
      const float4 qv = *reinterpret_cast<const float4*>(q + b * dp + 16 * s + 4 * t);
#pragma unroll
      for (int j = 0; j < R; ++j) {
        acc[j][b][0] = fma_term<L2>(qv.x, xv[j].x, acc[j][b][0]);
        acc[j][b][1] = fma_term<L2>(qv.y, xv[j].y, acc[j][b][1]);
        acc[j][b][2] = fma_term<L2>(qv.z, xv[j].z, acc[j][b][2]);
        acc[j][b][3] = fma_term<L2>(qv.w, xv[j].w, acc[j][b][3]);
      }
    }
  };
  if constexpr (MODE == kLoadStream && D16 > 0 && D16 % 8 == 0 && QB >= 4) {
    // Several queries per pass: enough FMAs per step that the compiler no longer hoists later
    // steps' loads above them. Two register sets of four steps each: the next four steps' loads
    // are issued before the current four steps' FMAs (past the last step: the last step again,
    // never read); the scheduling barriers keep that order.
    float4 xa[4][R], xb[4][R];
    auto load4 = [&](float4 (&dst)[4][R], uint32_t s0) __attribute__((always_inline)) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const uint32_t s = min(s0 + i, uint32_t(D16 - 1));
#pragma unroll
        for (int j = 0; j < R; ++j) dst[i][j] = load_step<MODE>(xr[j] + s * STEP + 4 * t, ok[j]);
      }
    };
    load4(xa, 0);
    for (uint32_t s = 0; s < uint32_t(D16); s += 8) {
      load4(xb, s + 4);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int i = 0; i < 4; ++i) fma_step(xa[i], s + i);
      __builtin_amdgcn_sched_barrier(0);
      load4(xa, s + 8);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int i = 0; i < 4; ++i) fma_step(xb[i], s + 4 + i);
      __builtin_amdgcn_sched_barrier(0);
    }
  } else {
#pragma unroll 4
    for (uint32_t s = 0; s < nb16; ++s) {
      float4 xv[R];
#pragma unroll
      for (int j = 0; j < R; ++j) xv[j] = load_step<MODE>(xr[j] + s * STEP + 4 * t, ok[j]);
      fma_step(xv, s);
    }
  }
  if (D16 == 0 && (d & 8)) {  // the 8-dim step adds into acc0 only (distance.h:64-65)
    const uint32_t s = nb16;
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const float4 xv = t < 2 ? load_step<MODE>(xr[j] + s * STEP + 4 * t, ok[j])
                              : make_float4(0.f, 0.f, 0.f, 0.f);
      if (t < 2) {
#pragma unroll
        for (int b = 0; b < QB; ++b) {
          const float4 qv = *reinterpret_cast<const float4*>(q + b * dp + 16 * s + 4 * t);
          acc[j][b][0] = fma_term<L2>(qv.x, xv.x, acc[j][b][0]);
          acc[j][b][1] = fma_term<L2>(qv.y, xv.y, acc[j][b][1]);
          acc[j][b][2] = fma_term<L2>(qv.z, xv.z, acc[j][b][2]);
          acc[j][b][3] = fma_term<L2>(qv.w, xv.w, acc[j][b][3]);
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < R; ++j) {
#pragma unroll
    for (int b = 0; b < QB; ++b) {
      float v[4], w[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) v[c] = acc[j][b][c] + dpp_f<kDppXor2>(acc[j][b][c]);
#pragma unroll
      for (int c = 0; c < 4; ++c) w[c] = v[c] + dpp_f<kDppXor1>(v[c]);
      float r = (w[0] + w[1]) + (w[2] + w[3]);
      if (D16 == 0 && (d & 7))
        r = scalar_tail<L2, STEP>(q + b * dp, xr[j], int(d & ~7u), int(d), r);
      dist[j][b] = L2 ? r : 1.0f - r;
    }
  }
}

// The same distances with eight lanes per row (d a multiple of 16): each lane holds two of
// the 16 accumulators, so a lane carries half the registers per (query, row) pair of the quad
// layout and twice the queries or rows fit. Lane t (0..7 within its row group) owns dims
// o(t), o(t)+1 of every 16-dim step, o(t) = 2t for t < 4 and 22 - 2t for t >= 4, so the
// lane that DPP row_half_mirror pairs with t (7 - t) owns the accumulators j + 8 of t's
// accumulators j. The tree of hsum256 (distance.h:18-27) then runs as:
//   v = acc + mirror(acc)         (v_j = acc_j + acc_{j+8})
//   w = v + xor2(v)               (w_j = v_j + v_{j+4})
//   p = w[0] + w[1]; r = p + xor1(p)   ((w0 + w1) + (w2 + w3))
// with every add commutative, so all eight lanes of a row hold the same bits. xr[j] points at
// row j; q holds QB queries at stride dp in LDS.
constexpr int kDppHalfMirror = 0x141;

// The distance of one query-row pair from its eight lanes' accumulators (a0: dim o(t),
// a1: dim o(t) + 1, lane t = lane & 7 of the row's group): the accumulator tree of hsum256
// (distance.h:18-27) with DPP exchanges, left in every lane of the group.
template <bool L2>
__device__ __forceinline__ float oct_reduce(float a0, float a1) {
  const float v0 = a0 + dpp_f<kDppHalfMirror>(a0);
  const float v1 = a1 + dpp_f<kDppHalfMirror>(a1);
  const float w0 = v0 + dpp_f<kDppXor2>(v0);
  const float w1 = v1 + dpp_f<kDppXor2>(v1);
  const float p = w0 + w1;
  const float r = p + dpp_f<kDppXor1>(p);
  return L2 ? r : 1.0f - r;
}

template <int QB, int R, int D16, bool L2>
__device__ __forceinline__ void oct_dists(const float* const (&xr)[R], const float* q, uint32_t dp,
                                          int t, float (&dist)[R][QB]) {
  static_assert(D16 > 0, "eight lanes per row need d to be a multiple of 16");
  const int o = t < 4 ? 2 * t : 22 - 2 * t;
  float acc[R][QB][2];
#pragma unroll
  for (int j = 0; j < R; ++j)
#pragma unroll
    for (int b = 0; b < QB; ++b) acc[j][b][0] = acc[j][b][1] = 0.f;
#pragma unroll 4
  for (int s = 0; s < D16; ++s) {
    float2 xv[R];
#pragma unroll
    for (int j = 0; j < R; ++j) xv[j] = *reinterpret_cast<const float2*>(xr[j] + 16 * s + o);
#pragma unroll
    for (int b = 0; b < QB; ++b) {
      const float2 qv = *reinterpret_cast<const float2*>(q + b * dp + 16 * s + o);
#pragma unroll
      for (int j = 0; j < R; ++j) {
        acc[j][b][0] = fma_term<L2>(qv.x, xv[j].x, acc[j][b][0]);
        acc[j][b][1] = fma_term<L2>(qv.y, xv[j].y, acc[j][b][1]);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < R; ++j)
#pragma unroll
    for (int b = 0; b < QB; ++b) dist[j][b] = oct_reduce<L2>(acc[j][b][0], acc[j][b][1]);
}

// LDS-DMA and counted waits, shared by the ring-buffered kernels.
typedef __attribute__((address_space(3))) void* lds_vptr;
typedef __attribute__((address_space(1))) void* glb_vptr;

// 16 bytes per lane from global straight into LDS at lds_base + 16 * lane (global_load_lds_dwordx4);
// lds_base must be the same for the whole wave.
__device__ __forceinline__ void load_lds16(const void* g, void* lds_base) {
  __builtin_amdgcn_global_load_lds((glb_vptr)g, (lds_vptr)lds_base, 16, 0, 0);
}
__device__ __forceinline__ void load_lds4(const void* g, void* lds_base) {
  __builtin_amdgcn_global_load_lds((glb_vptr)g, (lds_vptr)lds_base, 4, 0, 0);
}
// The same, non-temporal (cache policy nt): for data streamed once, such as a pass over the base.
__device__ __forceinline__ void load_lds16_nt(const void* g, void* lds_base) {
  __builtin_amdgcn_global_load_lds((glb_vptr)g, (lds_vptr)lds_base, 16, 0, 2);
}
__device__ __forceinline__ void load_lds4_nt(const void* g, void* lds_base) {
  __builtin_amdgcn_global_load_lds((glb_vptr)g, (lds_vptr)lds_base, 4, 0, 2);
}
// s_waitcnt vmcnt(N), N < 64 (gfx9 encoding: vmcnt bits [3:0] and [15:14]; expcnt and lgkmcnt
// left at their maxima so only vector-memory operations are waited for)
template <int N>
__device__ __forceinline__ void wait_vm() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// Workgroup barrier of the ring-buffered kernels: waits for this wave's LDS reads and writes
// (lgkmcnt(0)) but not for its loads in flight. Each caller first waits, with a counted vmcnt,
// for the ring stages it is about to read. (__syncthreads() would also wait for every LDS-DMA
// load in flight, vmcnt(0), draining the prefetch ring at every stage.) The memory clobber
// keeps the compiler from moving memory accesses across the barrier.
__device__ __forceinline__ void ring_barrier() {
  __builtin_amdgcn_s_waitcnt(0xC07F);  // vmcnt 63 (no wait), expcnt 7, lgkmcnt 0
  asm volatile("s_barrier" ::: "memory");
}

template <int I>
struct Slot {
  static constexpr int value = I;
};

// The kk-th smallest (1-based, 1 <= kk <= c) of the high words of buf[0, c) (the order-preserving
// bits of the estimates), by a block of NT >= 256 threads: radix select, 8 bits a pass; wave 0 finds
// each pass's bin with a prefix sum over the 256 bins (4 a lane). Uses hist (256) and sel (2).
template <int NT = 256>
__device__ __forceinline__ uint32_t block_select_hi32(const uint64_t* buf, uint32_t c, uint32_t kk, uint32_t* hist, uint32_t* sel,
                                      uint32_t tid) {
  static_assert(NT >= 256 && NT % 64 == 0, "a block of at least the 256 histogram bins");
  uint32_t prefix = 0, pmask = 0, rank = kk;
  for (int shift = 24; shift >= 0; shift -= 8) {
    if (NT == 256 || tid < 256) hist[tid] = 0;
    __syncthreads();
    for (uint32_t i = tid; i < c; i += NT) {
      const uint32_t x = uint32_t(buf[i] >> 32);
      if ((x & pmask) == prefix) atomicAdd(&hist[(x >> shift) & 255], 1u);
    }
    __syncthreads();
    if (tid < 64) {
      uint32_t h[4], sum = 0;
#pragma unroll
      for (int j = 0; j < 4; ++j) sum += (h[j] = hist[4 * tid + j]);
      uint32_t incl = sum;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off, 64);
        if (tid >= uint32_t(off)) incl += y;
      }
      uint32_t acc = incl - sum;
      if (acc < rank && rank <= incl) {  // exactly one lane
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          if (acc + h[j] >= rank) {
            sel[0] = 4 * tid + j;
            sel[1] = rank - acc;
            break;
          }
          acc += h[j];
        }
      }
    }
    __syncthreads();
    prefix |= sel[0] << shift;
    pmask |= 255u << shift;
    rank = sel[1];
  }
  return prefix;
}

// Merge of the block's per-wave sorted k-lists into one sorted k-list per query, by ranking:
// a real key's rank in the union is its slot in its own list plus, for every other list, the
// number of keys below it (binary search). Real keys are unique, so ranks are distinct; keys
// ranked >= k are dropped and unused slots keep kPadKey.
// lds: NW x k keys (wave w's list at lds + w * k); out: k keys (LDS or global).
template <int NW>
__device__ __forceinline__ void block_rank_merge(const uint64_t* lds, uint32_t k, uint64_t* out,
                                                 int tid, int nthreads) {
  for (uint32_t i = tid; i < k; i += nthreads) out[i] = kPadKey;
  __syncthreads();
  for (uint32_t i = tid; i < uint32_t(NW) * k; i += nthreads) {
    const uint32_t w = i / k, s = i - w * k;
    const uint64_t v = lds[i];
    if (v == kPadKey) continue;
    uint32_t rank = s;
#pragma unroll
    for (int o = 0; o < NW; ++o) {
      if (o == int(w)) continue;
      const uint64_t* l = lds + o * k;
      uint32_t lo = 0, hi = k;  // first index with l[idx] >= v
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (l[mid] < v) lo = mid + 1;
        else hi = mid;
      }
      rank += lo;
      if (rank >= k) break;
    }
    if (rank < k) out[rank] = v;
  }
}

__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
  const uint32_t lo = uint32_t(__shfl_xor(int(uint32_t(v)), m, 64));
  const uint32_t hi = uint32_t(__shfl_xor(int(uint32_t(v >> 32)), m, 64));
  return (uint64_t(hi) << 32) | lo;
}

// value of lane `src` (0..63, per lane) of v: ds_bpermute, no LDS storage
__device__ __forceinline__ uint64_t bperm64(uint64_t v, int src) {
  const uint32_t lo = uint32_t(__builtin_amdgcn_ds_bpermute(src << 2, int(uint32_t(v))));
  const uint32_t hi = uint32_t(__builtin_amdgcn_ds_bpermute(src << 2, int(uint32_t(v >> 32))));
  return (uint64_t(hi) << 32) | lo;
}

// Sorts the wave's 64 keys (one per lane) ascending across the lanes: a bitonic network of 21
// compare-exchange stages.
__device__ __forceinline__ uint64_t wave_sort64(uint64_t v, int lane) {
#pragma unroll
  for (int k = 2; k <= 64; k <<= 1)
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const uint64_t o = shfl_xor64(v, j);
      const bool keep_min = ((lane & j) == 0) == ((lane & k) == 0);
      v = keep_min ? (o < v ? o : v) : (o < v ? v : o);
    }
  return v;
}

// Slot q (0 <= q < KP) of a wave list: register q >> 6, lane q & 63, read by every lane.
template <int KP>
__device__ __forceinline__ uint64_t list_at(const uint64_t (&e)[KP / 64], int q) {
  uint64_t v = bperm64(e[0], q & 63);
#pragma unroll
  for (int r = 1; r < KP / 64; ++r) {
    const uint64_t w = bperm64(e[r], q & 63);
    v = (q >> 6) == r ? w : v;
  }
  return v;
}

// Merges 64 sorted keys (b: key of rank `lane`, kPadKey for none) into a sorted wave list and
// keeps its KP smallest; returns the new k-th key. Merge path: output slot p takes the p-th
// smallest of the union; i = the number of b keys among the outputs before p is the largest i
// in [0, min(p, 64)] with i == 0 or b[i - 1] < list[p - i] (the predicate is monotone; list keys
// come first among equals, and only pad keys can be equal). Every lane runs the same seven
// search steps (a range of at most 65 values), because a lane permute reads zero from a lane
// that is not active: nothing here may run under a partial mask.
template <int KP>
__device__ __forceinline__ uint64_t merge_sorted64(WaveList<KP>& lst, uint64_t b, int lane, int k) {
  uint64_t out[KP / 64];
#pragma unroll
  for (int r = 0; r < KP / 64; ++r) {
    const int p = r * 64 + lane;
    int lo = 0, hi = p < 64 ? p : 64;
#pragma unroll
    for (int it = 0; it < 7; ++it) {
      const int mid = (lo + hi + 1) >> 1;  // == lo once lo == hi
      const uint64_t bv = bperm64(b, mid > 0 ? mid - 1 : 0);
      const uint64_t av = list_at<KP>(lst.e, p - mid);  // 0 <= p - mid < KP
      const bool pred = bv < av;
      if (lo < hi) {
        if (pred) lo = mid;
        else hi = mid - 1;
      }
    }
    const uint64_t bv = bperm64(b, lo < 64 ? lo : 63);
    const uint64_t av = list_at<KP>(lst.e, p - lo);
    out[r] = (lo < 64 && bv < av) ? bv : av;
  }
#pragma unroll
  for (int r = 0; r < KP / 64; ++r) lst.e[r] = out[r];
  const int kk = k - 1;
  uint64_t thr = kPadKey;
#pragma unroll
  for (int r = 0; r < KP / 64; ++r)
    if ((kk >> 6) == r) thr = readlane64(lst.e[r], kk & 63);
  return thr;
}

// Offer keys of the calling wave to a query's k-list; `key` is kPadKey on lanes that carry no
// candidate. Returns the updated admission threshold (the list's k-th key).
template <int KP>
__device__ __forceinline__ uint64_t offer(WaveList<KP>& lst, uint64_t thr, uint64_t key, int lane,
                                          int k) {
  uint64_t m = ballot(key < thr);
  while (m) {
    const int src = __builtin_ctzll(m);
    m &= m - 1;
    const uint64_t ck = readlane64(key, src);
    if (ck < thr) {
      // the list's k-th key, or the (tighter) threshold the scan started from
      const uint64_t t = lst.insert(ck, lane, k);
      thr = t < thr ? t : thr;
    }
  }
  return thr;
}

}  // namespace vsg

namespace vsg {

// Offer the wave's 64 keys (one per lane, kPadKey on lanes without one) to a query's k-list:
// a few admitted keys are inserted one at a time, more at once by a sort and a merge (about
// the cost of 12 single inserts for 128 slots, 8 for 64). Either way the list ends up holding
// its KP smallest keys among the old list and the keys below thr, so its first k slots are the
// same. Returns the updated admission threshold.
template <int KP>
__device__ __forceinline__ uint64_t offer64(WaveList<KP>& lst, uint64_t thr, uint64_t key, int lane, int k) {
  const uint64_t m = ballot(key < thr);
  if (__popcll(m) <= (KP == 128 ? 12 : 8)) return offer<KP>(lst, thr, key, lane, k);
  const uint64_t sorted = wave_sort64(key < thr ? key : kPadKey, lane);
  const uint64_t t = merge_sorted64<KP>(lst, sorted, lane, k);
  return t < thr ? t : thr;
}

}  // namespace vsg
