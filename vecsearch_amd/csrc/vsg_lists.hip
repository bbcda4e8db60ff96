// Filtered exact scan over allowed-id lists, for a batch of queries with different filters
// and different k, in one persistent launch.
//
// Reference behaviour: HNSWIndex::search with SearchFilter (hnsw.h:25-33) takes the exact path
// when n_allowed <= exact_below (hnsw.cpp:322-334) and scans the allowed rows with distance()
// into a bounded heap (hnsw.cpp:336-354); search_batch pads missing results with kNone / +inf
// (hnsw.cpp:356-366). Here every query carries its own allowed set (a CSR range of row ids) and
// its own k; distances are the reference's bits (quad_dists) and ties go to the lower id.
//
// Work decomposition: queries whose CSR ranges are identical form a group and share every row
// read. A work item is (group, chunk of <= C allowed rows, block of <= QB of the group's
// queries). Waves pull items from a global counter until none is left, keep a k-list per
// query, and write it to the query's slot for that chunk; the merge kernel then reduces the
// chunk lists of each query.
//
// Chunks of one query share an admission threshold: qthr[q] (a key, kPadKey at the start) is
// lowered with atomicMin to the k-th key of every complete chunk list, and a chunk's list
// starts from it instead of from empty. A key left out is then above k keys of some other
// chunk, so the merged top-k is unchanged, and later chunks skip most of the list inserts
// that filling k slots from empty costs (k (1 + ln(C / k)) per chunk for k = 100).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <utility>
#include <vector>

#include "vsg.h"
#include "vsg_internal.h"
#include "vsg_kernels.h"

namespace vsg {

struct ListItem {
  uint32_t lb, le;  // range of allow_ids (direct: range of rows)
  uint32_t chunk;   // chunk index within the group
  uint32_t qoff;    // first entry in qperm
  uint32_t nq;      // queries in this block
  uint32_t direct;  // 1: unfiltered group, position p is row p
};

// LPR: lanes per row, 4 (quad_dists, any d) or 8 (oct_dists, d a multiple of 16).
template <int QB, int R, int D16, int KP, bool L2, int LPR>
__global__ __launch_bounds__(kBlock) void lists_scan_kernel(
    const float* __restrict__ base, uint32_t dp, uint32_t d, const float* __restrict__ queries,
    const uint32_t* __restrict__ ks, const ListItem* __restrict__ items, uint32_t n_items,
    const uint32_t* __restrict__ qperm, const uint32_t* __restrict__ allow_ids,
    const uint64_t* __restrict__ slot_off, uint32_t L, uint64_t* __restrict__ part,
    uint32_t* __restrict__ counter, uint64_t* __restrict__ qthr) {
  extern __shared__ float lds[];
  constexpr int kRowsPerStep = 64 / LPR;  // rows per wave instruction
  const int tid = threadIdx.x, lane = tid & 63, t = lane & (LPR - 1), g = lane / LPR;
  float* qw = lds + (tid >> 6) * QB * dp;  // this wave's queries
  // a launch with no work left (a captured fallback with nothing to do) ends on a plain load
  // instead of every wave's same-address atomic
  if (__hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= n_items) return;

  for (;;) {
    uint32_t it = 0;
    if (lane == 0) it = atomicAdd(counter, 1u);
    it = __builtin_amdgcn_readfirstlane(it);
    if (it >= n_items) break;
    const ListItem w = items[it];

    // previous item's LDS reads are done before the queries are replaced
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    for (uint32_t i = lane; i < uint32_t(QB) * dp; i += 64) {
      const uint32_t b = i / dp, c = i - b * dp;
      qw[i] = (b < w.nq && c < d) ? queries[size_t(qperm[w.qoff + b]) * d + c] : 0.f;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");

    WaveList<KP> lst[QB];
    uint64_t thr[QB];
    int kq[QB];
#pragma unroll
    for (int b = 0; b < QB; ++b) {
      lst[b].init();
      const uint32_t q = uint32_t(b) < w.nq ? qperm[w.qoff + b] : 0u;
      kq[b] = uint32_t(b) < w.nq ? int(min(ks[q], uint32_t(KP))) : 0;
      // k == 0 admits nothing; else the tightest bound other chunks of the query have published
      thr[b] = kq[b] > 0 ? __hip_atomic_load(qthr + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
    }

    for (uint32_t r0 = w.lb; r0 < w.le; r0 += kRowsPerStep * R) {
      uint32_t row[R];
      bool ok[R];
      const float* xr[R];
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const uint32_t idx = r0 + j * kRowsPerStep + g;
        ok[j] = idx < w.le;
        row[j] = w.direct ? (ok[j] ? idx : w.lb) : allow_ids[ok[j] ? idx : w.lb];
        xr[j] = base + size_t(row[j]) * dp;
      }
      float dist[R][QB];
      if constexpr (LPR == 8)
        oct_dists<QB, R, D16, L2>(xr, qw, dp, t, dist);
      else
        quad_dists<QB, R, D16, L2, 16, kLoadPlain>(xr, ok, qw, dp, d, t, dist);
#pragma unroll
      for (int j = 0; j < R; ++j)
#pragma unroll
        for (int b = 0; b < QB; ++b) {
          if (kq[b] == 0) continue;
          const uint64_t key = (ok[j] && t == 0) ? make_key(dist[j][b], row[j]) : kPadKey;
          thr[b] = offer<KP>(lst[b], thr[b], key, lane, kq[b]);
        }
    }

#pragma unroll
    for (int b = 0; b < QB; ++b) {
      if (uint32_t(b) >= w.nq) continue;
      const uint32_t q = qperm[w.qoff + b];
      if (lane == 0 && thr[b] != kPadKey && kq[b] > 0) atomicMin(reinterpret_cast<unsigned long long*>(qthr + q), thr[b]);
      uint64_t* dst = part + (slot_off[q] + w.chunk) * L;
#pragma unroll
      for (int r = 0; r < KP / 64; ++r) {
        const uint32_t s = r * 64 + lane;
        if (s < L) dst[s] = lst[b].e[r];
      }
    }
  }
}

// ---- one lane per row ------------------------------------------------------------------------
// A workgroup of 8 waves takes one work item at a time (a chunk of a group's rows and up to 8 of
// its queries) and runs it in rounds of 512 rows, wave w computing rows 64w .. 64w + 63 of the
// round. Each lane owns one row and computes its distances to the item's queries itself: the
// reference's 16 accumulators per pair (accumulator j sums dims j, j + 16, ...) and the hsum256
// tree (distance.h:18-27, 54-69), all in its registers. The queries are the same for every lane
// and every wave, so they come through scalar loads (their 12 KB stay in the scalar cache for
// the item) and enter the FMAs as scalar operands: no LDS traffic for them and no cross-lane
// reduction. They are read in item order from the copy lists_gather_queries_kernel makes.
//
// Rows reach the lanes through LDS: a wave loads each 16-dim step of its 64 rows with four
// LDS-DMA loads, four lanes per row (16 cache lines per instruction instead of 64), two steps
// ahead in a three-slot ring, and each lane then reads its own row's 64 bytes back. A row's four
// 16-byte chunks are stored XOR-swizzled by (row >> 2) & 3, so the 16 lanes of a read touch 16
// different bank groups.
//
// The round's keys (512 rows x 8 queries) go to LDS; then wave b alone offers query b's keys to
// the query's k-list, which lives in its registers for the whole item (offer64: one sorted merge
// for many admitted keys). One list per query sees every row of the chunk, so the inserts are
// those of one sequential scan. At the end of the item wave b writes the list as the query's
// chunk list and lowers the query's device-wide threshold, as the kernels above do; a workgroup
// runs its items one after another, so its later chunks start from it.
//
// Items come from eight queues, one per XCD. Workgroups are dealt to the XCDs in turn, so
// workgroup b starts on queue b % 8, where the planner put every query block of a chunk; a
// chunk's re-reads then find its rows in that XCD's L2. A drained queue sends the workgroup on
// to the next one.
constexpr int kLaneQ = 8;        // queries per item
constexpr int kLaneWaves = 8;    // waves per workgroup
constexpr uint32_t kQueues = 8;  // one per XCD
constexpr uint32_t kCounterWords = 32;  // [0] the single queue, [1..8] the XCD queues,
constexpr uint32_t kBoundsAt = 16;      // [16..24] the XCD queues' item ranges
constexpr uint32_t kLaneRound = 64 * kLaneWaves;                    // rows per round
constexpr size_t kLaneKeyBytes = size_t(kLaneQ) * kLaneRound * 8;  // 32 KB
constexpr size_t kLaneStage = 4096;                                // one step of 64 rows
constexpr int kLaneSlots = 3;                                     // ring slots per wave
// two key buffers (rounds alternate; the item index is passed in the second, which no round
// writes before every wave has read it) and the waves' rings: 160 KB, a CU's whole LDS
constexpr size_t kLaneLds = 2 * kLaneKeyBytes + size_t(kLaneWaves) * kLaneSlots * kLaneStage;
static_assert(kLaneLds <= 160 * 1024, "one workgroup a CU");

// Byte offset of 16-byte chunk j of row r (0..63) in a staged step.
__device__ __forceinline__ uint32_t lane_slot(uint32_t r, uint32_t j) {
  return (r >> 4) * 1024 + (r & 15) * 64 + ((j ^ ((r >> 2) & 3)) << 4);
}

// One 16-dim step of NQ queries against the lane's row. The query values come in pairs of
// queries (32 scalar registers): pair p + 1 is requested right after the first FMA of pair p,
// which is where the wait for pair p falls, so the remaining 31 FMAs of pair p cover the next
// pair's load; with an even number of pairs the last pair requests the next step's first pair
// (qs_next) the same way, so no step starts on a load in flight (1.60 -> 1.45 ms for the
// 4,096-query filtered batch). (Scalar loads complete out of order: any wait is a wait for all
// of them.)
template <int NQ, bool L2>
__device__ __forceinline__ void lane_fma_step(const float4 (&xv)[4], const float* __restrict__ qs,
                                              const float* __restrict__ qs_next, size_t d, float (&acc)[NQ][16],
                                              float (&qv)[2][32]) {
  static_assert(NQ % 2 == 0, "queries in pairs");
  // an even number of pairs: qv[0] holds this step's first pair on entry (loaded during the
  // previous step), and the last pair loads the next step's first pair into qv[0]
  constexpr bool kCarry = (NQ / 2) % 2 == 0;
  if constexpr (!kCarry) {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      qv[0][i] = qs[i];
      qv[0][16 + i] = qs[size_t(d) + i];
    }
  }
#pragma unroll
  for (int p = 0; p < NQ / 2; ++p) {
    const float* q = qv[p & 1];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int b = 2 * p + h;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const float x4[4] = {xv[c].x, xv[c].y, xv[c].z, xv[c].w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int j = 4 * c + e;
          acc[b][j] = fma_term<L2>(q[16 * h + j], x4[e], acc[b][j]);
          if (h == 0 && j == 0 && p + 1 < NQ / 2) {
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int i = 0; i < 16; ++i) {
              qv[(p + 1) & 1][i] = qs[size_t(2 * p + 2) * d + i];
              qv[(p + 1) & 1][16 + i] = qs[size_t(2 * p + 3) * d + i];
            }
            __builtin_amdgcn_sched_barrier(0);
          }
          if (kCarry && h == 0 && j == 0 && p + 1 == NQ / 2) {
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int i = 0; i < 16; ++i) {
              qv[0][i] = qs_next[i];
              qv[0][16 + i] = qs_next[size_t(d) + i];
            }
            __builtin_amdgcn_sched_barrier(0);
          }
        }
      }
    }
  }
}

// Keys of the wave's 64 rows x NQ queries into key_s[b * kLaneRound + 64 * wave + lane].
// src[c]: this lane's global address for DMA load c of step 0 (row 16c + lane / 4 of the
// wave's rows, the chunk that lands at this lane's slot); stage: the wave's three 4 KB slots.
template <int NQ, int D16, bool L2>
__device__ __forceinline__ void lane_rows(const float* const (&src)[4], uint8_t* stage, const float* __restrict__ qb,
                                          uint32_t d, uint32_t nd16, bool ok, uint32_t row, uint64_t* key_s,
                                          int wave, int lane) {
  float acc[NQ][16];
#pragma unroll
  for (int b = 0; b < NQ; ++b)
#pragma unroll
    for (int j = 0; j < 16; ++j) acc[b][j] = 0.f;
  const uint32_t last = nd16 - 1;
  auto issue = [&](uint32_t slot, uint32_t s) __attribute__((always_inline)) {
#pragma unroll
    for (int c = 0; c < 4; ++c) load_lds16(src[c] + 16 * s, stage + slot * kLaneStage + c * 1024);
  };
  issue(0, 0);
  issue(1, last >= 1 ? 1 : last);
  uint32_t slot = 0;  // slot of step s; step s + 2 goes to the slot read at step s - 1
  float qv[2][32];  // query values in scalar registers (lane_fma_step)
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    qv[0][i] = qb[i];
    qv[0][16 + i] = qb[size_t(d) + i];
  }
  for (uint32_t s = 0; s < nd16; ++s) {
    issue(slot == 0 ? 2 : slot - 1, s + 2 <= last ? s + 2 : last);  // past the end: unread slots
    wait_vm<8>();                                                   // step s has landed
    const uint8_t* cur = stage + slot * kLaneStage;
    slot = slot == 2 ? 0 : slot + 1;
    float4 xv[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) xv[j] = *reinterpret_cast<const float4*>(cur + lane_slot(uint32_t(lane), uint32_t(j)));
    // the query row stride as a constant where the width is one (the loads' offsets are then
    // immediates rather than scalar-register arithmetic)
    lane_fma_step<NQ, L2>(xv, qb + 16 * s, qb + 16 * (s < last ? s + 1 : s), D16 > 0 ? size_t(D16) * 16 : size_t(d),
                          acc, qv);
  }
  wait_vm<0>();  // the extra loads past the last step land before the slots are reused
  // the tree of hsum256 on (acc0 + acc1): v_j = a_j + a_{j+8}, w_c = v_c + v_{c+4},
  // r = (w0 + w1) + (w2 + w3)
#pragma unroll
  for (int b = 0; b < NQ; ++b) {
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = acc[b][j] + acc[b][j + 8];
    const float w0 = v[0] + v[4], w1 = v[1] + v[5], w2 = v[2] + v[6], w3 = v[3] + v[7];
    const float r = (w0 + w1) + (w2 + w3);
    key_s[b * kLaneRound + 64 * wave + lane] = ok ? make_key(L2 ? r : 1.0f - r, row) : kPadKey;
  }
}

template <int D16, int KP, bool L2>
__global__ __launch_bounds__(64 * kLaneWaves, 1) void lists_lane_kernel(
    const float* __restrict__ base, uint32_t dp, uint32_t d, const float* __restrict__ qsorted,
    const uint32_t* __restrict__ ks, const ListItem* __restrict__ items, const uint32_t* __restrict__ qperm,
    const uint32_t* __restrict__ allow_ids, const uint64_t* __restrict__ slot_off, uint32_t L,
    uint64_t* __restrict__ part, uint32_t* __restrict__ counter, uint64_t* __restrict__ qthr) {
  extern __shared__ __attribute__((aligned(16))) uint8_t lds_raw[];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  uint64_t* keys2 = reinterpret_cast<uint64_t*>(lds_raw);  // 2 x kLaneQ x kLaneRound
  uint8_t* stage = lds_raw + 2 * kLaneKeyBytes + size_t(wave) * kLaneSlots * kLaneStage;  // this wave's ring
  // the item index: in the second key buffer, first written by the item's second round, after
  // the first round's barrier (every wave has read the index by then)
  uint32_t* item_s = reinterpret_cast<uint32_t*>(lds_raw + kLaneKeyBytes);
  const uint32_t nd16 = D16 > 0 ? uint32_t(D16) : d >> 4;
  const uint32_t* bounds = counter + kBoundsAt;

  uint32_t qu = blockIdx.x % kQueues, tries = 0;
  for (;;) {
    __syncthreads();  // the previous item's LDS is no longer read
    if (tid == 0) {
      uint32_t it = ~0u;
      while (tries < kQueues) {
        const uint32_t i = atomicAdd(counter + 1 + qu, 1u);
        if (bounds[qu] + i < bounds[qu + 1]) {
          it = bounds[qu] + i;
          break;
        }
        qu = (qu + 1) % kQueues;  // this queue is drained: the next one
        ++tries;
      }
      *item_s = it;
    }
    __syncthreads();
    const uint32_t it = __builtin_amdgcn_readfirstlane(*item_s);  // the same in every lane
    if (it == ~0u) break;
    const ListItem w = items[it];
    const uint32_t nq = w.nq;
    const float* qb = qsorted + size_t(w.qoff) * d;  // the item's queries, consecutive

    // wave b (< nq) owns query b's k-list and threshold for the whole item
    const bool owner = uint32_t(wave) < nq;
    const uint32_t myq = owner ? qperm[w.qoff + wave] : 0u;
    const int kq = owner ? int(min(ks[myq], uint32_t(KP))) : 0;
    // k == 0 admits nothing; else the tightest bound the query's other chunks published
    uint64_t thr = kq > 0 ? __hip_atomic_load(qthr + myq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
    WaveList<KP> lst;
    lst.init();

    uint32_t par = 0;  // key buffer of the round
    for (uint32_t r0 = w.lb; r0 < w.le; r0 += kLaneRound, par ^= 1u) {
      // Round r writes buffer r % 2 and its owners then offer from it. Round r + 1 writes the
      // other buffer, last read by round r - 1's offers, which every wave finished before round
      // r's barrier: one barrier a round, and a wave with few keys to offer goes on to the next
      // round's rows while the others still offer.
      uint64_t* key_s = keys2 + size_t(par) * (kLaneKeyBytes / 8);
      const uint32_t my0 = r0 + 64u * uint32_t(wave);
      const uint32_t idx = my0 + uint32_t(lane);
      const bool ok = idx < w.le;
      if (my0 < w.le) {
        const uint32_t row = w.direct ? (ok ? idx : w.lb) : allow_ids[ok ? idx : w.lb];
        const float* src[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const uint32_t rr = 16u * c + (uint32_t(lane) >> 2);  // row of the wave that DMA load c fills here
          const uint32_t jj = (uint32_t(lane) & 3) ^ ((uint32_t(lane) >> 4) & 3);  // its chunk at this slot
          const uint32_t rid = uint32_t(__shfl(int(row), int(rr), 64));
          src[c] = base + size_t(rid) * dp + 4 * jj;
        }
        // the item's queries, rounded up to a pair
        if (nq <= 2) {
          lane_rows<2, D16, L2>(src, stage, qb, d, nd16, ok, row, key_s, wave, lane);
        } else if (kLaneQ <= 4 || nq <= 4) {
          lane_rows<4, D16, L2>(src, stage, qb, d, nd16, ok, row, key_s, wave, lane);
        } else if constexpr (kLaneQ > 4) {
          if (nq <= 6) lane_rows<6, D16, L2>(src, stage, qb, d, nd16, ok, row, key_s, wave, lane);
          else lane_rows<8, D16, L2>(src, stage, qb, d, nd16, ok, row, key_s, wave, lane);
        }
      }
      __syncthreads();  // the round's keys are in LDS
      if (owner && kq > 0) {
        // a tighter bound that the query's other chunks published since (a relaxed load of a
        // value that only decreases: any value read is a valid bound)
        const uint64_t pub = __hip_atomic_load(qthr + myq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t thr0 = pub < thr ? pub : thr;
        thr = thr0;
        const uint32_t nw = min(uint32_t(kLaneWaves), (w.le - r0 + 63) / 64);  // waves with rows
#pragma unroll 1
        for (uint32_t c = 0; c < nw; ++c) {
          const uint64_t key = key_s[uint32_t(wave) * kLaneRound + 64 * c + uint32_t(lane)];
          if (ballot(key < thr) == 0) continue;
          thr = offer64<KP>(lst, thr, key, lane, kq);
        }
        // publish this chunk's bound as soon as the round tightened it, not only at the item's
        // end, so that the query's chunks running at the same time admit fewer keys
        if (lane == 0 && thr < thr0) atomicMin(reinterpret_cast<unsigned long long*>(qthr + myq), thr);
      }
    }

    if (owner) {
      if (lane == 0 && thr != kPadKey && kq > 0) atomicMin(reinterpret_cast<unsigned long long*>(qthr + myq), thr);
      uint64_t* dst = part + (slot_off[myq] + w.chunk) * L;
#pragma unroll
      for (int r = 0; r < KP / 64; ++r) {
        const uint32_t sl = r * 64 + uint32_t(lane);
        if (sl < L) dst[sl] = lst.e[r];
      }
    }
  }
}

// The batch's queries in item order (qperm), so an item's queries are consecutive rows.
__global__ void lists_gather_queries_kernel(const float* __restrict__ q, const uint32_t* __restrict__ qperm,
                                            uint32_t nq, uint32_t d, float* __restrict__ out) {
  const uint32_t i = blockIdx.x;
  if (i >= nq) return;
  const float* src = q + size_t(qperm[i]) * d;
  float* dst = out + size_t(i) * d;
  for (uint32_t c = threadIdx.x; c < d; c += blockDim.x) dst[c] = src[c];
}

// ---- block-shared rows: 4 waves x 8 queries read every row once ---------------------------
// The kernel above gives each wave its own item, so a group's rows are read from memory once
// per QB queries. Here a work item belongs to a whole block: (group, chunk of rows, up to 32 of
// the group's queries), and every row read serves all 32.
//
// Rows and query dims stream through a 3-slot LDS ring filled by LDS-DMA (global_load_lds)
// two stages ahead. A stage is a window of 4 16-dim steps: the 64 rows of a row block, laid
// out [step][row][16 floats] (16 KB), and the item's 32 queries, [step][query][16 floats]
// (8 KB); wave w loads step w of the window for both. Wave w computes its queries 8w .. 8w+7
// against all 64 rows in the eight-lanes-per-row order (lane t of row group g owns dims
// 16s + o(t), +1 of row j * 8 + g; oct_reduce), accumulating over the windows of the row block;
// after its last window the pair sums are reduced and offered to the wave's k-lists.
//
// Row ids travel in registers: lane L holds the id of row L of a row block. Every stage's
// fetch first issues one plain load of the next row block's ids, then its 6 DMA loads; the
// counted wait for a stage allows only the next stage's 6 DMA loads in flight, so it also
// covers that stage's id load, issued before them (and stays exact where the compiler drops
// the id load of an unfiltered item). The fetch that starts a row block moves the loaded ids
// into `cur` (the previous ones into `prev`, still needed by the offers of the row block being
// finished, which trail the fetches by two stages). Lanes get other rows' ids with __shfl.
constexpr int kBlkWaves = 4, kBlkQ = 8, kBlkR = 8, kBlkRows = 8 * kBlkR, kBlkWin = 4;
constexpr int kBlkItemQ = kBlkWaves * kBlkQ;                 // queries per item
constexpr int kBlkRowBytes = kBlkWin * kBlkRows * 64;        // 16 KB
constexpr int kBlkQBytes = kBlkWin * kBlkItemQ * 64;         // 8 KB
constexpr int kBlkStage = kBlkRowBytes + kBlkQBytes;
constexpr int kBlkLoads = kBlkStage / 1024 / kBlkWaves;      // DMA loads per wave per stage
static_assert(kBlkRowBytes / 1024 == 4 * kBlkWaves && kBlkQBytes / 1024 == 2 * kBlkWaves, "stage split");

template <int D16, int KP, bool L2>
__global__ __launch_bounds__(64 * kBlkWaves, 2) void lists_block_kernel(
    const float* __restrict__ base, uint32_t dp, const float* __restrict__ queries,
    const uint32_t* __restrict__ ks, const ListItem* __restrict__ items, uint32_t n_items,
    const uint32_t* __restrict__ qperm, const uint32_t* __restrict__ allow_ids,
    const uint64_t* __restrict__ slot_off, uint32_t L, uint64_t* __restrict__ part,
    uint32_t* __restrict__ counter, uint64_t* __restrict__ qthr) {
  static_assert(D16 % kBlkWin == 0 && D16 / kBlkWin >= 2, "whole windows, at least two per row block");
  constexpr uint32_t NWIN = D16 / kBlkWin;
  constexpr uint32_t D = D16 * 16;  // dims; queries have stride D, rows stride dp
  __shared__ __attribute__((aligned(16))) uint8_t slot0[kBlkStage];
  __shared__ __attribute__((aligned(16))) uint8_t slot1[kBlkStage];
  __shared__ __attribute__((aligned(16))) uint8_t slot2[kBlkStage];
  __shared__ uint32_t item_s;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, t = lane & 7, g = lane >> 3;
  const int o = t < 4 ? 2 * t : 22 - 2 * t;
  if (__hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= n_items) return;  // no work

  for (;;) {
    __syncthreads();  // the previous item's LDS is no longer read
    if (tid == 0) item_s = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t it = item_s;
    if (it >= n_items) break;
    const ListItem w = items[it];
    const uint32_t nrows = w.le - w.lb, nrb = (nrows + kBlkRows - 1) / kBlkRows, nst = nrb * NWIN;
    if (nrows == 0) continue;  // (the planner makes no empty items)
    const uint32_t qw0 = uint32_t(wave) * kBlkQ;
    const bool active = qw0 < w.nq;  // wave-uniform: a wave without queries only loads

    // query rows of this lane's DMA pieces (query 16 m + lane / 4 of the item, m = 0, 1)
    const float* qsrc[2];
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      const uint32_t qi = 16 * m + (lane >> 2);
      qsrc[m] = queries + size_t(qperm[w.qoff + (qi < w.nq ? qi : 0)]) * D + (lane & 3) * 4;
    }
    WaveList<KP> lst[kBlkQ];
    uint64_t thr[kBlkQ];
    int kq[kBlkQ];
#pragma unroll
    for (int b = 0; b < kBlkQ; ++b) {
      lst[b].init();
      const uint32_t qi = qw0 + b;
      const uint32_t q = qi < w.nq ? qperm[w.qoff + qi] : 0u;
      kq[b] = qi < w.nq ? int(min(ks[q], uint32_t(KP))) : 0;
      thr[b] = kq[b] > 0 ? __hip_atomic_load(qthr + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
    }
    // ids of row block rb, lane L: row L (clamped into the chunk; a direct group's rows are lb + i)
    auto id_load = [&](uint32_t rb) -> uint32_t {
      const uint32_t i = rb * kBlkRows + uint32_t(lane), ic = i < nrows ? i : 0u;
      return w.direct ? allow_ids[0] : allow_ids[w.lb + ic];
    };
    auto id_of = [&](uint32_t v, uint32_t rb, uint32_t r) -> uint32_t {  // row r of row block rb
      const uint32_t i = rb * kBlkRows + r;
      return w.direct ? w.lb + (i < nrows ? i : 0u) : uint32_t(__shfl(int(v), int(r), 64));
    };
    uint32_t id_next = id_load(0), id_cur = 0, id_prev = 0;
    wait_vm<0>();  // qperm / ks / the first ids (and the previous item's stores)

    auto fetch = [&](uint8_t* dst, uint32_t st) __attribute__((always_inline)) {
      const uint32_t rb = st / NWIN, win = st - rb * NWIN;
      if (win == 0) {
        id_prev = id_cur;
        id_cur = id_next;
      }
      uint32_t row[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) row[m] = id_of(id_cur, rb, 16 * m + (lane >> 2));
      id_next = id_load(rb + 1 < nrb ? rb + 1 : rb);
      const uint32_t s = win * kBlkWin + uint32_t(wave);  // the step this wave loads
#pragma unroll
      for (int m = 0; m < 4; ++m)
        load_lds16(base + size_t(row[m]) * dp + s * 16 + (lane & 3) * 4, dst + wave * 4096 + m * 1024);
#pragma unroll
      for (int m = 0; m < 2; ++m) load_lds16(qsrc[m] + s * 16, dst + kBlkRowBytes + (2 * wave + m) * 1024);
    };

    float acc[kBlkR][kBlkQ][2];
    auto compute = [&](const uint8_t* src, uint32_t st) __attribute__((always_inline)) {
      const uint32_t rb = st / NWIN, win = st - rb * NWIN;
      if (win == 0) {
#pragma unroll
        for (int j = 0; j < kBlkR; ++j)
#pragma unroll
          for (int b = 0; b < kBlkQ; ++b) acc[j][b][0] = acc[j][b][1] = 0.f;
      }
      const float* xs = reinterpret_cast<const float*>(src);
      const float* qs = reinterpret_cast<const float*>(src + kBlkRowBytes) + qw0 * 16;
#pragma unroll
      for (int sw = 0; sw < kBlkWin; ++sw) {
        float2 xv[kBlkR];
#pragma unroll
        for (int j = 0; j < kBlkR; ++j)
          xv[j] = *reinterpret_cast<const float2*>(xs + (sw * kBlkRows + j * 8 + g) * 16 + o);
#pragma unroll
        for (int b = 0; b < kBlkQ; ++b) {
          const float2 qv = *reinterpret_cast<const float2*>(qs + (sw * kBlkItemQ + b) * 16 + o);
#pragma unroll
          for (int j = 0; j < kBlkR; ++j) {
            acc[j][b][0] = fma_term<L2>(qv.x, xv[j].x, acc[j][b][0]);
            acc[j][b][1] = fma_term<L2>(qv.y, xv[j].y, acc[j][b][1]);
          }
        }
      }
      if (win + 1 == NWIN) {  // the row block is complete: distances to the k-lists
#pragma unroll
        for (int j = 0; j < kBlkR; ++j) {
          const uint32_t r = uint32_t(j * 8 + g), i = rb * kBlkRows + r;
          const bool ok = i < nrows && t == 0;
          // the fetches moved on to row block rb + 1 two stages ago, if there is one
          const uint32_t row = id_of(rb + 1 < nrb ? id_prev : id_cur, rb, r);
#pragma unroll
          for (int b = 0; b < kBlkQ; ++b) {
            const float dist = oct_reduce<L2>(acc[j][b][0], acc[j][b][1]);
            if (kq[b] == 0) continue;
            thr[b] = offer<KP>(lst[b], thr[b], ok ? make_key(dist, row) : kPadKey, lane, kq[b]);
          }
        }
      }
    };
    auto slot = [&](auto I) -> uint8_t* {
      constexpr int v = decltype(I)::value % 3;
      return v == 0 ? slot0 : v == 1 ? slot1 : slot2;
    };
    // Consume stage st from slot I; refill slot I + 2 (consumed at st - 1) with stage st + 2.
    // Every step issues the same loads (past the last stage it fetches the last stage again,
    // into a slot nobody reads; the next item's first barrier drains them), so the counted
    // wait is the same everywhere and the compiler's own load tracking stays exact.
    auto step = [&](auto I, uint32_t st) __attribute__((always_inline)) {
      wait_vm<kBlkLoads>();  // stage st + 1's DMA loads may stay in flight
      ring_barrier();
      fetch(slot(Slot<decltype(I)::value + 2>()), min(st + 2, nst - 1));
      if (active) compute(slot(I), st);
    };
    fetch(slot0, 0);  // nst >= NWIN >= 2: an item has at least one row
    fetch(slot1, 1);
    for (uint32_t st = 0;;) {  // leaves the loop right after the last stage
      step(Slot<0>(), st);
      if (++st == nst) break;
      step(Slot<1>(), st);
      if (++st == nst) break;
      step(Slot<2>(), st);
      if (++st == nst) break;
    }

    wait_vm<0>();  // the refetches past the last stage land before the slots are reused
    if (active) {
#pragma unroll
      for (int b = 0; b < kBlkQ; ++b) {
        const uint32_t qi = qw0 + b;
        if (qi >= w.nq) continue;
        const uint32_t q = qperm[w.qoff + qi];
        if (lane == 0 && thr[b] != kPadKey && kq[b] > 0) atomicMin(reinterpret_cast<unsigned long long*>(qthr + q), thr[b]);
        uint64_t* dst = part + (slot_off[q] + w.chunk) * L;
#pragma unroll
        for (int r = 0; r < KP / 64; ++r) {
          const uint32_t sl = r * 64 + lane;
          if (sl < L) dst[sl] = lst[b].e[r];
        }
      }
    }
  }
}

// ---- host side ---------------------------------------------------------------------------

namespace {

enum { kDQ = 0, kDKs, kDItems, kDQperm, kDSlot, kDIds, kDPart, kDOut, kDThr, kDQmap, kDQsorted };

hipError_t dev_reserve(ListsWorkspace& ws, int i, size_t bytes) {
  if (bytes <= ws.dev_cap[i]) return hipSuccess;
  if (ws.dev[i]) (void)hipFree(ws.dev[i]);
  ws.dev[i] = nullptr;
  ws.dev_cap[i] = 0;
  const size_t want = std::max(bytes, size_t(4096));
  hipError_t e = hipMalloc(&ws.dev[i], want);
  if (e == hipSuccess) ws.dev_cap[i] = want;
  return e;
}

hipError_t host_reserve(ListsWorkspace& ws, int i, size_t bytes) {
  if (bytes <= ws.host_cap[i]) return hipSuccess;
  if (ws.host[i]) (void)hipHostFree(ws.host[i]);
  ws.host[i] = nullptr;
  ws.host_cap[i] = 0;
  const size_t want = std::max(bytes, size_t(4096));
  hipError_t e = hipHostMalloc(&ws.host[i], want, hipHostMallocDefault);
  if (e == hipSuccess) ws.host_cap[i] = want;
  return e;
}

struct ListsArgs {
  dim3 grid;
  size_t lds;
  hipStream_t st;
  const float* base;
  uint32_t dp, d;
  const float* q;
  const float* qsorted;  // the queries in item order (one lane per row)
  const uint32_t* ks;
  const ListItem* items;
  uint32_t n_items;
  const uint32_t* qperm;
  const uint32_t* ids;
  const uint64_t* slot_off;
  uint32_t L;
  uint64_t* part;
  uint32_t* counter;
  uint64_t* qthr;
};

template <int QB, int R, int D16, int KP, bool L2, int LPR>
hipError_t launch_lists_t(const ListsArgs& a) {
  if (a.lds > 65536) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&lists_scan_kernel<QB, R, D16, KP, L2, LPR>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(a.lds));
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((lists_scan_kernel<QB, R, D16, KP, L2, LPR>), a.grid, dim3(kBlock), a.lds, a.st,
                     a.base, a.dp, a.d, a.q, a.ks, a.items, a.n_items, a.qperm, a.ids, a.slot_off,
                     a.L, a.part, a.counter, a.qthr);
  return hipGetLastError();
}

template <int QB, int R, int D16, int LPR>
hipError_t launch_lists_k(bool l2, int kp, const ListsArgs& a) {
  if (kp == 64)
    return l2 ? launch_lists_t<QB, R, D16, 64, true, LPR>(a) : launch_lists_t<QB, R, D16, 64, false, LPR>(a);
  return l2 ? launch_lists_t<QB, R, D16, 128, true, LPR>(a) : launch_lists_t<QB, R, D16, 128, false, LPR>(a);
}

// quad layout (4 lanes per row): any d
template <int QB, int R>
hipError_t launch_lists_d(bool l2, int kp, const ListsArgs& a) {
  switch (a.d) {
    case 384: return launch_lists_k<QB, R, 24, 4>(l2, kp, a);
    case 768: return launch_lists_k<QB, R, 48, 4>(l2, kp, a);
    case 128: return launch_lists_k<QB, R, 8, 4>(l2, kp, a);
    default: return launch_lists_k<QB, R, 0, 4>(l2, kp, a);
  }
}

// eight lanes per row: d = 128, 384, 768
template <int QB, int R>
hipError_t launch_lists_oct(bool l2, int kp, const ListsArgs& a) {
  switch (a.d) {
    case 384: return launch_lists_k<QB, R, 24, 8>(l2, kp, a);
    case 768: return launch_lists_k<QB, R, 48, 8>(l2, kp, a);
    case 128: return launch_lists_k<QB, R, 8, 8>(l2, kp, a);
    default: return hipErrorInvalidValue;
  }
}

template <int D16, int KP, bool L2>
hipError_t launch_block_t(const ListsArgs& a) {
  hipLaunchKernelGGL((lists_block_kernel<D16, KP, L2>), a.grid, dim3(64 * kBlkWaves), 0, a.st, a.base, a.dp, a.q,
                     a.ks, a.items, a.n_items, a.qperm, a.ids, a.slot_off, a.L, a.part, a.counter, a.qthr);
  return hipGetLastError();
}

template <int D16>
hipError_t launch_block_k(bool l2, int kp, const ListsArgs& a) {
  if (kp == 64) return l2 ? launch_block_t<D16, 64, true>(a) : launch_block_t<D16, 64, false>(a);
  return l2 ? launch_block_t<D16, 128, true>(a) : launch_block_t<D16, 128, false>(a);
}

template <int D16, int KP, bool L2>
hipError_t launch_lane_t(const ListsArgs& a) {
  const void* f = reinterpret_cast<const void*>(&lists_lane_kernel<D16, KP, L2>);
  const size_t lds = kLaneLds;
  hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((lists_lane_kernel<D16, KP, L2>), a.grid, dim3(64 * kLaneWaves), lds, a.st, a.base, a.dp,
                     a.d, a.qsorted, a.ks, a.items, a.qperm, a.ids, a.slot_off, a.L, a.part, a.counter, a.qthr);
  return hipGetLastError();
}

template <int D16>
hipError_t launch_lane_k(bool l2, int kp, const ListsArgs& a) {
  if (kp == 64) return l2 ? launch_lane_t<D16, 64, true>(a) : launch_lane_t<D16, 64, false>(a);
  return l2 ? launch_lane_t<D16, 128, true>(a) : launch_lane_t<D16, 128, false>(a);
}

hipError_t launch_lists_tile(int qb, int r, int lpr, bool l2, int kp, const ListsArgs& a) {
  if (lpr == 1) {  // one lane per row (d a multiple of 32)
    switch (a.d) {
      case 384: return launch_lane_k<24>(l2, kp, a);
      case 768: return launch_lane_k<48>(l2, kp, a);
      case 128: return launch_lane_k<8>(l2, kp, a);
      default: return launch_lane_k<0>(l2, kp, a);
    }
  }
  if (qb == kBlkItemQ) {  // block-shared rows (r = kBlkR, eight lanes per row)
    switch (a.d) {
      case 384: return launch_block_k<24>(l2, kp, a);
      case 768: return launch_block_k<48>(l2, kp, a);
      case 128: return launch_block_k<8>(l2, kp, a);
      default: return hipErrorInvalidValue;
    }
  }
  if (lpr == 8) {
    if (qb == 8 && r == 4) return launch_lists_oct<8, 4>(l2, kp, a);
    if (qb == 8 && r == 2) return launch_lists_oct<8, 2>(l2, kp, a);
    if (qb == 16 && r == 2) return launch_lists_oct<16, 2>(l2, kp, a);
    return hipErrorInvalidValue;
  }
  if (qb == 8 && r == 2) return launch_lists_d<8, 2>(l2, kp, a);
  if (qb == 4 && r == 4) return launch_lists_d<4, 4>(l2, kp, a);
  if (qb == 2 && r == 4) return launch_lists_d<2, 4>(l2, kp, a);
  if (qb == 1 && r == 4) return launch_lists_d<1, 4>(l2, kp, a);
  return hipErrorInvalidValue;
}

std::string hip_msg(const char* what, hipError_t e) {
  return std::string(what) + ": " + hipGetErrorString(e);
}

int hip_status(hipError_t e) {
  return e == hipErrorOutOfMemory ? VSG_ERR_OUT_OF_MEMORY : VSG_ERR_DEVICE;
}

}  // namespace

bool lists_tile_supported(int qb, int r, int lpr, size_t dim) {
  if (lpr == 1) return r == 1 && dim % 32 == 0;  // kLaneQ queries per item, whatever qb
  if (qb == kBlkItemQ) return r == kBlkR && lpr == 8 && (dim == 128 || dim == 384 || dim == 768);
  if (lpr == 8)
    return (dim == 128 || dim == 384 || dim == 768) && ((qb == 8 && (r == 2 || r == 4)) || (qb == 16 && r == 2));
  return (qb == 8 && r == 2) || (qb == 4 && r == 4) || (qb == 2 && r == 4) || (qb == 1 && r == 4);
}

void ListsWorkspace::release() {
  for (int i = 0; i < 11; ++i) {
    if (dev[i]) (void)hipFree(dev[i]);
    dev[i] = nullptr;
    dev_cap[i] = 0;
  }
  for (int i = 0; i < 3; ++i) {
    if (host[i]) (void)hipHostFree(host[i]);
    host[i] = nullptr;
    host_cap[i] = 0;
  }
  if (counter) (void)hipFree(counter);
  counter = nullptr;
}

#define LISTS_HIP(expr)                        \
  do {                                         \
    hipError_t e_ = (expr);                    \
    if (e_ != hipSuccess) {                    \
      err = hip_msg(#expr, e_);                \
      return hip_status(e_);                   \
    }                                          \
  } while (0)

namespace {

// Host-side plan of a filtered batch: kept alive until the uploads that read it are complete.
struct ListsPlanHost {
  std::vector<ListItem> items;
  std::vector<uint32_t> qperm;
  std::vector<uint64_t> slot_off;
  uint32_t bounds[kQueues + 1] = {0};  // item range of each XCD queue
};

// Validates and plans a filtered batch, reserves its workspace and uploads the plan, ks and
// the allowed ids on st (and the host queries, when given). The uploads read `ph`.
int lists_prepare(ListsWorkspace& ws, size_t n, size_t dim, size_t dp, int metric, int num_cus,
                  hipStream_t st, const float* queries, size_t nq, const uint32_t* ks, size_t kmax,
                  const uint64_t* range, const uint32_t* allow_ids, size_t n_ids, std::string& err,
                  const ListsTiles& tiles, ListsTiles* used, ListsPlanHost& ph, ListsLaunch& out) {
  // ---- validate (hnsw.cpp:326: an empty allowed set returns nothing) ----
  const uint64_t total = n_ids;
  for (size_t i = 0; i < nq; ++i) {
    if (range[2 * i] == VSG_ALL_ROWS) {
      if (ks[i] > kmax) {
        err = "ks[i] must be <= kmax";
        return VSG_ERR_INVALID_ARGUMENT;
      }
      continue;
    }
    if (range[2 * i + 1] < range[2 * i] || range[2 * i + 1] > n_ids) {
      err = "each query's range must satisfy begin <= end <= len(allow_ids)";
      return VSG_ERR_INVALID_ARGUMENT;
    }
    if (ks[i] > kmax) {
      err = "ks[i] must be <= kmax";
      return VSG_ERR_INVALID_ARGUMENT;
    }
  }
  if (total > 0 && !allow_ids) {
    err = "null allow_ids";
    return VSG_ERR_INVALID_ARGUMENT;
  }
  // Group queries by identical range; each group's ids must be increasing and < n.
  // (unfiltered queries form the group keyed (VSG_ALL_ROWS, 0))
  std::map<std::pair<uint64_t, uint64_t>, std::vector<uint32_t>> groups;
  for (size_t i = 0; i < nq; ++i) {
    if (range[2 * i] == VSG_ALL_ROWS)
      groups[{VSG_ALL_ROWS, 0}].push_back(uint32_t(i));
    else
      groups[{range[2 * i], range[2 * i + 1]}].push_back(uint32_t(i));
  }
  auto span_of = [&](const std::pair<uint64_t, uint64_t>& key) {
    return key.first == VSG_ALL_ROWS ? std::make_pair(uint64_t(0), uint64_t(n)) : key;
  };
  for (const auto& kv : groups) {
    if (kv.first.first == VSG_ALL_ROWS) continue;
    const uint64_t lb = kv.first.first, le = kv.first.second;
    for (uint64_t p = lb; p < le; ++p) {
      if (allow_ids[p] >= n || (p > lb && allow_ids[p] <= allow_ids[p - 1])) {
        err = "each query's allowed ids must be strictly increasing and < size()";
        return VSG_ERR_INVALID_ARGUMENT;
      }
    }
  }
  if (total >= (uint64_t(1) << 32)) {
    err = "too many allowed ids in one call (limit 2^32 - 1)";
    return VSG_ERR_LENGTH;
  }

  // ---- plan: chunks and work items ----
  // queries per item: 4 waves x qb x dp floats of LDS must fit in 64 KB (160 KB for a
  // requested tile)
  int qb = tiles.qb ? tiles.qb : 4, r = tiles.r;
  const bool block = qb == kBlkItemQ;  // block-shared rows: its ring is static LDS
  const size_t lds_cap = tiles.qb ? size_t(160) << 10 : size_t(65536);
  if (block && !lists_tile_supported(qb, kBlkR, 8, dim)) qb = 8;
  while (!block && qb > 1 && size_t(kWavesPerBlock) * qb * dp * sizeof(float) > lds_cap) qb /= 2;
  if (!block && size_t(kWavesPerBlock) * qb * dp * sizeof(float) > lds_cap) {
    err = "dim too large for the filtered scan (at most 4096)";
    return VSG_ERR_INVALID_ARGUMENT;
  }
  int lpr = tiles.lanes_per_row == 8 || block ? 8 : 4;
  const bool lane_tile = !block && tiles.lanes_per_row == 1 && dim % 32 == 0;
  if (lane_tile) {
    lpr = 1;
    qb = kLaneQ;
    r = 1;
  }
  if (block) r = kBlkR;
  if (!r) r = lpr == 8 ? (qb >= 16 ? 2 : 4) : (qb >= 8 ? 2 : 4);
  if (!lists_tile_supported(qb, r, lpr, dim)) {  // fall back to the nearest quad tile
    lpr = 4;
    qb = std::min(qb, 8);
    r = qb >= 8 ? 2 : 4;
  }
  // chunk rows: aim for ~2 items per wave slot, within [256, 8192], and at most
  // kMaxListsPerQuery chunks per query
  uint64_t visits = 0;
  for (const auto& kv : groups) {
    const auto sp = span_of(kv.first);
    visits += (sp.second - sp.first) * ((kv.second.size() + qb - 1) / qb);
  }
  // item slots: workgroups of one item (block-shared rows, one lane per row: several items per
  // workgroup in turn) or waves of one item each
  const uint64_t wave_slots = uint64_t(num_cus) * (block ? 2 : lane_tile ? 3 : 16);
  uint64_t C = visits / (wave_slots * 2) + 1;
  C = std::min<uint64_t>(8192, std::max<uint64_t>(256, (C + 63) / 64 * 64));
  if (tiles.chunk) C = std::max<uint64_t>(64, (tiles.chunk + 63) / 64 * 64);
  uint64_t max_len = 0;
  for (const auto& kv : groups) {
    const auto sp = span_of(kv.first);
    max_len = std::max(max_len, sp.second - sp.first);
  }
  while ((max_len + C - 1) / C > kMaxListsPerQuery) C *= 2;

  if (used) {
    used->qb = qb;
    used->r = r;
    used->lanes_per_row = lpr;
    used->chunk = C;
    used->blocks_per_cu = lane_tile ? 1 : tiles.blocks_per_cu > 0 ? tiles.blocks_per_cu : block ? 2 : 4;
  }
  std::vector<ListItem>& items = ph.items;
  std::vector<uint32_t>& qperm = ph.qperm;
  std::vector<uint64_t>& slot_off = ph.slot_off;
  items.clear();
  qperm.clear();
  qperm.reserve(nq);
  slot_off.assign(nq + 1, 0);
  std::vector<uint32_t> nchunks_of(nq, 0);
  // The query blocks of one chunk are consecutive items, so its rows are re-read from cache
  // rather than from HBM. Chunks are dealt to eight queues (one per XCD, for the kernels that
  // use them), each to the queue with the least work so far; the items are then laid out queue
  // by queue, and a kernel with one queue runs them in that order. Chunks are dealt by chunk
  // index across the groups (every group's first chunk, then every second chunk, ...), so a
  // query's later chunks mostly start after its earlier ones have published their k-th key
  // (qthr) and admit few keys, instead of all of a large group's chunks starting from empty
  // lists at once.
  struct GroupPlan {
    uint64_t lb, le;
    uint32_t direct, nch, qoff0;
    const std::vector<uint32_t>* qs;
  };
  std::vector<GroupPlan> gp;
  gp.reserve(groups.size());
  uint32_t max_nch = 0;
  for (const auto& kv : groups) {
    const auto sp = span_of(kv.first);
    GroupPlan g;
    g.lb = sp.first;
    g.le = sp.second;
    g.direct = kv.first.first == VSG_ALL_ROWS ? 1u : 0u;
    g.nch = uint32_t((g.le - g.lb + C - 1) / C);
    g.qs = &kv.second;
    for (uint32_t q : kv.second) nchunks_of[q] = g.nch;
    g.qoff0 = uint32_t(qperm.size());
    for (uint32_t q : kv.second) qperm.push_back(q);
    max_nch = std::max(max_nch, g.nch);
    gp.push_back(g);
  }
  std::vector<std::vector<ListItem>> queue(kQueues);
  std::vector<uint64_t> qwork(kQueues, 0);
  for (uint32_t c = 0; c < max_nch; ++c)
    for (const GroupPlan& g : gp) {
      if (c >= g.nch) continue;
      const auto& qs = *g.qs;
      const uint32_t qi = uint32_t(std::min_element(qwork.begin(), qwork.end()) - qwork.begin());
      const uint64_t clb = g.lb + uint64_t(c) * C, cle = std::min<uint64_t>(g.le, clb + C);
      qwork[qi] += (cle - clb) * qs.size();
      for (size_t b0 = 0; b0 < qs.size(); b0 += qb) {
        ListItem it;
        it.lb = uint32_t(clb);
        it.le = uint32_t(cle);
        it.chunk = c;
        it.qoff = g.qoff0 + uint32_t(b0);
        it.nq = uint32_t(std::min<size_t>(qb, qs.size() - b0));
        it.direct = g.direct;
        queue[qi].push_back(it);
      }
    }
  uint32_t bounds[kQueues + 1] = {0};
  for (uint32_t qi = 0; qi < kQueues; ++qi) {
    items.insert(items.end(), queue[qi].begin(), queue[qi].end());
    bounds[qi + 1] = uint32_t(items.size());
  }
  for (size_t i = 0; i < nq; ++i) slot_off[i + 1] = slot_off[i] + nchunks_of[i];
  const uint32_t L = uint32_t(kmax);
  const uint64_t n_slots = slot_off[nq];

  // ---- upload ----
  const size_t qbytes = nq * dim * sizeof(float);
  if (queries) {
    LISTS_HIP(dev_reserve(ws, kDQ, qbytes));
    LISTS_HIP(dev_reserve(ws, kDOut, nq * kmax * (sizeof(uint32_t) + sizeof(float))));
    LISTS_HIP(host_reserve(ws, 0, qbytes));
    LISTS_HIP(host_reserve(ws, 1, nq * kmax * sizeof(uint32_t)));
    LISTS_HIP(host_reserve(ws, 2, nq * kmax * sizeof(float)));
  }
  LISTS_HIP(dev_reserve(ws, kDKs, nq * sizeof(uint32_t)));
  LISTS_HIP(dev_reserve(ws, kDItems, std::max<size_t>(1, items.size()) * sizeof(ListItem)));
  LISTS_HIP(dev_reserve(ws, kDQperm, nq * sizeof(uint32_t)));
  LISTS_HIP(dev_reserve(ws, kDSlot, (nq + 1) * sizeof(uint64_t)));
  LISTS_HIP(dev_reserve(ws, kDIds, std::max<uint64_t>(1, total) * sizeof(uint32_t)));
  LISTS_HIP(dev_reserve(ws, kDPart, std::max<uint64_t>(1, n_slots) * L * sizeof(uint64_t)));
  LISTS_HIP(dev_reserve(ws, kDThr, nq * sizeof(uint64_t)));
  // + kLaneQ rows: an item of fewer queries still reads (and ignores) up to kLaneQ rows
  if (lane_tile) LISTS_HIP(dev_reserve(ws, kDQsorted, (nq + kLaneQ) * dim * sizeof(float)));
  if (!ws.counter) LISTS_HIP(hipMalloc(&ws.counter, kCounterWords * sizeof(uint32_t)));
  std::memcpy(ph.bounds, bounds, sizeof(bounds));
  LISTS_HIP(hipMemcpyAsync(ws.counter + kBoundsAt, ph.bounds, sizeof(ph.bounds), hipMemcpyHostToDevice, st));

  if (queries) {
    std::memcpy(ws.host[0], queries, qbytes);
    LISTS_HIP(hipMemcpyAsync(ws.dev[kDQ], ws.host[0], qbytes, hipMemcpyHostToDevice, st));
  }
  LISTS_HIP(hipMemcpyAsync(ws.dev[kDKs], ks, nq * sizeof(uint32_t), hipMemcpyHostToDevice, st));
  if (!items.empty())
    LISTS_HIP(hipMemcpyAsync(ws.dev[kDItems], items.data(), items.size() * sizeof(ListItem),
                             hipMemcpyHostToDevice, st));
  LISTS_HIP(hipMemcpyAsync(ws.dev[kDQperm], qperm.data(), nq * sizeof(uint32_t), hipMemcpyHostToDevice, st));
  LISTS_HIP(hipMemcpyAsync(ws.dev[kDSlot], slot_off.data(), (nq + 1) * sizeof(uint64_t),
                           hipMemcpyHostToDevice, st));
  if (total)
    LISTS_HIP(hipMemcpyAsync(ws.dev[kDIds], allow_ids, total * sizeof(uint32_t), hipMemcpyHostToDevice, st));

  const uint64_t want_blocks = block ? items.size() : (items.size() + kWavesPerBlock - 1) / kWavesPerBlock;
  const int bpc = tiles.blocks_per_cu > 0 ? tiles.blocks_per_cu : block ? 2 : 4;
  out.nq = uint32_t(nq);
  out.kmax = uint32_t(kmax);
  out.n_items = uint32_t(items.size());
  out.grid = uint32_t(std::min<uint64_t>(lane_tile ? items.size() : want_blocks,
                                          uint64_t(num_cus) * (lane_tile ? 1 : bpc)));
  out.lds = (block || lane_tile) ? 0 : size_t(kWavesPerBlock) * qb * dp * sizeof(float);
  out.dp = uint32_t(dp);
  out.d = uint32_t(dim);
  out.qb = qb;
  out.r = r;
  out.lpr = lpr;
  out.kp = kmax <= 64 ? 64 : 128;
  out.l2 = metric == VSG_METRIC_L2;
  out.chunk = C;
  return VSG_OK;
}

// work counters = 0 (the single queue's and the XCD queues') and every per-query threshold =
// the pad key (a kernel, so that a captured step holds no memset node)
__global__ void lists_reset_kernel(uint32_t* counter, uint64_t* qthr, uint32_t nq) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= kQueues) counter[i] = 0u;
  if (i < nq) qthr[i] = ~uint64_t(0);
}

ListsArgs lists_args(ListsWorkspace& ws, const ListsLaunch& L, const float* base, const float* d_q, hipStream_t st) {
  ListsArgs a;
  a.grid = dim3(L.grid);
  a.lds = L.lds;
  a.st = st;
  a.base = base;
  a.dp = L.dp;
  a.d = L.d;
  a.q = d_q;
  a.qsorted = static_cast<const float*>(ws.dev[kDQsorted]);
  a.ks = static_cast<const uint32_t*>(ws.dev[kDKs]);
  a.items = static_cast<const ListItem*>(ws.dev[kDItems]);
  a.n_items = L.n_items;
  a.qperm = static_cast<const uint32_t*>(ws.dev[kDQperm]);
  a.ids = static_cast<const uint32_t*>(ws.dev[kDIds]);
  a.slot_off = static_cast<const uint64_t*>(ws.dev[kDSlot]);
  a.L = L.kmax;
  a.part = static_cast<uint64_t*>(ws.dev[kDPart]);
  a.counter = ws.counter;
  a.qthr = static_cast<uint64_t*>(ws.dev[kDThr]);
  return a;
}

// Work items of the unfiltered queries i < rq with flags[i] != 0 (all of them when *all != 0),
// in the layout the host planner gives one unfiltered group: chunk-major, qb queries per item.
// The n_act items fill the END of the items array (max_items entries) and the work counter
// starts at max_items - n_act, so the scan kernel's bound stays the capture-time max_items;
// with no query chosen every block exits on its first fetch. qmap lists the chosen queries
// for the merge, unused slots kNoneId.
__global__ __launch_bounds__(1024) void lists_fallback_plan_kernel(
    const uint32_t* __restrict__ flags, const uint32_t* __restrict__ all, uint32_t rq, uint32_t rows,
    uint32_t C, uint32_t nch, uint32_t qb, uint32_t max_items, ListItem* __restrict__ items,
    uint32_t* __restrict__ qperm, uint32_t* __restrict__ qmap, uint64_t* __restrict__ slot_off,
    uint32_t* __restrict__ counter, uint64_t* __restrict__ qthr) {
  __shared__ uint32_t wsum[16];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const bool every = *all != 0u;
  const uint32_t per = (rq + 1023) / 1024;  // consecutive queries per thread (rq <= 4096)
  const uint32_t q0 = min(rq, tid * per), q1 = min(rq, q0 + per);
  uint32_t cnt = 0;
  for (uint32_t q = q0; q < q1; ++q) cnt += (every || flags[q] != 0u) ? 1u : 0u;
  uint32_t incl = cnt;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(incl, off, 64);
    if (lane >= uint32_t(off)) incl += y;
  }
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  uint32_t before = 0, nf = 0;
  for (uint32_t w = 0; w < 16; ++w) {
    before += w < wave ? wsum[w] : 0u;
    nf += wsum[w];
  }
  uint32_t pos = before + incl - cnt;
  for (uint32_t q = q0; q < q1; ++q) {
    slot_off[q] = uint64_t(pos) * nch;
    if (every || flags[q] != 0u) {
      qperm[pos] = q;
      qmap[pos] = q;
      ++pos;
    }
    qthr[q] = ~uint64_t(0);
  }
  if (tid == 0) slot_off[rq] = uint64_t(nf) * nch;
  for (uint32_t i = nf + tid; i < rq; i += 1024) qmap[i] = kNoneId;
  const uint32_t nb = (nf + qb - 1) / qb, n_act = nch * nb;
  for (uint32_t j = tid; j < n_act; j += 1024) {
    const uint32_t c = j / nb, b = j - c * nb;
    ListItem it;
    it.lb = uint32_t(uint64_t(c) * C);
    it.le = uint32_t(min(uint64_t(rows), uint64_t(c + 1) * C));
    it.chunk = c;
    it.qoff = b * qb;
    it.nq = min(qb, nf - b * qb);
    it.direct = 1;
    items[max_items - n_act + j] = it;
  }
  if (tid == 0) *counter = max_items - n_act;
}

}  // namespace

hipError_t lists_enqueue(ListsWorkspace& ws, const ListsLaunch& L, const float* base, const float* d_q,
                         uint32_t* d_ids, float* d_dists, hipStream_t st, TimingSink* timing) {
  hipLaunchKernelGGL(lists_reset_kernel, dim3((L.nq + 255) / 256 + (L.nq == 0)), dim3(256), 0, st, ws.counter,
                     static_cast<uint64_t*>(ws.dev[kDThr]), L.nq);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (L.n_items) {
    const ListsArgs a = lists_args(ws, L, base, d_q, st);
    if (L.lpr == 1) {
      hipLaunchKernelGGL(lists_gather_queries_kernel, dim3(L.nq), dim3(128), 0, st, d_q, a.qperm, L.nq, L.d,
                         static_cast<float*>(ws.dev[kDQsorted]));
      if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (timing) timing->begin(st);
    if ((e = launch_lists_tile(L.qb, L.r, L.lpr, L.l2, L.kp, a)) != hipSuccess) return e;
    if (timing) timing->end(st, VSG_KERNEL_LISTS, 1);
  }
  if (timing) timing->begin(st);
  e = launch_merge(L.nq, static_cast<const uint64_t*>(ws.dev[kDPart]), static_cast<const uint64_t*>(ws.dev[kDSlot]),
                   0, L.kmax, static_cast<const uint32_t*>(ws.dev[kDKs]), 0, nullptr, d_ids, d_dists, L.kmax, st);
  if (e != hipSuccess) return e;
  if (timing) timing->end(st, VSG_KERNEL_MERGE, 1);
  return hipSuccess;
}

int search_lists_host(ListsWorkspace& ws, const float* base, size_t n, size_t dim, size_t dp,
                      int metric, int num_cus, hipStream_t st, const float* queries, size_t nq,
                      const uint32_t* ks, size_t kmax, const uint64_t* range,
                      const uint32_t* allow_ids, size_t n_ids, uint32_t* out_ids,
                      float* out_dists, std::string& err, TimingSink* timing,
                      const ListsTiles& tiles, ListsTiles* used) {
  ListsPlanHost ph;
  ListsLaunch L;
  const int rc = lists_prepare(ws, n, dim, dp, metric, num_cus, st, queries, nq, ks, kmax, range, allow_ids,
                               n_ids, err, tiles, used, ph, L);
  if (rc) return rc;
  uint32_t* d_ids = static_cast<uint32_t*>(ws.dev[kDOut]);
  float* d_dists = reinterpret_cast<float*>(d_ids + nq * kmax);
  LISTS_HIP(lists_enqueue(ws, L, base, static_cast<const float*>(ws.dev[kDQ]), d_ids, d_dists, st, timing));
  LISTS_HIP(hipMemcpyAsync(ws.host[1], d_ids, nq * kmax * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  LISTS_HIP(hipMemcpyAsync(ws.host[2], d_dists, nq * kmax * sizeof(float), hipMemcpyDeviceToHost, st));
  LISTS_HIP(hipStreamSynchronize(st));
  std::memcpy(out_ids, ws.host[1], nq * kmax * sizeof(uint32_t));
  std::memcpy(out_dists, ws.host[2], nq * kmax * sizeof(float));
  return VSG_OK;
}

int prepare_lists_fallback(ListsWorkspace& ws, size_t n, size_t dim, size_t dp, int metric, int num_cus,
                           hipStream_t st, size_t nq, size_t k, std::string& err, ListsLaunch& out) {
  if (nq == 0 || nq > 4096 || k == 0) {
    err = "the exact fallback plans 1 to 4096 queries with k > 0";
    return VSG_ERR_INVALID_ARGUMENT;
  }
  // every query chosen: the largest plan, whose buffers then fit every subset
  const std::vector<uint64_t> range(2 * nq, VSG_ALL_ROWS);
  const std::vector<uint32_t> ks(nq, uint32_t(k));
  ListsPlanHost ph;
  int rc = lists_prepare(ws, n, dim, dp, metric, num_cus, st, nullptr, nq, ks.data(), k, range.data(), nullptr, 0,
                         err, ListsTiles{}, nullptr, ph, out);
  if (rc) return rc;
  out.rows = uint32_t(n);
  out.nch = uint32_t((n + out.chunk - 1) / out.chunk);
  if (uint64_t(out.nch) * ((nq + out.qb - 1) / out.qb) != out.n_items) {
    err = "internal: fallback plan layout";
    return VSG_ERR_DEVICE;
  }
  LISTS_HIP(dev_reserve(ws, kDQmap, nq * sizeof(uint32_t)));
  LISTS_HIP(hipStreamSynchronize(st));  // the plan's host vectors are read by the uploads
  return VSG_OK;
}

hipError_t lists_fallback_enqueue(ListsWorkspace& ws, const ListsLaunch& L, const float* base, const float* d_q,
                                  uint32_t rq, const uint32_t* flags, const uint32_t* all, uint32_t* d_ids,
                                  float* d_dists, uint32_t ldo, hipStream_t st) {
  if (rq == 0 || rq > L.nq) return hipErrorInvalidValue;
  hipLaunchKernelGGL(lists_fallback_plan_kernel, dim3(1), dim3(1024), 0, st, flags, all, rq, L.rows,
                     uint32_t(L.chunk), L.nch, uint32_t(L.qb), L.n_items, static_cast<ListItem*>(ws.dev[kDItems]),
                     static_cast<uint32_t*>(ws.dev[kDQperm]), static_cast<uint32_t*>(ws.dev[kDQmap]),
                     static_cast<uint64_t*>(ws.dev[kDSlot]), ws.counter, static_cast<uint64_t*>(ws.dev[kDThr]));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  ListsArgs a = lists_args(ws, L, base, d_q, st);
  if ((e = launch_lists_tile(L.qb, L.r, L.lpr, L.l2, L.kp, a)) != hipSuccess) return e;
  return launch_merge(rq, static_cast<const uint64_t*>(ws.dev[kDPart]), static_cast<const uint64_t*>(ws.dev[kDSlot]),
                      0, L.kmax, static_cast<const uint32_t*>(ws.dev[kDKs]), 0,
                      static_cast<const uint32_t*>(ws.dev[kDQmap]), d_ids, d_dists, ldo, st);
}

int prepare_lists_device(ListsWorkspace& ws, size_t n, size_t dim, size_t dp, int metric, int num_cus,
                         hipStream_t st, size_t nq, const uint32_t* ks, size_t kmax, const uint64_t* range,
                         const uint32_t* allow_ids, size_t n_ids, std::string& err, const ListsTiles& tiles,
                         ListsLaunch& out) {
  ListsPlanHost ph;
  const int rc = lists_prepare(ws, n, dim, dp, metric, num_cus, st, nullptr, nq, ks, kmax, range, allow_ids,
                               n_ids, err, tiles, nullptr, ph, out);
  if (rc) return rc;
  LISTS_HIP(hipStreamSynchronize(st));  // the plan's host vectors are read by the uploads
  return VSG_OK;
}

}  // namespace vsg
