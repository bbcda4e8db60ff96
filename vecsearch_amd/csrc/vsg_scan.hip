// Exact streaming scan with a fused per-wave top-k, and the per-query merge of wave lists.
//
// What it computes: for each query q and base row x, the reference CPU distance
// (include/vecsearch/distance.h:30-73 via src/brute_force.cpp:14-74) in the same FP32
// operation order, so every distance is bit-identical to the reference's; then the k smallest
// by (dist, id) (brute_force.cpp:152-168). It takes the place of the reference's skinny/naive
// kernels, its distance matrix and its bitonic select passes (src/gpu_brute_force.cu:75-140,
// 220-314, 442-482): no distance is ever written to HBM.
//
// Layout: the scan reads the index's tiled copy of the base: 16-row tiles, rows padded to
// dp = ceil(d/16)*16 floats, and inside a tile the 64-byte blocks of one 16-dim step of all 16
// rows are contiguous. A quad of lanes shares a row, so each wave instruction reads one
// contiguous 1 KB block (quad_dists in vsg_kernels.h, STEP = 256), with non-temporal loads:
// the base is streamed once per pass. A lane works on R rows and QB queries per step; the QB
// queries of a pass sit in LDS.
//
// Top-k: each wave keeps a sorted k-list per query in registers (WaveList). A row is admitted
// only if its (dist, id) key beats the list's k-th key; admissions come from one ballot per
// (row group, query) and are rare after the first rows. Each wave writes its k-list; the
// merge kernel picks the k smallest keys of a query over all its lists.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "vsg_internal.h"
#include "vsg_kernels.h"

namespace vsg {

// One pass: queries [q0, q0 + nqp) of the batch (nqp <= QB) against the base. Wave gw scans
// rows [gw * rpw, min(n, (gw + 1) * rpw)) (rpw a multiple of 64); the block then merges its
// waves' k-lists and writes one k-list per query of the pass to part[(q * W + block) * k + slot]
// (W = number of blocks). mask (nullable): rows with mask[row] == 0 carry no result.
// D16 > 0 specialises d = 16 * D16; D16 == 0 handles any d.
template <int QB, int R, int D16, int KP, bool L2>
__global__ __launch_bounds__(kBlock) void scan_exact_kernel(
    const float* __restrict__ tiles, uint32_t n, uint32_t dp, uint32_t d,
    const float* __restrict__ queries, uint32_t q0, uint32_t nqp, uint32_t k, uint32_t rpw,
    uint32_t W, const uint8_t* __restrict__ mask, uint64_t* __restrict__ part) {
  extern __shared__ float q_lds[];  // QB x dp, then (after the scan) QB x 4 x k merge keys
  const int tid = threadIdx.x, lane = tid & 63, t = lane & 3, g = lane >> 2;
  const uint32_t gw = blockIdx.x * kWavesPerBlock + (tid >> 6);

  for (uint32_t i = tid; i < uint32_t(QB) * dp; i += kBlock) {
    const uint32_t b = i / dp, c = i - b * dp;
    q_lds[i] = (b < nqp && c < d) ? queries[size_t(q0 + b) * d + c] : 0.f;
  }
  __syncthreads();

  WaveList<KP> lst[QB];
  uint64_t thr[QB];
#pragma unroll
  for (int b = 0; b < QB; ++b) {
    lst[b].init();
    thr[b] = kPadKey;
  }

  const uint32_t row_begin = min(n, gw * rpw), row_end = min(n, row_begin + rpw);
  for (uint32_t r0 = row_begin; r0 < row_end; r0 += 16 * R) {
    uint32_t row[R];
    bool ok[R];
    const float* xr[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
      row[j] = r0 + j * 16 + g;
      const bool in = row[j] < row_end;
      ok[j] = in && (mask == nullptr || mask[row[j]] != 0);
      const uint32_t rr = in ? row[j] : row_begin;  // readable row for lanes past the end
      xr[j] = tiles + size_t(rr >> 4) * 16 * dp + (rr & 15) * 16;
    }
    float dist[R][QB];
    quad_dists<QB, R, D16, L2, 256, kLoadStream>(xr, ok, q_lds, dp, d, t, dist);
#pragma unroll
    for (int j = 0; j < R; ++j)
#pragma unroll
      for (int b = 0; b < QB; ++b) {
        if (uint32_t(b) >= nqp) continue;
        const uint64_t key = (ok[j] && t == 0) ? make_key(dist[j][b], row[j]) : kPadKey;
        thr[b] = offer<KP>(lst[b], thr[b], key, lane, int(k));
      }
  }

  // block merge: the wave lists go to LDS (reusing the query area), one list per query leaves
  __syncthreads();
  uint64_t* keys = reinterpret_cast<uint64_t*>(q_lds);
  const int wave = tid >> 6;
#pragma unroll
  for (int b = 0; b < QB; ++b) {
    if (uint32_t(b) >= nqp) continue;
#pragma unroll
    for (int r = 0; r < KP / 64; ++r) {
      const uint32_t s = r * 64 + lane;
      if (s < k) keys[(b * kWavesPerBlock + wave) * k + s] = lst[b].e[r];
    }
  }
  __syncthreads();
  for (uint32_t b = 0; b < nqp; ++b)
    block_rank_merge<kWavesPerBlock>(keys + b * kWavesPerBlock * k, k,
                                     part + (size_t(q0 + b) * W + blockIdx.x) * k, tid, kBlock);
}

// ---- merge -------------------------------------------------------------------------------
// One block of 256 threads per query reduces that query's sorted lists to its k smallest keys.
// Lists of query q: part + (list_off ? list_off[q] : q * W) * L, count list_off[q+1] -
// list_off[q] (or W), L >= k entries each. Query q returns kq = (ks ? ks[q] : k) results into
// row q of the outputs (stride ldo), padding the rest of the row with (kNone, +inf).
//
// 1. Threshold: m = ceil(kq / #lists) leading keys of every list form a sample of >= kq keys;
//    its kq-th smallest key T is >= the answer's kq-th key (a subset's kq-th smallest is never
//    below the whole set's), and because the lists are sorted and cover disjoint rows, T is
//    close to it. The sample is ranked in LDS (by counting up to 128 keys, else by a bitonic
//    sort of up to kGather keys).
// 2. Every thread walks its lists from the head while keys are <= T and appends them to LDS.
// 3. The gathered keys (about kq of them) are ranked: by counting (<= 256 keys) or by a
//    bitonic sort (<= kGather keys). Keys of real rows are unique (a row is scanned once per
//    query), so ranks are distinct.
// 4. A sample or gather larger than kGather keys (tiny lists, huge list counts, or many exact
//    ties) falls back to kq rounds of a block-wide minimum over the list heads.
constexpr int kGather = 4096;

__device__ __forceinline__ uint64_t wave_min64(uint64_t v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const uint32_t lo = __shfl_xor(uint32_t(v), off, 64);
    const uint32_t hi = __shfl_xor(uint32_t(v >> 32), off, 64);
    const uint64_t o = (uint64_t(hi) << 32) | lo;
    v = o < v ? o : v;
  }
  return v;
}

__device__ __forceinline__ uint64_t block_min64(uint64_t v, uint64_t* red) {
  const uint64_t wm = wave_min64(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = wm;
  __syncthreads();
  uint64_t m = red[0];
#pragma unroll
  for (int i = 1; i < 4; ++i) m = red[i] < m ? red[i] : m;
  __syncthreads();
  return m;
}

// Ascending bitonic sort of buf[0, n2) (n2 a power of two) by a 256-thread block.
__device__ void bitonic_sort(uint64_t* buf, uint32_t n2, uint32_t tid) {
  for (uint32_t size = 2; size <= n2; size <<= 1) {
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      for (uint32_t i = tid; i < n2 / 2; i += 256) {
        const uint32_t lo = 2 * i - (i & (stride - 1));
        const uint32_t hi = lo + stride;
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
}

__device__ __forceinline__ void put_key(uint32_t* oi, float* od, uint32_t r, uint64_t key) {
  oi[r] = key == kPadKey ? kNoneId : key_id(key);
  od[r] = key_dist(key);
}

__global__ __launch_bounds__(256) void merge_lists_kernel(
    const uint64_t* __restrict__ part, const uint64_t* __restrict__ list_off, uint32_t W,
    uint32_t L, const uint32_t* __restrict__ ks, uint32_t k, const uint32_t* __restrict__ qmap,
    uint32_t* __restrict__ out_ids, float* __restrict__ out_dists, uint32_t ldo) {
  __shared__ uint64_t red[4];
  __shared__ uint64_t buf[kGather];
  __shared__ uint32_t cnt;
  const uint32_t q = qmap ? qmap[blockIdx.x] : blockIdx.x, tid = threadIdx.x;
  if (q == kNoneId) return;  // an unused output slot of a device-chosen query set
  const uint64_t l0 = list_off ? list_off[q] : uint64_t(q) * W;
  const uint32_t nl = list_off ? uint32_t(list_off[q + 1] - l0) : W;
  const uint32_t kq = min(min(ks ? ks[q] : k, ldo), L);
  const uint64_t* lists = part + l0 * L;
  uint32_t* oi = out_ids + size_t(q) * ldo;
  float* od = out_dists + size_t(q) * ldo;
  for (uint32_t c = kq + tid; c < ldo; c += 256) {  // columns beyond this query's k
    oi[c] = kNoneId;
    od[c] = __builtin_inff();
  }
  if (kq == 0 || nl == 0) {
    for (uint32_t c = tid; c < kq; c += 256) put_key(oi, od, c, kPadKey);
    return;
  }

  // 1. threshold from a sample of every list's leading keys
  const uint32_t m = (kq + nl - 1) / nl;
  const uint32_t ns = nl * m;
  uint64_t T = kPadKey;
  if (ns <= uint32_t(kGather)) {
    for (uint32_t i = tid; i < ns; i += 256) {
      const uint32_t l = i / m, p = i - l * m;
      buf[i] = lists[size_t(l) * L + p];
    }
    __syncthreads();
    if (ns <= 128) {  // rank by counting: ns LDS reads per key
      for (uint32_t i = tid; i < ns; i += 256) {
        const uint64_t v = buf[i];
        uint32_t below = 0, equal_before = 0;
        for (uint32_t j = 0; j < ns; ++j) {
          const uint64_t u = buf[j];
          below += u < v;
          equal_before += (u == v) & (j < i);
        }
        if (below + equal_before == kq - 1) red[0] = v;  // the unique key of rank kq - 1
      }
      __syncthreads();
      T = red[0];
    } else {  // radix select of the kq-th smallest high word (the distance bits): every key with
              // that high word or a smaller one is kept, at least kq of them
      __shared__ uint32_t hist[256], sel[2];
      const uint32_t hi = block_select_hi32(buf, ns, kq, hist, sel, tid);
      T = (uint64_t(hi) << 32) | 0xFFFFFFFFull;
    }
    __syncthreads();
  }
  if (tid == 0) cnt = 0;
  __syncthreads();

  // 2. gather real keys <= T
  for (uint32_t l = tid; l < nl; l += 256) {
    const uint64_t* lp = lists + size_t(l) * L;
    for (uint32_t p = 0; p < L; ++p) {
      const uint64_t v = lp[p];
      if (v > T || v == kPadKey) break;
      const uint32_t slot = atomicAdd(&cnt, 1u);
      if (slot < uint32_t(kGather)) buf[slot] = v;
    }
  }
  __syncthreads();
  const uint32_t c = cnt;

  if (c <= uint32_t(kGather)) {
    if (c <= 256) {  // 3a. rank by counting
      if (tid < c) {
        const uint64_t v = buf[tid];
        uint32_t rank = 0;
        for (uint32_t j = 0; j < c; ++j) rank += buf[j] < v;
        if (rank < kq) put_key(oi, od, rank, v);
      }
    } else {  // 3b. bitonic sort of the next power of two
      uint32_t n2 = 512;
      while (n2 < c) n2 <<= 1;
      for (uint32_t i = c + tid; i < n2; i += 256) buf[i] = kPadKey;
      __syncthreads();
      bitonic_sort(buf, n2, tid);
      for (uint32_t r = tid; r < kq && r < c; r += 256) put_key(oi, od, r, buf[r]);
    }
    for (uint32_t r = c + tid; r < kq; r += 256) put_key(oi, od, r, kPadKey);
    return;
  }

  // 4. fallback: k rounds over the list heads (heads cached in registers)
  uint32_t pos[kMergeHeads];
  uint64_t head[kMergeHeads];
  uint64_t best = kPadKey;
  int best_h = -1;
#pragma unroll
  for (int h = 0; h < kMergeHeads; ++h) {
    pos[h] = 0;
    const uint32_t l = tid + h * 256;
    head[h] = l < nl ? lists[size_t(l) * L] : kPadKey;
    if (head[h] < best) {
      best = head[h];
      best_h = h;
    }
  }
  for (uint32_t r = 0; r < kq; ++r) {
    const uint64_t m = block_min64(best, red);
    if (m == kPadKey) {
      for (uint32_t rr = r + tid; rr < kq; rr += 256) put_key(oi, od, rr, kPadKey);
      return;
    }
    if (tid == 0) put_key(oi, od, r, m);
    if (best == m) {  // this thread owns the winner: advance that list
      best = kPadKey;
      int nh = -1;
#pragma unroll
      for (int h = 0; h < kMergeHeads; ++h) {
        if (h == best_h) {
          pos[h]++;
          head[h] = pos[h] < L ? lists[size_t(tid + h * 256) * L + pos[h]] : kPadKey;
        }
        if (head[h] < best) {
          best = head[h];
          nh = h;
        }
      }
      best_h = nh;
    }
  }
}

// Writes rows [r0, r1) of the row-major base into the tiled copy (see the layout note above).
__global__ void tile_rows_kernel(const float* __restrict__ base, float* __restrict__ tiles,
                                 uint32_t r0, uint32_t r1, uint32_t dp) {
  const size_t total = size_t(r1 - r0) * dp;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += size_t(gridDim.x) * blockDim.x) {
    const uint32_t r = r0 + uint32_t(i / dp), j = uint32_t(i % dp);
    tiles[size_t(r >> 4) * 16 * dp + (j >> 4) * 256 + (r & 15) * 16 + (j & 15)] = base[size_t(r) * dp + j];
  }
}

// Row gather / scatter of 32-bit words: dst[i] = src[idx[i]] (gather) or dst[idx[i]] = src[i]
// (scatter), rows of `width` words, for i < rows.
__global__ void rows_move_kernel(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                 const uint32_t* __restrict__ idx, uint32_t rows, uint32_t width,
                                 int scatter) {
  const size_t total = size_t(rows) * width;
  for (size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
       t += size_t(gridDim.x) * blockDim.x) {
    const uint32_t i = uint32_t(t / width), c = uint32_t(t - size_t(i) * width);
    if (scatter)
      dst[size_t(idx[i]) * width + c] = src[t];
    else
      dst[t] = src[size_t(idx[i]) * width + c];
  }
}

__global__ void fill_missing_kernel(uint32_t* __restrict__ ids, float* __restrict__ dists,
                                    size_t count) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
       i += size_t(gridDim.x) * blockDim.x) {
    ids[i] = kNoneId;
    dists[i] = __builtin_inff();
  }
}

// ---- launch helpers (host) ---------------------------------------------------------------

namespace {

struct ScanArgs {
  dim3 grid;
  size_t lds;
  hipStream_t st;
  const float* tiles;
  uint32_t n, dp, d;
  const float* q;
  uint32_t q0, nqp, k, rpw, W;
  const uint8_t* mask;
  uint64_t* part;
};

template <int QB, int R, int D16, int KP, bool L2>
hipError_t launch_scan_t(const ScanArgs& a) {
  hipLaunchKernelGGL((scan_exact_kernel<QB, R, D16, KP, L2>), a.grid, dim3(kBlock), a.lds, a.st,
                     a.tiles, a.n, a.dp, a.d, a.q, a.q0, a.nqp, a.k, a.rpw, a.W, a.mask, a.part);
  return hipGetLastError();
}

template <int QB, int R, int D16>
hipError_t launch_scan_k(bool l2, int kp, const ScanArgs& a) {
  if (kp == 64) return l2 ? launch_scan_t<QB, R, D16, 64, true>(a) : launch_scan_t<QB, R, D16, 64, false>(a);
  return l2 ? launch_scan_t<QB, R, D16, 128, true>(a) : launch_scan_t<QB, R, D16, 128, false>(a);
}

template <int QB, int R>
hipError_t launch_scan_d(bool l2, int kp, const ScanArgs& a) {
  switch (a.d) {
    case 384: return launch_scan_k<QB, R, 24>(l2, kp, a);
    case 768: return launch_scan_k<QB, R, 48>(l2, kp, a);
    case 128: return launch_scan_k<QB, R, 8>(l2, kp, a);
    default: return launch_scan_k<QB, R, 0>(l2, kp, a);
  }
}

}  // namespace

hipError_t launch_scan(int qb, bool l2, uint32_t k, uint32_t n_blocks, hipStream_t st,
                       const float* tiles, uint32_t n, uint32_t dp, uint32_t d, const float* q,
                       uint32_t q0, uint32_t nqp, uint32_t rpw, const uint8_t* mask,
                       uint64_t* part) {
  ScanArgs a;
  a.grid = dim3(n_blocks);
  a.lds = std::max(size_t(qb) * dp * sizeof(float), size_t(qb) * kWavesPerBlock * k * sizeof(uint64_t));
  a.st = st;
  a.tiles = tiles;
  a.n = n;
  a.dp = dp;
  a.d = d;
  a.q = q;
  a.q0 = q0;
  a.nqp = nqp;
  a.k = k;
  a.rpw = rpw;
  a.W = n_blocks;
  a.mask = mask;
  a.part = part;
  const int kp = k <= 64 ? 64 : 128;
  switch (qb) {
    case 1: return launch_scan_d<1, 4>(l2, kp, a);
    case 2: return launch_scan_d<2, 4>(l2, kp, a);
    case 4: return launch_scan_d<4, 2>(l2, kp, a);
    default: return launch_scan_d<8, 2>(l2, kp, a);
  }
}

hipError_t launch_merge(uint32_t n_out, const uint64_t* part, const uint64_t* list_off, uint32_t W,
                        uint32_t L, const uint32_t* ks, uint32_t k, const uint32_t* qmap,
                        uint32_t* ids, float* dists, uint32_t ldo, hipStream_t st) {
  if (n_out == 0) return hipSuccess;
  hipLaunchKernelGGL(merge_lists_kernel, dim3(n_out), dim3(256), 0, st, part, list_off, W, L, ks, k,
                     qmap, ids, dists, ldo);
  return hipGetLastError();
}

hipError_t launch_tile_rows(hipStream_t st, const float* base, float* tiles, uint32_t r0, uint32_t r1,
                            uint32_t dp) {
  if (r1 <= r0) return hipSuccess;
  const size_t total = size_t(r1 - r0) * dp;
  const unsigned blocks = unsigned(std::min<size_t>(8192, (total + 255) / 256));
  hipLaunchKernelGGL(tile_rows_kernel, dim3(blocks), dim3(256), 0, st, base, tiles, r0, r1, dp);
  return hipGetLastError();
}

hipError_t launch_rows_move(hipStream_t st, const void* src, void* dst, const uint32_t* idx,
                            uint32_t rows, uint32_t width, bool scatter) {
  if (rows == 0 || width == 0) return hipSuccess;
  const size_t total = size_t(rows) * width;
  const unsigned blocks = unsigned(std::min<size_t>((total + 255) / 256, 4096));
  hipLaunchKernelGGL(rows_move_kernel, dim3(blocks), dim3(256), 0, st, static_cast<const uint32_t*>(src),
                     static_cast<uint32_t*>(dst), idx, rows, width, scatter ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_fill_missing(hipStream_t st, uint32_t* ids, float* dists, size_t count) {
  if (count == 0) return hipSuccess;
  const unsigned blocks = unsigned(std::min<size_t>(1024, (count + 255) / 256));
  hipLaunchKernelGGL(fill_missing_kernel, dim3(blocks), dim3(256), 0, st, ids, dists, count);
  return hipGetLastError();
}

}  // namespace vsg
