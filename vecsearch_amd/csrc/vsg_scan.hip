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
// rows [gw * rpw, min(n, (gw + 1) * rpw)) (rpw a multiple of 64) and writes its k-list of each
// query of the pass to part[(q * W + gw) * k + slot]. mask (nullable): rows with
// mask[row] == 0 carry no result. D16 > 0 specialises d = 16 * D16; D16 == 0 handles any d.
template <int QB, int R, int D16, int KP, bool L2>
__global__ __launch_bounds__(kBlock) void scan_exact_kernel(
    const float* __restrict__ tiles, uint32_t n, uint32_t dp, uint32_t d,
    const float* __restrict__ queries, uint32_t q0, uint32_t nqp, uint32_t k, uint32_t rpw,
    uint32_t W, const uint8_t* __restrict__ mask, uint64_t* __restrict__ part) {
  extern __shared__ float q_lds[];  // QB x dp
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

#pragma unroll
  for (int b = 0; b < QB; ++b) {
    if (uint32_t(b) >= nqp) continue;
    uint64_t* dst = part + (size_t(q0 + b) * W + gw) * k;
#pragma unroll
    for (int r = 0; r < KP / 64; ++r) {
      const uint32_t s = r * 64 + lane;
      if (s < k) dst[s] = lst[b].e[r];
    }
  }
}

// ---- merge -------------------------------------------------------------------------------
// One block of 256 threads per query over that query's sorted lists: k rounds of a
// block-wide minimum over the list heads. Keys of real rows are unique (a row is scanned by
// one wave of a query), so exactly one thread owns the winner and advances that list.
// Lists of query q: part + (list_off ? list_off[q] : q * W) * L, count list_off[q+1] -
// list_off[q] (or W), L entries each. Query q returns ks ? ks[q] : k results into row q of
// the outputs (stride ldo), padding the rest of the row with (kNone, +inf).
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

__global__ __launch_bounds__(256) void merge_lists_kernel(
    const uint64_t* __restrict__ part, const uint64_t* __restrict__ list_off, uint32_t W,
    uint32_t L, const uint32_t* __restrict__ ks, uint32_t k, const uint32_t* __restrict__ qmap,
    uint32_t* __restrict__ out_ids, float* __restrict__ out_dists, uint32_t ldo) {
  __shared__ uint64_t red[4];
  const uint32_t q = qmap ? qmap[blockIdx.x] : blockIdx.x, tid = threadIdx.x;
  const uint64_t l0 = list_off ? list_off[q] : uint64_t(q) * W;
  const uint32_t nl = list_off ? uint32_t(list_off[q + 1] - l0) : W;
  const uint32_t kq = min(ks ? ks[q] : k, ldo);
  const uint64_t* lists = part + l0 * L;
  uint32_t* oi = out_ids + size_t(q) * ldo;
  float* od = out_dists + size_t(q) * ldo;
  for (uint32_t c = kq + tid; c < ldo; c += 256) {  // columns beyond this query's k
    oi[c] = kNoneId;
    od[c] = __builtin_inff();
  }

  // heads of this thread's lists stay in registers; advancing a list loads one new key
  uint32_t pos[kMergeHeads];
  uint64_t head[kMergeHeads];
  uint64_t best = kPadKey;
  int best_h = -1;
#pragma unroll
  for (int h = 0; h < kMergeHeads; ++h) {
    pos[h] = 0;
    const uint32_t l = tid + h * 256;
    head[h] = (l < nl && L > 0) ? lists[size_t(l) * L] : kPadKey;
    if (head[h] < best) {
      best = head[h];
      best_h = h;
    }
  }
  for (uint32_t r = 0; r < kq; ++r) {
    const uint64_t wm = wave_min64(best);
    if ((tid & 63) == 0) red[tid >> 6] = wm;
    __syncthreads();
    uint64_t m = red[0];
#pragma unroll
    for (int i = 1; i < 4; ++i) m = red[i] < m ? red[i] : m;
    __syncthreads();
    if (m == kPadKey) {
      for (uint32_t rr = r + tid; rr < kq; rr += 256) {
        oi[rr] = kNoneId;
        od[rr] = __builtin_inff();
      }
      return;
    }
    if (tid == 0) {
      oi[r] = key_id(m);
      od[r] = key_dist(m);
    }
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
  a.lds = size_t(qb) * dp * sizeof(float);
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
  a.W = n_blocks * kWavesPerBlock;
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

hipError_t launch_fill_missing(hipStream_t st, uint32_t* ids, float* dists, size_t count) {
  if (count == 0) return hipSuccess;
  const unsigned blocks = unsigned(std::min<size_t>(1024, (count + 255) / 256));
  hipLaunchKernelGGL(fill_missing_kernel, dim3(blocks), dim3(256), 0, st, ids, dists, count);
  return hipGetLastError();
}

}  // namespace vsg
