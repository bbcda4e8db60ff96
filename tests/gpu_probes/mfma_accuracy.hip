// Accuracy of the MFMA instructions the certified shortlist uses, against FP64 sums.
//
// One wave runs chains of 1, 4, 12 and 24 instructions (a dot product of 32 x chain dims) of
// v_mfma_f32_16x16x32_fp8_fp8 and of v_mfma_f32_16x16x32_bf16, and chains of 1, 3 and 6 (128,
// 384 and 768 dims) of v_mfma_f32_16x16x128_f8f6f4 with e4m3 operands, on seeded random
// operands (e4m3: every finite code; bf16: exponents 2^-27 .. 2^28, random sign and mantissa),
// with mixed signs and with one sign. Lane l holds A[l % 16][8 (l / 16) .. +7] and B[column
// l % 16][same K] of each 32-wide k-step, as in the shortlist's fragment copies; a 128-dim
// instruction takes a lane's pieces of four consecutive k-steps side by side, as the shortlist's
// kernels do. For each of the 16 x 16 outputs it reports |mfma - exact| / sum |products|; the
// JSON line gives the largest ratio per instruction and chain, which the GPU tests compare with
// the error bound's terms.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef int i32x8 __attribute__((ext_vector_type(8)));
typedef long i64x4 __attribute__((ext_vector_type(4)));

__global__ void fp8_chain(const uint8_t* A, const uint8_t* B, float* C, int chain) {
  const int l = threadIdx.x;
  f32x4 c = {0.f, 0.f, 0.f, 0.f};
  for (int s = 0; s < chain; ++s) {
    long a, b;
    memcpy(&a, A + s * 512 + (l % 16) * 32 + 8 * (l / 16), 8);
    memcpy(&b, B + s * 512 + (l % 16) * 32 + 8 * (l / 16), 8);
    c = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a, b, c, 0, 0, 0);
  }
  for (int i = 0; i < 4; ++i) C[(4 * (l / 16) + i) * 16 + (l % 16)] = c[i];
}

// chain instructions of 128 dims: k-steps 4t .. 4t + 3 of the 32-dim layout in one operand
__global__ void fp8x128_chain(const uint8_t* A, const uint8_t* B, float* C, int chain) {
  const int l = threadIdx.x;
  f32x4 c = {0.f, 0.f, 0.f, 0.f};
  for (int t = 0; t < chain; ++t) {
    i64x4 a, b;
    for (int s = 0; s < 4; ++s) {
      long x, y;
      memcpy(&x, A + (4 * t + s) * 512 + (l % 16) * 32 + 8 * (l / 16), 8);
      memcpy(&y, B + (4 * t + s) * 512 + (l % 16) * 32 + 8 * (l / 16), 8);
      a[s] = x;
      b[s] = y;
    }
    c = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(__builtin_bit_cast(i32x8, a), __builtin_bit_cast(i32x8, b),
                                                         c, 0, 0, 0, 0, 0, 0);
  }
  for (int i = 0; i < 4; ++i) C[(4 * (l / 16) + i) * 16 + (l % 16)] = c[i];
}

__global__ void bf16_chain(const __bf16* A, const __bf16* B, float* C, int chain) {
  const int l = threadIdx.x;
  f32x4 c = {0.f, 0.f, 0.f, 0.f};
  for (int s = 0; s < chain; ++s) {
    bf16x8 a, b;
    for (int j = 0; j < 8; ++j) {
      a[j] = A[s * 512 + (l % 16) * 32 + 8 * (l / 16) + j];
      b[j] = B[s * 512 + (l % 16) * 32 + 8 * (l / 16) + j];
    }
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
  for (int i = 0; i < 4; ++i) C[(4 * (l / 16) + i) * 16 + (l % 16)] = c[i];
}

static double e4m3_value(uint8_t v) {
  const int s = v >> 7, e = (v >> 3) & 15, m = v & 7;
  const double x = e == 0 ? (m / 8.0) * std::ldexp(1.0, -6) : (1 + m / 8.0) * std::ldexp(1.0, e - 7);
  return s ? -x : x;
}

static double bf16_value(uint16_t h) {
  const uint32_t u = uint32_t(h) << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

#define CHECK(x)                                                    \
  do {                                                              \
    if ((x) != hipSuccess) {                                        \
      fprintf(stderr, "%s failed\n", #x);                           \
      return 1;                                                     \
    }                                                               \
  } while (0)

int main() {
  std::mt19937 rng(2024);
  printf("{");
  bool first = true;
  const std::vector<int> chains32 = {1, 4, 12, 24}, chains128 = {1, 3, 6};
  for (int prec = 0; prec < 3; ++prec) {  // 0: fp8 16x16x32, 1: bf16, 2: fp8 16x16x128
    for (int chain : prec == 2 ? chains128 : chains32) {
      const int steps = prec == 2 ? 4 * chain : chain;  // 32-dim k-steps
      double worst = 0;
      for (int same_sign = 0; same_sign < 2; ++same_sign) {
        const int n = steps * 512;
        std::vector<uint8_t> a8(n), b8(n);
        std::vector<uint16_t> a16(n), b16(n);
        for (int i = 0; i < n; ++i) {
          uint8_t v;
          do v = uint8_t(rng() & 0xFF); while ((v & 0x7F) == 0x7F);  // no NaN code
          a8[i] = same_sign ? (v & 0x7F) : v;
          do v = uint8_t(rng() & 0xFF); while ((v & 0x7F) == 0x7F);
          b8[i] = same_sign ? (v & 0x7F) : v;
          const uint16_t sa = same_sign ? 0 : uint16_t(rng() & 0x8000), sb = same_sign ? 0 : uint16_t(rng() & 0x8000);
          a16[i] = uint16_t(sa | (uint16_t(100 + rng() % 56) << 7) | (rng() & 0x7F));
          b16[i] = uint16_t(sb | (uint16_t(100 + rng() % 56) << 7) | (rng() & 0x7F));
        }
        void *dA, *dB;
        float* dC;
        const bool fp8 = prec != 1;
        const size_t bytes = fp8 ? size_t(n) : size_t(n) * 2;
        CHECK(hipMalloc(&dA, bytes));
        CHECK(hipMalloc(&dB, bytes));
        CHECK(hipMalloc(&dC, 256 * sizeof(float)));
        CHECK(hipMemcpy(dA, fp8 ? (void*)a8.data() : (void*)a16.data(), bytes, hipMemcpyHostToDevice));
        CHECK(hipMemcpy(dB, fp8 ? (void*)b8.data() : (void*)b16.data(), bytes, hipMemcpyHostToDevice));
        if (prec == 2)
          hipLaunchKernelGGL(fp8x128_chain, dim3(1), dim3(64), 0, 0, static_cast<const uint8_t*>(dA),
                             static_cast<const uint8_t*>(dB), dC, chain);
        else if (prec == 0)
          hipLaunchKernelGGL(fp8_chain, dim3(1), dim3(64), 0, 0, static_cast<const uint8_t*>(dA),
                             static_cast<const uint8_t*>(dB), dC, chain);
        else
          hipLaunchKernelGGL(bf16_chain, dim3(1), dim3(64), 0, 0, static_cast<const __bf16*>(dA),
                             static_cast<const __bf16*>(dB), dC, chain);
        CHECK(hipGetLastError());
        float c[256];
        CHECK(hipMemcpy(c, dC, sizeof c, hipMemcpyDeviceToHost));
        CHECK(hipFree(dA));
        CHECK(hipFree(dB));
        CHECK(hipFree(dC));
        for (int m = 0; m < 16; ++m)
          for (int col = 0; col < 16; ++col) {
            double exact = 0, sum_abs = 0;
            for (int s = 0; s < steps; ++s)
              for (int k = 0; k < 32; ++k) {
                const int ia = s * 512 + m * 32 + k, ib = s * 512 + col * 32 + k;
                const double p = fp8 ? e4m3_value(a8[ia]) * e4m3_value(b8[ib])
                                           : bf16_value(a16[ia]) * bf16_value(b16[ib]);
                exact += p;
                sum_abs += std::fabs(p);
              }
            worst = std::fmax(worst, std::fabs(c[m * 16 + col] - exact) / sum_abs);
          }
      }
      printf("%s\"%s_%d\": %.6e", first ? "" : ", ", prec == 0 ? "fp8" : prec == 1 ? "bf16" : "fp8x128", chain, worst);
      first = false;
    }
  }
  printf("}\n");
  return 0;
}
