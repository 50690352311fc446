// Measured MFMA peak of this device for the three element kinds the filter
// uses: every wave keeps four independent 32x32 accumulators busy with
// back-to-back MFMAs (no memory traffic inside the loop), over a grid of 8
// waves per CU. Operands are random finite values (`random`) or a constant
// pattern: this part's MFMA throughput is power-limited, so random operands
// (what real data feeds the filter) give the lower, realistic ceiling.
//
// Built into libcurplsh_probe.so for the benchmark only; not part of the
// library's C ABI:
//   rplsh_status rplsh_bench_mfma_peak(int kind, int random_operands, double* tflops);
// kind 0 = bf16 (32x32x16), 1 = fp8 e4m3 (32x32x64 scaled), 2 = fp32 (32x32x2).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"
#include "rplsh.h"

namespace rplsh {
namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef int i32x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef int i32x4 __attribute__((ext_vector_type(4)));

constexpr int kChains = 4;  // independent accumulators per wave
constexpr int kOps = 8;     // operand registers cycled through
// Launches of ~0.6 ms each (BF16 / FP8 at the random-operand rate; FP32 about
// as long) before and inside the timed region.
constexpr int kWarmLaunches = 40;
constexpr int kTimedLaunches = 40;

__device__ inline uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// 32 random bits shaped into a finite value of the element kind (two bf16,
// four e4m3 or one fp32 per word; magnitudes 2^-4..1, random sign/mantissa).
template <int KIND>
__device__ inline int shape_word(uint32_t r, bool random) {
  if (!random) return KIND == 0 ? 0x3f803f80 : (KIND == 1 ? 0x38383838 : 0x3f800000);
  if (KIND == 0) {  // bf16 pair
    uint32_t w = 0;
    for (int j = 0; j < 2; ++j) {
      const uint32_t b = r >> (16 * j);
      const uint32_t e = 123u + (b % 5u), m = (b >> 3) & 0x7fu, sg = (b >> 15) & 1u;
      w |= ((sg << 15) | (e << 7) | m) << (16 * j);
    }
    return (int)w;
  }
  if (KIND == 1) {  // four e4m3: exponent field 3..7, never the NaN code
    uint32_t w = 0;
    for (int j = 0; j < 4; ++j) {
      const uint32_t b = (r >> (8 * j)) & 0xffu;
      const uint32_t e = 3u + (b % 5u), m = (b >> 3) & 7u, sg = (b >> 7) & 1u;
      w |= ((sg << 7) | (e << 3) | m) << (8 * j);
    }
    return (int)w;
  }
  const uint32_t e = 123u + (r % 5u);  // fp32
  return (int)(((r >> 31) << 31) | (e << 23) | ((r >> 3) & 0x7fffffu));
}

template <int KIND>
__global__ __launch_bounds__(256) void mfma_peak_kernel(int iters, int random, float* out) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  int a[kOps][8], b[kOps][8];
#pragma unroll
  for (int o = 0; o < kOps; ++o)
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      a[o][i] = shape_word<KIND>(mix32(tid * 131u + o * 17u + i), random != 0);
      b[o][i] = shape_word<KIND>(mix32(tid * 977u + o * 29u + i + 7u), random != 0);
    }
  f32x16 acc[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c)
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[c][i] = 0.0f;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int o = 0; o < kOps; ++o) {
#pragma unroll
      for (int c = 0; c < kChains; ++c) {
        const int ao = (o + c) % kOps;
        if (KIND == 0) {
          const i32x4 av = {a[ao][0], a[ao][1], a[ao][2], a[ao][3]};
          const i32x4 bv = {b[o][0], b[o][1], b[o][2], b[o][3]};
          acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, av),
                                                          __builtin_bit_cast(bf16x8, bv), acc[c], 0, 0, 0);
        } else if (KIND == 1) {
          const i32x8 av = {a[ao][0], a[ao][1], a[ao][2], a[ao][3], a[ao][4], a[ao][5], a[ao][6], a[ao][7]};
          const i32x8 bv = {b[o][0], b[o][1], b[o][2], b[o][3], b[o][4], b[o][5], b[o][6], b[o][7]};
          acc[c] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, bv, acc[c], 0, 0, 0, 127, 0, 127);
        } else {
          acc[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(__int_as_float(a[ao][0]), __int_as_float(b[o][0]),
                                                         acc[c], 0, 0, 0);
        }
      }
    }
  }
  float s = 0.0f;
#pragma unroll
  for (int c = 0; c < kChains; ++c)
#pragma unroll
    for (int i = 0; i < 16; ++i) s += acc[c][i];
  out[tid] = s;
}

// FLOPs of one MFMA of the kind.
double mfma_flops(int kind) { return kind == 0 ? 2.0 * 32 * 32 * 16 : (kind == 1 ? 2.0 * 32 * 32 * 64 : 2.0 * 32 * 32 * 2); }

}  // namespace
}  // namespace rplsh

using namespace rplsh;

extern "C" rplsh_status rplsh_bench_mfma_peak(int kind, int random_operands, double* tflops) {
  return guard([&] {
    RPLSH_REQUIRE(kind >= 0 && kind <= 2, "kind: 0 bf16, 1 fp8 e4m3, 2 fp32");
    RPLSH_REQUIRE(tflops, "NULL output");
    int dev = 0, cus = 0;
    RPLSH_HIP(hipGetDevice(&dev));
    RPLSH_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int blocks = cus * 2;  // two 4-wave workgroups per CU: two waves per SIMD
    const int iters = kind == 2 ? 512 : 1024;
    float* out = nullptr;
    RPLSH_HIP(hipMalloc(&out, (size_t)blocks * 256 * sizeof(float)));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.0f;
    try {
      RPLSH_HIP(hipEventCreate(&e0));
      RPLSH_HIP(hipEventCreate(&e1));
      auto launch = [&](int n_iters) {
        if (kind == 0) mfma_peak_kernel<0><<<blocks, 256>>>(n_iters, random_operands, out);
        else if (kind == 1) mfma_peak_kernel<1><<<blocks, 256>>>(n_iters, random_operands, out);
        else mfma_peak_kernel<2><<<blocks, 256>>>(n_iters, random_operands, out);
        RPLSH_HIP(hipGetLastError());
      };
      // The clock settles only after ~25 ms of sustained MFMA work (a first
      // 25 ms window runs ~8 % below the later ones): warm up that long, then
      // time as long again.
      for (int w = 0; w < kWarmLaunches; ++w) launch(iters);
      RPLSH_HIP(hipEventRecord(e0, nullptr));
      for (int r = 0; r < kTimedLaunches; ++r) launch(iters);
      RPLSH_HIP(hipEventRecord(e1, nullptr));
      RPLSH_HIP(hipEventSynchronize(e1));
      RPLSH_HIP(hipEventElapsedTime(&ms, e0, e1));
    } catch (...) {
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      (void)hipFree(out);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    RPLSH_HIP(hipFree(out));
    const double mfmas = (double)kTimedLaunches * blocks * 4.0 * iters * kOps * kChains;  // 4 waves per block
    *tflops = mfmas * mfma_flops(kind) / (ms / 1e3) / 1e12;
  });
}
