// Element encoders shared by the add-time and query-time kernels (norm.hip,
// filter_packed.hip) and the persistent serving kernel (persist.hip): BF16
// round-to-nearest-even, OCP FP8 e4m3 (FN) round-to-nearest-even with
// saturation, and the per-row power-of-two e4m3 scale.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rplsh {

// bf16 round-to-nearest-even of one float (NaN stays a quiet NaN).
__device__ inline uint16_t bf16_rne(float v) {
  uint32_t u = __float_as_uint(v);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// OCP FP8 e4m3 (FN: no infinities, max 448) round-to-nearest-even with
// saturation. Scaling by powers of two is exact, so rintf on the scaled
// magnitude is the RNE step of the target binade.
__device__ inline uint8_t e4m3_rne_sat(float v) {
  if (isnan(v)) return 0x7f;
  const uint8_t sign = signbit(v) ? 0x80 : 0x00;
  const float a = fabsf(v);
  if (a >= 448.0f) return sign | 0x7e;
  if (a < 0.015625f) {  // below 2^-6: subnormal quantum 2^-9; m == 8 is 2^-6
    const int m = (int)rintf(a * 512.0f);
    return sign | (uint8_t)m;
  }
  int e2;
  frexpf(a, &e2);  // a = f * 2^e2, f in [0.5, 1): binade 2^(e2-1)
  int eb = e2 - 1;
  int r = (int)rintf(ldexpf(a, 3 - eb));  // in [8, 16]
  if (r == 16) {
    r = 8;
    ++eb;
  }
  int code = ((eb + 7) << 3) | (r - 8);
  if (code > 0x7e) code = 0x7e;
  return sign | (uint8_t)code;
}

// Smallest power of two >= max|x| / 448 (1 for an all-zero row), so that
// max|x| / s <= 448 and the scale is exact as an E8M0 exponent.
__device__ inline float e4m3_row_scale(float max_abs) {
  float s = 1.0f;
  if (max_abs > 0.0f) {
    const float s0 = max_abs / 448.0f;
    int e;
    const float f = frexpf(s0, &e);  // s0 = f * 2^e, f in [0.5, 1)
    s = f == 0.5f ? s0 : ldexpf(1.0f, e);
  }
  return s;
}

}  // namespace rplsh
