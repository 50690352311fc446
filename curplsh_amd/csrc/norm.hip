// L2 norms and element encoders.
//
// l2norm_kernel follows computeL2Norm (src/internal/norm.cu:15-202) in what it
// computes -- the (squared) L2 norm of every row -- but not in how: one lane
// owns one row and runs the canonical chain n = fmaf(x[j], x[j], n) over
// j = 0..d-1, while each wave stages its 64 rows through LDS in 32-column
// chunks so global reads stay coalesced (consecutive lanes read consecutive
// 16-byte pieces of a row). Any n and d are handled (reference defects D4-D6:
// partial warps, under-sized shared memory and out-of-bounds rows do not
// apply).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

#include "common.hpp"
#include "device.hpp"
#include "encode.hpp"
#include "kernels.hpp"

namespace rplsh {

namespace {

constexpr int kNormCols = 32;

template <bool kVec4>
__global__ __launch_bounds__(256) void l2norm_kernel(const float* __restrict__ x, int64_t n, int d,
                                                     int64_t ldx, float* __restrict__ out,
                                                     int squared) {
  __shared__ float tile[4][64][kNormCols + 1];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t row0 = ((int64_t)blockIdx.x * 4 + wave) * 64;
  float s = 0.0f;
  for (int c0 = 0; c0 < d; c0 += kNormCols) {
    if (kVec4) {
      // 64 rows x 8 float4 = 512 float4 per wave, 8 per lane.
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int u = t * 64 + lane;
        const int r = u >> 3, c4 = (u & 7) * 4;
        const int64_t row = row0 + r;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (row < n && c0 + c4 < d) v = *reinterpret_cast<const float4*>(x + row * ldx + c0 + c4);
        tile[wave][r][c4 + 0] = v.x;
        tile[wave][r][c4 + 1] = v.y;
        tile[wave][r][c4 + 2] = v.z;
        tile[wave][r][c4 + 3] = v.w;
      }
    } else {
#pragma unroll 4
      for (int t = 0; t < 32; ++t) {
        const int u = t * 64 + lane;
        const int r = u >> 5, c = u & 31;
        const int64_t row = row0 + r;
        tile[wave][r][c] = (row < n && c0 + c < d) ? x[row * ldx + c0 + c] : 0.0f;
      }
    }
    __syncthreads();
    const int cend = min(kNormCols, d - c0);
    for (int j = 0; j < cend; ++j) {
      const float v = tile[wave][lane][j];
      s = fmaf(v, v, s);
    }
    __syncthreads();
  }
  const int64_t row = row0 + lane;
  if (row < n) out[row] = squared ? s : sqrtf(s);
}

__global__ void encode_bf16_kernel(const float* __restrict__ x, int64_t n, int d, int64_t ldx,
                                   int ld, uint16_t* __restrict__ out) {
  const int64_t total = n * ld;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / ld;
    const int j = (int)(e - i * ld);
    out[e] = j < d ? bf16_rne(x[i * ldx + j]) : (uint16_t)0;
  }
}

// One wave per row: power-of-two scale s >= max|x| / 448, then codes of x / s.
__global__ __launch_bounds__(256) void encode_e4m3_kernel(const float* __restrict__ x, int64_t n,
                                                          int d, int64_t ldx, int ld,
                                                          uint8_t* __restrict__ out,
                                                          float* __restrict__ scales) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * 4 + wave;
  if (row >= n) return;
  const float* xr = x + row * ldx;
  float mx = 0.0f;
  for (int j = lane; j < d; j += 64) mx = fmaxf(mx, fabsf(xr[j]));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
  const float s = e4m3_row_scale(mx);
  for (int j = lane; j < ld; j += 64) out[row * ld + j] = j < d ? e4m3_rne_sat(xr[j] / s) : 0;
  if (lane == 0) scales[row] = s;
}

__global__ void fill_f32_kernel(float* __restrict__ p, int64_t n, float v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void fill_i32_kernel(int* __restrict__ p, int64_t n, int v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Row-major copy with zero padding: out[i][0..ld) <- x[i][0..d).
__global__ void pad_rows_kernel(const float* __restrict__ x, int64_t n, int d, int64_t ldx, int ld,
                                float* __restrict__ out) {
  const int64_t total = n * ld;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / ld;
    const int j = (int)(e - i * ld);
    out[e] = j < d ? x[i * ldx + j] : 0.0f;
  }
}

__global__ __launch_bounds__(256) void max_reduce_kernel(const float* __restrict__ v, int64_t n,
                                                         float* __restrict__ out) {
  float m = 0.0f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    m = fmaxf(m, v[i]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  __shared__ float part[4];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = fmaxf(fmaxf(part[0], part[1]), fmaxf(part[2], part[3]));
    // Non-negative floats order like their bit patterns.
    atomicMax(reinterpret_cast<unsigned int*>(out), __float_as_uint(m));
  }
}

inline unsigned grid_for(int64_t work, int block) {
  int64_t g = (work + block - 1) / block;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (unsigned)g;
}

}  // namespace

void launch_l2norm(const float* x, int64_t n, int d, int64_t ldx, float* out, int squared,
                   hipStream_t st) {
  if (n <= 0) return;
  const unsigned grid = (unsigned)div_up(n, 256);
  const bool vec = (d % 4 == 0) && (ldx % 4 == 0) && ((uintptr_t)x % 16 == 0);
  if (vec) {
    l2norm_kernel<true><<<grid, 256, 0, st>>>(x, n, d, ldx, out, squared);
  } else {
    l2norm_kernel<false><<<grid, 256, 0, st>>>(x, n, d, ldx, out, squared);
  }
  RPLSH_HIP(hipGetLastError());
}

void launch_encode_bf16(const float* x, int64_t n, int d, int64_t ldx, int ld, uint16_t* out,
                        hipStream_t st) {
  if (n <= 0) return;
  encode_bf16_kernel<<<grid_for(n * ld, 256), 256, 0, st>>>(x, n, d, ldx, ld, out);
  RPLSH_HIP(hipGetLastError());
}

void launch_encode_e4m3(const float* x, int64_t n, int d, int64_t ldx, int ld, uint8_t* out,
                        float* scales, hipStream_t st) {
  if (n <= 0) return;
  encode_e4m3_kernel<<<(unsigned)div_up(n, 4), 256, 0, st>>>(x, n, d, ldx, ld, out, scales);
  RPLSH_HIP(hipGetLastError());
}

void launch_fill_f32(float* p, int64_t n, float v, hipStream_t st) {
  if (n <= 0) return;
  fill_f32_kernel<<<grid_for(n, 256), 256, 0, st>>>(p, n, v);
  RPLSH_HIP(hipGetLastError());
}

void launch_fill_i32(int* p, int64_t n, int v, hipStream_t st) {
  if (n <= 0) return;
  fill_i32_kernel<<<grid_for(n, 256), 256, 0, st>>>(p, n, v);
  RPLSH_HIP(hipGetLastError());
}

void launch_pad_rows(const float* x, int64_t n, int d, int64_t ldx, int ld, float* out,
                     hipStream_t st) {
  if (n <= 0) return;
  pad_rows_kernel<<<grid_for(n * ld, 256), 256, 0, st>>>(x, n, d, ldx, ld, out);
  RPLSH_HIP(hipGetLastError());
}

void launch_max_reduce(const float* v, int64_t n, float* out, hipStream_t st) {
  if (n <= 0) return;
  max_reduce_kernel<<<grid_for(n, 256) > 1024 ? 1024 : grid_for(n, 256), 256, 0, st>>>(v, n, out);
  RPLSH_HIP(hipGetLastError());
}

}  // namespace rplsh

using namespace rplsh;

extern "C" {

rplsh_status rplsh_l2norm_dev(const float* x, int64_t n, int d, float* out, int squared,
                              void* stream) {
  return guard([&] {
    RPLSH_REQUIRE(n >= 0 && d > 0, "bad shape n=%lld d=%d", (long long)n, d);
    RPLSH_REQUIRE(n == 0 || (x && out), "NULL pointer");
    launch_l2norm(x, n, d, d, out, squared, (hipStream_t)stream);
  });
}

rplsh_status rplsh_encode_bf16_dev(const float* x, int64_t n, int d, int ld, uint16_t* out,
                                   void* stream) {
  return guard([&] {
    RPLSH_REQUIRE(n >= 0 && d > 0 && ld >= d, "bad shape");
    RPLSH_REQUIRE(n == 0 || (x && out), "NULL pointer");
    launch_encode_bf16(x, n, d, d, ld, out, (hipStream_t)stream);
  });
}

rplsh_status rplsh_encode_e4m3_dev(const float* x, int64_t n, int d, int ld, uint8_t* out,
                                   float* scales, void* stream) {
  return guard([&] {
    RPLSH_REQUIRE(n >= 0 && d > 0 && ld >= d, "bad shape");
    RPLSH_REQUIRE(n == 0 || (x && out && scales), "NULL pointer");
    launch_encode_e4m3(x, n, d, d, ld, out, scales, (hipStream_t)stream);
  });
}

}  // extern "C"
