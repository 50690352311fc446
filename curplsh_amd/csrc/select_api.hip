// The reference's internal selection and broadcast functions as device-pointer
// entry points (the C++ wrappers with the reference's Tensor signatures live in
// curplsh_amd/host/curplsh.cpp):
//   l2Select            include/internal/select.h:29-34, src/internal/select.cu:189-224
//   blockSelect (x2)    include/internal/select.h:8-24, src/internal/select.cu:55-96
//   broadcastAlongRowsSum  include/internal/broadcast.h:7-8, src/internal/broadcast.cu:15-74
// Selection runs on the search path's select kernels: every (key, label) pair of
// a row becomes one 64-bit word (ordered key bits, label), so the k smallest
// words are the k smallest keys with ties broken by the lower label -- the
// order the reference leaves unspecified (merge_network.cuh:89-92). The k
// largest are the k smallest of the negated keys.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "common.hpp"
#include "device.hpp"
#include "kernels.hpp"

namespace rplsh {
namespace {

constexpr int kMaxSelectK = 4096;

// words[r][i] = (key, label): key = (norms ? norms[i] + p[r][i] : p[r][i]),
// negated when `negate`; label = labels ? labels[r][i] : i.
__global__ void pack_rows_kernel(const float* __restrict__ p, int64_t ldp, const float* __restrict__ norms,
                                 const int32_t* __restrict__ labels, int64_t rows, int64_t n, int negate,
                                 uint64_t* __restrict__ words) {
  const int64_t total = rows * n;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / n, i = e - r * n;
    float key = p[r * ldp + i];
    if (norms) key = norms[i] + key;
    if (negate) key = -key;
    words[e] = pack_cand(key, labels ? labels[e] : (int32_t)i);
  }
}

__global__ void unpack_rows_kernel(const uint64_t* __restrict__ words, int64_t total, int negate,
                                   float* __restrict__ out_k, int32_t* __restrict__ out_v) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t w = words[e];
    if (w == kEmptyCand) {
      out_k[e] = negate ? -FLT_MAX : FLT_MAX;
      out_v[e] = -1;
    } else {
      const float key = cand_key(w);
      out_k[e] = negate ? -key : key;
      out_v[e] = cand_idx(w);
    }
  }
}

__global__ void broadcast_rows_sum_kernel(const float* __restrict__ v, float* __restrict__ m, int64_t rows,
                                          int64_t cols) {
  const int64_t total = rows * cols;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x)
    m[e] += v[e / cols];
}

inline unsigned grid_of(int64_t work) {
  int64_t g = (work + 255) / 256;
  if (g > 65536) g = 65536;
  return (unsigned)(g < 1 ? 1 : g);
}

// Row-wise top-k of packed words (the search's top-K select), unpacked.
void select_rows(const float* p, int64_t ldp, const float* norms, const int32_t* labels, int64_t rows,
                 int64_t n, int k, int negate, float* out_k, int32_t* out_v, hipStream_t st) {
  RPLSH_REQUIRE(rows >= 0 && n >= 0 && k > 0, "bad sizes");
  RPLSH_REQUIRE(k <= kMaxSelectK, "k limited to %d", kMaxSelectK);
  RPLSH_REQUIRE(rows < (1ll << 31) && n < (1ll << 31), "sizes limited to 2^31");
  if (rows == 0) return;
  uint64_t *words = nullptr, *top = nullptr;
  int* counts = nullptr;
  float *tau = nullptr, *bound = nullptr;
  auto release = [&] {
    for (void* q : {(void*)words, (void*)top, (void*)counts, (void*)tau, (void*)bound})
      if (q) (void)hipFree(q);
  };
  try {
    RPLSH_HIP(hipMalloc(&words, (size_t)(rows * (n > 0 ? n : 1)) * sizeof(uint64_t)));
    RPLSH_HIP(hipMalloc(&top, (size_t)rows * k * sizeof(uint64_t)));
    RPLSH_HIP(hipMalloc(&counts, (size_t)rows * sizeof(int)));
    RPLSH_HIP(hipMalloc(&tau, (size_t)rows * sizeof(float)));
    RPLSH_HIP(hipMalloc(&bound, (size_t)rows * sizeof(float)));
    if (n > 0) {
      pack_rows_kernel<<<grid_of(rows * n), 256, 0, st>>>(p, ldp, norms, labels, rows, n, negate, words);
      RPLSH_HIP(hipGetLastError());
    }
    launch_fill_i32(counts, rows, (int)n, st);
    launch_fill_f32(tau, rows, INFINITY, st);
    SelectArgs sa{};
    sa.cand = words;
    sa.counts = counts;
    sa.cap = (int)(n > 0 ? n : 1);
    sa.nq = (int)rows;
    sa.K = k;
    sa.topk_mode = 1;
    sa.tau_in = tau;
    sa.topk_out = top;
    sa.bound_out = bound;
    launch_select(sa, st);
    unpack_rows_kernel<<<grid_of(rows * k), 256, 0, st>>>(top, rows * k, negate, out_k, out_v);
    RPLSH_HIP(hipGetLastError());
    RPLSH_HIP(hipStreamSynchronize(st));
  } catch (...) {
    (void)hipStreamSynchronize(st);
    release();
    throw;
  }
  release();
}

}  // namespace
}  // namespace rplsh

using namespace rplsh;

extern "C" {

rplsh_status rplsh_l2_select_dev(const float* products, int64_t rows, int64_t n, int64_t ldp,
                                 const float* base_norms, int k, float* out_dist, int32_t* out_idx,
                                 void* stream) {
  return guard([&] {
    RPLSH_REQUIRE(products || rows * n == 0, "products is NULL");
    RPLSH_REQUIRE(base_norms || n == 0, "base_norms is NULL");
    RPLSH_REQUIRE(out_dist && out_idx, "NULL output");
    RPLSH_REQUIRE(ldp >= n, "row stride < n");
    select_rows(products, ldp, base_norms, nullptr, rows, n, k, 0, out_dist, out_idx, (hipStream_t)stream);
  });
}

rplsh_status rplsh_block_select_dev(const float* in, const int32_t* labels, int64_t rows, int64_t n, int k,
                                    int select_max, float* out_k, int32_t* out_v, void* stream) {
  return guard([&] {
    RPLSH_REQUIRE(in || rows * n == 0, "input is NULL");
    RPLSH_REQUIRE(out_k && out_v, "NULL output");
    select_rows(in, n, nullptr, labels, rows, n, k, select_max ? 1 : 0, out_k, out_v, (hipStream_t)stream);
  });
}

rplsh_status rplsh_broadcast_rows_sum_dev(const float* v, float* m, int64_t rows, int64_t cols, void* stream) {
  return guard([&] {
    RPLSH_REQUIRE(rows >= 0 && cols >= 0, "bad sizes");
    if (rows * cols == 0) return;
    RPLSH_REQUIRE(v && m, "NULL pointer");
    broadcast_rows_sum_kernel<<<grid_of(rows * cols), 256, 0, (hipStream_t)stream>>>(v, m, rows, cols);
    RPLSH_HIP(hipGetLastError());
  });
}

}  // extern "C"
