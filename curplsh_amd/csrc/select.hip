// Candidate selection and exact re-ranking.
//
// select_kernel: one workgroup per query. Finds the K-th smallest candidate
// word (key, index) by an 8-pass radix select over the 64-bit words (8 bits
// per pass, histogram in LDS), then either reports its key (the threshold of
// the next filter level) or gathers the K smallest words and bitonic-sorts
// them. Because words are unique (indices are), the K smallest words are
// exactly the K smallest candidates in (key, index) order -- the order the
// reference's BlockSelect leaves unspecified for ties (merge_network.cuh:89-92)
// and the search defines as "lower index first". Role of the reference's
// kernelL2SelectK / kernelBlockSelect (select.cu:55-96,189-224) and of the
// cross-tile merge (search.cu:203-208).
//
// rerank_kernel: one workgroup per query. Recomputes every candidate's
// relative distance in the canonical FP32 order
//     dot = fmaf chain over j,  rel = fmaf(-2, dot, ||x||^2)
// (rows gathered cooperatively through LDS, score_rows_staged), sorts by
// (rel, index), keeps k, and writes  D = rel + ||q||^2
// (broadcast.cu:15-38 folded into the write; index offsets of
// search.cu:21-42 are not needed: candidates carry global indices).
// Plain path: certifies that no non-candidate can enter the top k. Precision
// forms: re-orders the boundary band around the k-th distance by FP64
// distance, reports the band members' FP64 distances (rounded to FP32) and
// returns the k results ascending by reported distance, then index.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "common.hpp"
#include "device.hpp"
#include "kernels.hpp"
#include "select_body.hpp"

namespace rplsh {

namespace {

__global__ __launch_bounds__(kSelBlock) void select_kernel(SelectArgs a) {
  __shared__ uint64_t stage[kSelLds];
  __shared__ uint64_t outl[kMaxSortK];
  __shared__ int hist[256];
  __shared__ int sh[4];
  const int q = blockIdx.x;
  if (q >= a.nq || (a.nq_dev && q >= *a.nq_dev)) return;
  const int count = a.counts[q];
  const int n = count < a.cap ? count : a.cap;
  int st = 0;
  if (count > a.cap) st |= kStatusOverflow;
  const uint64_t* gsrc = a.cand + (int64_t)q * a.cap;
  const uint64_t* src = gsrc;
  if (n <= kSelLds) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) stage[i] = gsrc[i];
    __syncthreads();
    src = stage;
  }
  const int K = a.K;
  if (n < K) st |= kStatusShort;

  if (!a.topk_mode) {
    // Too few survivors for rank K: keep the previous threshold, which every
    // survivor already met (an infinite one would let every row of the next
    // level through).
    float tau = a.tau_in[q];
    if (n >= K && K > 0) tau = cand_key(radix_kth(src, n, K, hist, sh));
    if (threadIdx.x == 0) {
      a.tau_out[q] = tau;
      if (st && a.status) atomicOr(a.status + q, st);
    }
    return;
  }

  // Top-K mode: the K smallest words, sorted.
  const int take = n < K ? n : K;
  uint64_t thr = ~0ull;
  if (n > K) thr = radix_kth(src, n, K, hist, sh);
  if (threadIdx.x == 0) sh[2] = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const uint64_t v = src[i];
    if (n <= K || v <= thr) {
      const int p = atomicAdd(&sh[2], 1);
      if (p < kMaxSortK) outl[p] = v;
    }
  }
  __syncthreads();
  const int P = pow2_at_least(take > 0 ? take : 1);
  for (int i = take + threadIdx.x; i < P; i += blockDim.x) outl[i] = kEmptyCand;
  __syncthreads();
  bitonic_sort_u64(outl, P);
  uint64_t* dst = a.topk_out + (int64_t)q * K;
  for (int i = threadIdx.x; i < K; i += blockDim.x) dst[i] = i < take ? outl[i] : kEmptyCand;
  if (threadIdx.x == 0) {
    // Every row outside the list has key >= bound: either a survivor that lost
    // the select (only when there were more than K), or a row the filter
    // rejected (key > tau_in; +inf when the level kept every row).
    a.bound_out[q] = count > K && K > 0 ? cand_key(outl[K - 1]) : a.tau_in[q];
    if (st && a.status) atomicOr(a.status + q, st);
  }
}


__global__ __launch_bounds__(64 * kWaveSelQ) void select_wave_kernel(SelectArgs a) {
  __shared__ __attribute__((aligned(16))) int hist_all[kWaveSelQ][256];
  __shared__ uint64_t out_all[kWaveSelQ][kWaveSelK];
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int q = blockIdx.x * kWaveSelQ + w;
  if (q >= a.nq || (a.nq_dev && q >= *a.nq_dev)) return;
  select_wave_query(a, q, hist_all[w], out_all[w]);
}

template <int NP>
__global__ __launch_bounds__(kSelBlock) void rerank_kernel(RerankArgs a) {
  extern __shared__ __attribute__((aligned(16))) uint64_t rerank_lds[];
  __shared__ float sq_fixed[NP < 0 ? 4 : 256];
  __shared__ float sxn[kSelWaves * 64];
  __shared__ int sh[4];
  const int q = blockIdx.x;
  if (q >= a.nq || (a.nq_dev && q >= *a.nq_dev)) return;
  rerank_query<NP>(a, q, rerank_lds, sq_fixed, sxn, sh);
}

__global__ void collect_failed_kernel(const int* __restrict__ status, int nq,
                                      const int* __restrict__ nq_dev, int* __restrict__ list,
                                      int* __restrict__ n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int live = nq_dev ? *nq_dev : nq;
  if (i < nq && i < live && status[i] != 0) {
    const int p = atomicAdd(n, 1);
    list[p] = i;
  }
}

__global__ void sum_counts_kernel(const int* __restrict__ counts, int nq, int cap,
                                  unsigned long long* __restrict__ out) {
  unsigned long long s = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x)
    s += (unsigned long long)counts[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0 && s) atomicAdd(out, s);
}

__global__ void gather_rows_kernel(const float* __restrict__ src, int64_t ld, const int* __restrict__ rows,
                                   int nrows, int d, float* __restrict__ dst) {
  const int64_t total = (int64_t)nrows * d;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(e / d);
    const int j = (int)(e - (int64_t)r * d);
    dst[e] = src[(int64_t)rows[r] * ld + j];
  }
}

__global__ void scatter_results_kernel(const int32_t* __restrict__ si, const float* __restrict__ sd,
                                       const int* __restrict__ rows, int nrows, int k,
                                       int32_t* __restrict__ di, float* __restrict__ dd) {
  const int64_t total = (int64_t)nrows * k;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(e / k);
    const int j = (int)(e - (int64_t)r * k);
    di[(int64_t)rows[r] * k + j] = si[e];
    dd[(int64_t)rows[r] * k + j] = sd[e];
  }
}

inline unsigned grid_for(int64_t work) {
  int64_t g = (work + 255) / 256;
  if (g > 65536) g = 65536;
  return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace

void launch_select(const SelectArgs& a, hipStream_t st) {
  if (a.nq <= 0) return;
  if (a.topk_mode && a.K > kMaxSortK) fail(RPLSH_ERR_UNSUPPORTED, "top-K select limited to %d", kMaxSortK);
  if (!a.topk_mode || a.K <= kWaveSelK) {
    select_wave_kernel<<<(unsigned)div_up(a.nq, kWaveSelQ), 64 * kWaveSelQ, 0, st>>>(a);
  } else {
    select_kernel<<<a.nq, kSelBlock, 0, st>>>(a);
  }
  RPLSH_HIP(hipGetLastError());
}

void launch_rerank(const RerankArgs& a, hipStream_t st) {
  if (a.nq <= 0) return;
  if (a.kp > kMaxSortK) fail(RPLSH_ERR_UNSUPPORTED, "re-rank list limited to %d", kMaxSortK);
  if (a.d > kMaxKLoopDim) fail(RPLSH_ERR_UNSUPPORTED, "re-rank dimension limited to %d", kMaxKLoopDim);
  const int P = pow2_at_least(a.kp < 1 ? 1 : a.kp);
  if (a.d > 256) {
    // Query in dynamic LDS behind the candidate / band regions (16-byte aligned).
    const size_t lds = rerank_lds_bytes(P, a.fp64_boundary != 0, a.d, false) + (size_t)round_up(a.d, 4) * 4;
    rerank_kernel<-1><<<a.nq, kSelBlock, lds, st>>>(a);
    RPLSH_HIP(hipGetLastError());
    return;
  }
  const bool staged = rerank_staged(a.x, a.ldx, a.d) && a.nq >= kStagedMinQueries;
  const size_t lds = rerank_lds_bytes(P, a.fp64_boundary != 0, a.d, staged);
  if (!staged)
    rerank_kernel<0><<<a.nq, kSelBlock, lds, st>>>(a);
  else if (a.d <= 128)
    rerank_kernel<1><<<a.nq, kSelBlock, lds, st>>>(a);
  else
    rerank_kernel<2><<<a.nq, kSelBlock, lds, st>>>(a);
  RPLSH_HIP(hipGetLastError());
}

void launch_collect_failed(const int* status, int nq, const int* nq_dev, int* list, int* n,
                           hipStream_t st) {
  if (nq <= 0) return;
  collect_failed_kernel<<<(unsigned)div_up(nq, 256), 256, 0, st>>>(status, nq, nq_dev, list, n);
  RPLSH_HIP(hipGetLastError());
}

void launch_sum_counts(const int* counts, int nq, int cap, long long* out, hipStream_t st) {
  if (nq <= 0) return;
  unsigned g = grid_for(nq);
  if (g > 256) g = 256;
  sum_counts_kernel<<<g, 256, 0, st>>>(counts, nq, cap, reinterpret_cast<unsigned long long*>(out));
  RPLSH_HIP(hipGetLastError());
}

void launch_gather_rows(const float* src, int64_t ld, const int* rows, int nrows, int d, float* dst,
                        hipStream_t st) {
  if (nrows <= 0) return;
  gather_rows_kernel<<<grid_for((int64_t)nrows * d), 256, 0, st>>>(src, ld, rows, nrows, d, dst);
  RPLSH_HIP(hipGetLastError());
}

void launch_scatter_results(const int32_t* si, const float* sd, const int* rows, int nrows, int k,
                            int32_t* di, float* dd, hipStream_t st) {
  if (nrows <= 0 || k <= 0) return;
  scatter_results_kernel<<<grid_for((int64_t)nrows * k), 256, 0, st>>>(si, sd, rows, nrows, k, di, dd);
  RPLSH_HIP(hipGetLastError());
}

}  // namespace rplsh
