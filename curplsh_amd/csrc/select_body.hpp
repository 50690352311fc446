// Device bodies of the candidate selects and the exact re-rank, shared by
// their own kernels (select.hip) and the persistent serving kernel
// (persist.hip), so both run the same arithmetic. See select.hip for what
// they compute.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "device.hpp"

namespace rplsh {

namespace {

constexpr int kSelBlock = 256;
constexpr int kSelLds = 4096;     // candidates staged in LDS when they fit
constexpr int kMaxSortK = 4096;   // longest sorted list

// Bitonic sort of s[0..P) ascending (P a power of two), whole block.
__device__ inline void bitonic_sort_u64(uint64_t* s, int P) {
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < (P >> 1); i += blockDim.x) {
        const int lo = ((i & ~(j - 1)) << 1) | (i & (j - 1));
        const int hi = lo + j;
        const bool up = (lo & k) == 0;
        const uint64_t x = s[lo], y = s[hi];
        if ((x > y) == up) {
          s[lo] = y;
          s[hi] = x;
        }
      }
      __syncthreads();
    }
  }
}

__host__ __device__ inline int pow2_at_least(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

// K-th smallest (1-based) of the n unique words at src, whole block.
__device__ uint64_t radix_kth(const uint64_t* src, int n, int K, int* hist, int* sh) {
  uint64_t prefix = 0;
  int kr = K;
  for (int pass = 7; pass >= 0; --pass) {
    const int shift = 8 * pass;
    const uint64_t mask = pass == 7 ? 0ull : (~0ull << (shift + 8));
    for (int i = threadIdx.x; i < 256; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const uint64_t v = src[i];
      if ((v & mask) == prefix) atomicAdd(&hist[(v >> shift) & 255], 1);
    }
    __syncthreads();
    // Inclusive scan over the 256 bins (Hillis-Steele in place).
    for (int off = 1; off < 256; off <<= 1) {
      int v = 0;
      if (threadIdx.x < 256 && threadIdx.x >= off) v = hist[threadIdx.x - off];
      __syncthreads();
      if (threadIdx.x < 256) hist[threadIdx.x] += v;
      __syncthreads();
    }
    if (threadIdx.x < 256) {
      const int before = threadIdx.x == 0 ? 0 : hist[threadIdx.x - 1];
      if (before < kr && kr <= hist[threadIdx.x]) {
        sh[0] = threadIdx.x;
        sh[1] = before;
      }
    }
    __syncthreads();
    prefix |= (uint64_t)sh[0] << shift;
    kr -= sh[1];
    __syncthreads();
  }
  return prefix;
}


// Wave-per-query select (threshold mode, and top-K mode for K <= kWaveSelK):
// four 8-bit radix passes over the 32-bit key find the key T of the K-th
// smallest word; only when several words share T, four more passes over
// their indices find the K-th word exactly. A histogram lives in LDS, its
// scan is a wave prefix sum: no workgroup barriers. Candidates are read from
// global memory (L2) once per pass.
constexpr int kWaveSelK = 512;
constexpr int kWaveSelQ = 4;  // queries (waves) per 256-thread block

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Inclusive prefix sum over the 64 lanes.
__device__ inline int wave_incl_scan(int v, int lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int u = __shfl_up(v, off, 64);
    if (lane >= off) v += u;
  }
  return v;
}

// One 8-bit radix step: among values v (selected by `pick`) finds the bin
// holding the kr-th smallest; returns the bin, subtracts the counts of lower
// bins from kr and sets cnt to the bin's count.
template <typename Get>
__device__ inline int wave_radix_step(int* hist, int lane, int n, int& kr, int& cnt, Get get) {
  reinterpret_cast<int4*>(hist)[lane] = int4{0, 0, 0, 0};
  wave_sync();
  for (int i = lane; i < n; i += 64) {
    int b;
    if (get(i, b)) atomicAdd(&hist[b], 1);
  }
  wave_sync();
  const int4 h4 = reinterpret_cast<const int4*>(hist)[lane];
  const int local = h4.x + h4.y + h4.z + h4.w;
  const int incl = wave_incl_scan(local, lane);
  const int excl = incl - local;
  const bool mine = excl < kr && kr <= incl;
  int bin = lane * 4, before = excl, c = h4.x;
  if (mine) {
    if (kr > before + h4.x) { before += h4.x; ++bin; c = h4.y;
      if (kr > before + h4.y) { before += h4.y; ++bin; c = h4.z;
        if (kr > before + h4.z) { before += h4.z; ++bin; c = h4.w; } } }
  }
  const unsigned long long m = __ballot(mine);
  const int src = m ? __builtin_ctzll(m) : 0;
  bin = __shfl(bin, src, 64);
  before = __shfl(before, src, 64);
  cnt = __shfl(c, src, 64);
  kr -= before;
  wave_sync();
  return bin;
}


// The select of query q by the calling wave; hist: 256 ints, out:
// kWaveSelK words of LDS owned by the wave.
__device__ void select_wave_query(const SelectArgs& a, int q, int* hist, uint64_t* out) {
  const int lane = threadIdx.x & 63;
  const int count = a.counts[q];
  const int n = count < a.cap ? count : a.cap;
  int st = 0;
  if (count > a.cap) st |= kStatusOverflow;
  const int K = a.K;
  if (n < K) st |= kStatusShort;
  const uint64_t* src = a.cand + (int64_t)q * a.cap;

  // Key T of the K-th smallest word; kr: its rank among the words with key T.
  uint32_t T = 0xffffffffu;
  int kr = K, cnt = 0;
  const bool ranked = n >= K && K > 0;
  if (ranked) {
    uint32_t prefix = 0;
#pragma unroll 1
    for (int pass = 3; pass >= 0; --pass) {
      const int shift = 8 * pass;
      const uint32_t hmask = pass == 3 ? 0u : (0xffffffffu << (shift + 8));
      const int bin = wave_radix_step(hist, lane, n, kr, cnt, [&](int i, int& b) {
        const uint32_t k = (uint32_t)(src[i] >> 32);
        b = (int)((k >> shift) & 255u);
        return (k & hmask) == prefix;
      });
      prefix |= (uint32_t)bin << shift;
    }
    T = prefix;
  }

  if (!a.topk_mode) {
    if (lane == 0) {
      a.tau_out[q] = ranked ? key_from_bits(T) : a.tau_in[q];
      if (st && a.status) atomicOr(a.status + q, st);
    }
    return;
  }

  // Top-K mode: W = the K-th smallest word; every word <= W is selected.
  uint64_t W = kEmptyCand;
  if (ranked && n > K) {
    uint32_t idx = 0xffffffffu;
    if (kr < cnt) {  // ties at T: the kr-th smallest index among them
      uint32_t prefix = 0;
#pragma unroll 1
      for (int pass = 3; pass >= 0; --pass) {
        const int shift = 8 * pass;
        const uint32_t hmask = pass == 3 ? 0u : (0xffffffffu << (shift + 8));
        int c2 = 0;
        const int bin = wave_radix_step(hist, lane, n, kr, c2, [&](int i, int& b) {
          const uint64_t v = src[i];
          const uint32_t lo = (uint32_t)v;
          b = (int)((lo >> shift) & 255u);
          return (uint32_t)(v >> 32) == T && (lo & hmask) == prefix;
        });
        prefix |= (uint32_t)bin << shift;
      }
      idx = prefix;
    }
    W = ((uint64_t)T << 32) | idx;
  }
  const int take = n < K ? n : K;
  int base = 0;
  for (int i0 = 0; i0 < n; i0 += 64) {
    const int i = i0 + lane;
    const uint64_t v = i < n ? src[i] : kEmptyCand;
    const bool sel = i < n && v <= W;
    const unsigned long long m = __ballot(sel);
    if (sel) {
      const int pos = base + __popcll(m & ((1ull << lane) - 1ull));
      if (pos < kWaveSelK) out[pos] = v;
    }
    base += __popcll(m);
  }
  int P = 1;
  while (P < take) P <<= 1;
  for (int i = take + lane; i < P; i += 64) out[i] = kEmptyCand;
  wave_sync();
  // Bitonic sort of out[0..P), one wave.
  for (int k2 = 2; k2 <= P; k2 <<= 1) {
    for (int j = k2 >> 1; j > 0; j >>= 1) {
      for (int i = lane; i < (P >> 1); i += 64) {
        const int lo = ((i & ~(j - 1)) << 1) | (i & (j - 1));
        const int hi = lo + j;
        const bool up = (lo & k2) == 0;
        const uint64_t x = out[lo], y = out[hi];
        if ((x > y) == up) {
          out[lo] = y;
          out[hi] = x;
        }
      }
      wave_sync();
    }
  }
  uint64_t* dst = a.topk_out + (int64_t)q * K;
  for (int i = lane; i < K; i += 64) dst[i] = i < take ? out[i] : kEmptyCand;
  if (lane == 0) {
    a.bound_out[q] = count > K && K > 0 ? key_from_bits(T) : a.tau_in[q];
    if (st && a.status) atomicOr(a.status + q, st);
  }
}

// Canonical FP32 relative distance of query sq (in LDS) to row idx.
// Loads of one row are issued kRowBatch float4 at a time before the
// arithmetic consumes them; the order over j is unchanged.
constexpr int kRowBatch = 8;

// Canonical relative distance of one row computed by one lane: 16-byte loads
// kRowBatch at a time when rows are 16-byte aligned, else scalar loads.
__device__ inline float canonical_rel(const float* __restrict__ x, int64_t ldx,
                                      const float* __restrict__ xnorm, int d, const float* sq,
                                      int32_t idx, bool vec4) {
  const float* xr = x + (int64_t)idx * ldx;
  const float xn = xnorm[idx];
  float s = 0.0f;
  if (vec4) {
    for (int j0 = 0; j0 < d; j0 += 4 * kRowBatch) {
      float4 v[kRowBatch];
#pragma unroll
      for (int t = 0; t < kRowBatch; ++t)
        if (j0 + 4 * t < d) v[t] = *reinterpret_cast<const float4*>(xr + j0 + 4 * t);
#pragma unroll
      for (int t = 0; t < kRowBatch; ++t) {
        if (j0 + 4 * t < d) {
          const int j = j0 + 4 * t;
          s = fmaf(sq[j + 0], v[t].x, s);
          s = fmaf(sq[j + 1], v[t].y, s);
          s = fmaf(sq[j + 2], v[t].z, s);
          s = fmaf(sq[j + 3], v[t].w, s);
        }
      }
    }
  } else {
    for (int j = 0; j < d; ++j) s = fmaf(sq[j], xr[j], s);
  }
  return fmaf(-2.0f, s, xn);
}

// Rows a wave stages per step of score_rows_staged, and the staging row
// stride in floats (padding keeps the four chain lanes on different banks).
constexpr int kRowGroup = 4;
constexpr int kSelWaves = kSelBlock / 64;
constexpr int kStagedMinQueries = 16;
__host__ __device__ inline int stage_ld(int d) { return d + 4; }

// Scores candidates [0, kp) of one query into s[] as (canonical rel, index)
// words; empty candidates give kEmptyCand. Rows need 16-byte alignment and
// d <= 128 * NP. Each wave owns a contiguous slice of the list and walks it
// in steps of kRowGroup rows: half a wave loads one row, 16 B per lane (NP
// loads per lane and row), so a load instruction touches two rows (a
// random-row gather is limited by the rows one instruction touches, not by
// bytes); the rows go through a double-buffered LDS stage, and while
// kRowGroup lanes run the sequential FMA chains from LDS the next step's rows
// are in flight. The loads are unconditional (absent rows read a present
// row of the same chunk, parts past d re-read the last part) so that no
// branch around them forces a wait.
// One step's rows in registers: row slots half (r0) and 2 + half (r1), parts
// p0 = lane % 32 and (NP == 2) p1 = p0 + 32, clamped to the row.
struct StageRegs {
  float4 r0p0, r0p1, r1p0, r1p1;
};

template <int NP>
__device__ __forceinline__ StageRegs stage_load(const RerankArgs& a, int32_t rowi, int half, int p0, int p1,
                                                int g) {
  StageRegs v;
  const int r = g * kRowGroup + half;
  const float* x0 = a.x + (int64_t)__shfl(rowi, r & 63, 64) * a.ldx;
  const float* x1 = a.x + (int64_t)__shfl(rowi, (r + 2) & 63, 64) * a.ldx;
  v.r0p0 = *reinterpret_cast<const float4*>(x0 + 4 * p0);
  v.r1p0 = *reinterpret_cast<const float4*>(x1 + 4 * p0);
  if constexpr (NP == 2) {
    v.r0p1 = *reinterpret_cast<const float4*>(x0 + 4 * p1);
    v.r1p1 = *reinterpret_cast<const float4*>(x1 + 4 * p1);
  }
  return v;
}

template <int NP>
__device__ void score_rows_staged(const RerankArgs& a, const float* sq, const uint64_t* src,
                                  int kp, uint64_t* s, float* stage, float* sxn) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int d = a.d, ldr = stage_ld(d), parts = d >> 2;
  float* const wstage = stage + (size_t)w * 2 * kRowGroup * ldr;
  const int per = (kp + kSelWaves - 1) / kSelWaves;
  const int c0 = min(kp, w * per), c1 = min(kp, c0 + per);
  const int half = lane >> 5;
  const int p0 = min(lane & 31, parts - 1), p1 = min((lane & 31) + 32, parts - 1);
  const bool w0 = (lane & 31) < parts, w1 = NP == 2 && (lane & 31) + 32 < parts;
  for (int cb = c0; cb < c1; cb += 64) {
    const int nc = min(64, c1 - cb);
    const uint64_t word = lane < nc ? src[cb + lane] : kEmptyCand;
    const int32_t myidx = word != kEmptyCand ? cand_idx(word) : -1;
    const uint64_t present = __ballot(myidx >= 0);
    if (present == 0) {
      if (lane < nc) s[cb + lane] = kEmptyCand;
      continue;
    }
    const int32_t any_row = __shfl(myidx, __ffsll((unsigned long long)present) - 1, 64);
    const int32_t rowi = myidx >= 0 ? myidx : any_row;
    // The chunk's norms wait here, before any row load is in flight: read
    // inside the loop, a register load would wait for the prefetched rows too.
    sxn[w * 64 + lane] = a.xnorm[rowi];
    const int ng = (nc + kRowGroup - 1) / kRowGroup;
    StageRegs v = stage_load<NP>(a, rowi, half, p0, p1, 0);
    for (int g = 0; g < ng; ++g) {
      float* const rw = wstage + (g & 1) * kRowGroup * ldr;
      float* const s0 = rw + half * ldr;
      float* const s1 = rw + (2 + half) * ldr;
      if (w0) {
        *reinterpret_cast<float4*>(s0 + 4 * p0) = v.r0p0;
        *reinterpret_cast<float4*>(s1 + 4 * p0) = v.r1p0;
      }
      if (w1) {
        *reinterpret_cast<float4*>(s0 + 4 * p1) = v.r0p1;
        *reinterpret_cast<float4*>(s1 + 4 * p1) = v.r1p1;
      }
      if (g + 1 < ng) v = stage_load<NP>(a, rowi, half, p0, p1, g + 1);
      wave_sync();
      const int r = g * kRowGroup + lane;
      const int idx = __shfl(myidx, r & 63, 64);
      if (lane < kRowGroup && r < nc) {
        const float xn = sxn[w * 64 + r];
        uint64_t out = kEmptyCand;
        if (idx >= 0) {
          const float* xr = rw + lane * ldr;
          float acc = 0.0f;
#pragma unroll 4
          for (int j = 0; j < d; j += 4) {
            const float4 t = *reinterpret_cast<const float4*>(xr + j);
            acc = fmaf(sq[j + 0], t.x, acc);
            acc = fmaf(sq[j + 1], t.y, acc);
            acc = fmaf(sq[j + 2], t.z, acc);
            acc = fmaf(sq[j + 3], t.w, acc);
          }
          out = pack_cand(fmaf(-2.0f, acc, xn), idx);
        }
        s[cb + r] = out;
      }
    }
    wave_sync();
  }
}

// sum_j (q_j - x_j)^2 in double, sequential j, no contraction: the same
// arithmetic as the FP64 oracle. Loads batched as in canonical_rel.
__device__ inline double fp64_dist(const float* __restrict__ x, int64_t ldx, int d,
                                   const float* sq, int32_t idx, bool vec4) {
  const float* xr = x + (int64_t)idx * ldx;
  double s = 0.0;
  if (vec4) {
    for (int j0 = 0; j0 < d; j0 += 4 * kRowBatch) {
      float4 v[kRowBatch];
#pragma unroll
      for (int t = 0; t < kRowBatch; ++t)
        if (j0 + 4 * t < d) v[t] = *reinterpret_cast<const float4*>(xr + j0 + 4 * t);
#pragma unroll
      for (int t = 0; t < kRowBatch; ++t) {
        if (j0 + 4 * t < d) {
          const int j = j0 + 4 * t;
          const float e[4] = {v[t].x, v[t].y, v[t].z, v[t].w};
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const double dt = __dsub_rn((double)sq[j + u], (double)e[u]);
            s = __dadd_rn(s, __dmul_rn(dt, dt));
          }
        }
      }
    }
  } else {
    for (int j = 0; j < d; ++j) {
      const double t = __dsub_rn((double)sq[j], (double)xr[j]);
      s = __dadd_rn(s, __dmul_rn(t, t));
    }
  }
  return s;
}

constexpr int kMaxBand = 512;

// Dynamic LDS, sized by the launch to the candidate list: s[P] words, then
// one region used first as the row stage of score_rows_staged (when rows are
// 16-byte aligned) and afterwards, for the FP64 band, as bw[B] words, bd[B]
// doubles and bi[B] ints, B = min(kMaxBand, P).
__host__ __device__ inline int rerank_band_cap(int P) { return P < kMaxBand ? P : kMaxBand; }
__host__ __device__ inline int rerank_words(int P) { return P < 2 ? 2 : P; }
__host__ __device__ inline size_t rerank_lds_bytes(int P, bool band, int d, bool staged) {
  const size_t band_bytes = band ? (size_t)rerank_band_cap(P) * (8 + 8 + 4) : 0;
  const size_t stage_bytes = staged ? (size_t)kSelWaves * 2 * kRowGroup * stage_ld(d) * 4 : 0;
  return (size_t)rerank_words(P) * 8 + (band_bytes > stage_bytes ? band_bytes : stage_bytes);
}
__host__ __device__ inline bool rerank_staged(const float* x, int64_t ldx, int d) {
  return d >= 4 && d % 4 == 0 && ldx % 4 == 0 && ((uintptr_t)x & 15) == 0;
}

// NP = 0: one lane per candidate (rows not 16-byte aligned, or a batch too
// small to fill the device, where the staged walk's serial steps per wave
// cost more than they save); NP = 1, 2: score_rows_staged<NP> (d <= 128,
// d <= 256); NP = -1: one lane per candidate for d > 256, the query kept in
// dynamic LDS after the candidate and band regions.
// The re-rank of query q by the whole (kSelBlock-thread) workgroup.
// rerank_lds: the dynamic region (rerank_lds_bytes, plus the query for
// NP < 0); sq_fixed: 256 floats (NP >= 0); sxn: kSelWaves * 64 floats; sh: 4.
template <int NP>
__device__ void rerank_query(const RerankArgs& a, int q, uint64_t* rerank_lds, float* sq_fixed,
                             float* sxn, int* sh) {
  uint64_t* s = rerank_lds;
  const int d = a.d;
  const int kp = a.kp;
  const int P = pow2_at_least(kp);
  float* sq = sq_fixed;
  if constexpr (NP < 0)
    sq = reinterpret_cast<float*>(reinterpret_cast<char*>(rerank_lds) +
                                  rerank_lds_bytes(P, a.fp64_boundary != 0, d, false));
  for (int j = threadIdx.x; j < d; j += blockDim.x) sq[j] = a.q[(int64_t)q * a.ldq + j];
  __syncthreads();
  const bool vec4 = rerank_staged(a.x, a.ldx, d);  // 16-byte loads also in the FP64 band
  const int band_cap = rerank_band_cap(P);
  uint64_t* bw = s + rerank_words(P);  // 16-byte aligned
  double* bd = reinterpret_cast<double*>(bw + band_cap);
  int* bi = reinterpret_cast<int*>(bd + band_cap);
  const uint64_t* src = a.topk + (int64_t)q * kp;
  if constexpr (NP > 0) {
    score_rows_staged<NP>(a, sq, src, kp, s, reinterpret_cast<float*>(bw), sxn);
    for (int i = kp + threadIdx.x; i < P; i += blockDim.x) s[i] = kEmptyCand;
  } else {
    for (int i = threadIdx.x; i < P; i += blockDim.x) {
      uint64_t w = kEmptyCand;
      if (i < kp) {
        const uint64_t c = src[i];
        if (c != kEmptyCand) {
          const int32_t idx = cand_idx(c);
          w = pack_cand(canonical_rel(a.x, a.ldx, a.xnorm, d, sq, idx, vec4), idx);
        }
      }
      s[i] = w;
    }
  }
  __syncthreads();
  bitonic_sort_u64(s, P);

  if (threadIdx.x == 0) {
    int valid = 0;
    // Empty words sort last; count the real ones.
    int lo = 0, hi = kp;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (s[mid] == kEmptyCand) hi = mid; else lo = mid + 1;
    }
    valid = lo;
    sh[0] = valid;
  }
  __syncthreads();
  const int valid = sh[0];
  const int k = a.k;
  const float qn = a.qnorm[q];

  // Plain path: certify the candidate margin.
  if (a.bound && threadIdx.x == 0 && valid >= k && k > 0) {
    const float relk = cand_key(s[k - 1]);
    const float eps = a.eps_coeff * (qn + 2.0f * a.xn_max);
    const float b = a.bound[q];
    if (!(relk < b - eps)) {
      atomicOr(a.status + q, kStatusCertFail);
      a.relk_out[q] = relk;
    }
  }

  // Precision forms: FP64 order inside the band around the k-th distance.
  bool banded = false;
  int band0 = 0, band1 = -1;
  if (a.fp64_boundary && valid > k && k > 0) {
    if (threadIdx.x == 0) {
      const float relk = cand_key(s[k - 1]);
      const float E = a.band_coeff * (qn + 2.0f * a.xn_max) + 4.0f * 5.96e-8f * (fabsf(relk) + qn);
      int b0 = k - 1, b1 = k - 1;
      while (b0 > 0 && cand_key(s[b0 - 1]) >= relk - 2.0f * E && (k - 1) - (b0 - 1) < band_cap / 2) --b0;
      while (b1 + 1 < valid && cand_key(s[b1 + 1]) <= relk + 2.0f * E && (b1 + 1) - b0 < band_cap) ++b1;
      sh[1] = b0;
      sh[2] = b1;
    }
    __syncthreads();
    const int b0 = sh[1], b1 = sh[2];
    const int nb = b1 - b0 + 1;
    if (nb > 1 && b1 >= k) {
      for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        const int32_t idx = cand_idx(s[b0 + i]);
        bd[i] = fp64_dist(a.x, a.ldx, d, sq, idx, vec4);
        bi[i] = b0 + i;
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        // Insertion sort of the band by (fp64 distance, index); nb is small.
        for (int i = 1; i < nb; ++i) {
          const double kd = bd[i];
          const int kb = bi[i];
          const int32_t kx = cand_idx(s[kb]);
          int j = i - 1;
          while (j >= 0 && (bd[j] > kd || (bd[j] == kd && cand_idx(s[bi[j]]) > kx))) {
            bd[j + 1] = bd[j];
            bi[j + 1] = bi[j];
            --j;
          }
          bd[j + 1] = kd;
          bi[j + 1] = kb;
        }
        atomicAdd(a.fp64_count, nb);
      }
      __syncthreads();
      // Rewrite the band positions in FP64 order (words keep their FP32 key).
      for (int i = threadIdx.x; i < nb; i += blockDim.x) bw[i] = s[bi[i]];
      __syncthreads();
      for (int i = threadIdx.x; i < nb; i += blockDim.x) s[b0 + i] = bw[i];
      __syncthreads();
      banded = true;
      band0 = b0;
      band1 = b1;
    }
  }

  if (banded) {
    // The band decided membership by FP64 distance; report its members' FP64
    // distances (rounded to FP32) and order the k results by the reported
    // distance, then index, so the list stays ascending.
    const int Pk = pow2_at_least(k);
    for (int j = threadIdx.x; j < Pk; j += blockDim.x) {
      uint64_t w = kEmptyCand;
      if (j < k && j < valid) {
        const int32_t idx = cand_idx(s[j]);
        float dist;
        if (j >= band0 && j <= band1) {
          dist = a.exact ? (float)bd[j - band0] : (float)(bd[j - band0] - (double)qn);
        } else {
          const float rel = cand_key(s[j]);
          dist = a.exact ? rel + qn : rel;
        }
        w = pack_cand(dist, idx);
      }
      s[j] = w;
    }
    __syncthreads();
    bitonic_sort_u64(s, Pk);
    for (int j = threadIdx.x; j < k; j += blockDim.x) {
      const int64_t o = (int64_t)q * k + j;
      const uint64_t w = s[j];
      a.out_idx[o] = w == kEmptyCand ? -1 : cand_idx(w);
      a.out_dist[o] = w == kEmptyCand ? FLT_MAX : cand_key(w);
    }
    return;
  }

  for (int j = threadIdx.x; j < k; j += blockDim.x) {
    const int64_t o = (int64_t)q * k + j;
    if (j < valid) {
      const uint64_t w = s[j];
      const float rel = cand_key(w);
      a.out_idx[o] = cand_idx(w);
      a.out_dist[o] = a.exact ? rel + qn : rel;
    } else {
      a.out_idx[o] = -1;
      a.out_dist[o] = FLT_MAX;
    }
  }
}


}  // namespace

}  // namespace rplsh
