// Random-projection LSH front end (IndexRPLSH): the "RPLSH" step the
// reference's name promises but does not implement. Its only LSH piece is the
// Gaussian projection-matrix initialiser (src/internal/matrix_ops.cu:71-76,
// curandGenerateNormal with mean / stddev / seed); everything after it is
// this library's own design, placed in front of the exact re-rank of the
// brute-force path.
//
// Index (add):
//   W [L*K][d]   standard normal projections from the library's seeded
//                generator (row r = table r / K, bit r % K)
//   code_t(x)    = sum_b [s_{t,b}(x) >= o_{t,b}] << b,  s_{t,b}(x) = canonical
//                fmaf chain over j of W[t*K+b][j] * x[j] (as the distance dot)
//   offsets o    0 (hyperplanes through the origin: angular hashing), or, with
//                `center`, each projection's median over a sample of the first
//                add (the first min(n, 65536) rows; element m/2 of the m sorted
//                dots), so every bit halves the data and hyperplanes cut
//                through clusters: collisions then follow L2 proximity
//   table t      rows stably sorted by code_t (ties: ascending row), as
//                sorted codes + row permutation, so a bucket is a range
// Search, per query:
//   probes       bucket code_t(q) and, for p = 1..P, code_t(q) with the bit of
//                the p-th smallest |s_{t,b}(q) - o_{t,b}| flipped (ties: lower
//                bit; query-directed multi-probe, Lv et al., VLDB 2007)
//   candidates   buckets concatenated in (probe, table) order, rows in bucket
//                order, cut after B entries, then de-duplicated
//   result       canonical FP32 re-rank of the distinct candidates (the
//                plain path's rerank_kernel): k smallest by (rel, index),
//                D = rel + ||q||^2; fewer than k candidates pad with -1/FLT_MAX
//
// Kernels: lsh_hash_kernel (VALU: a workgroup stages 64 rows in LDS, each
// wave hashes a quarter of the tables for them, a lane's K chains of a table
// interleaved, projections through the scalar cache), lsh_gather_kernel (one
// workgroup per query: probe codes, bucket ranges by binary search, a scan,
// the gather into an LDS hash set that keeps each row once; bucket searches
// start from a directory over the codes' top 16 bits). The table build is a
// stable LSD radix sort per table (hipCUB) at add time, off the search path.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cstdint>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "engine.hpp"
#include "kernels.hpp"

namespace rplsh {
namespace {

constexpr int kLshMaxBits = 24;
constexpr int kLshMaxEntries = 1024;  // (probes + 1) * tables
constexpr int kLshMaxBudget = 8192;   // candidates gathered per query
constexpr int kLshMaxExact = 4096;    // the exact re-rank's longest list
constexpr int kLshMaxTable = 11261;   // hash-set slots, prime (44 KB: the block stays under 64 KB of LDS)

// Slots of the gather's de-duplication hash set for a budget of B entries:
// the smallest prime >= twice B rounded up to a power of two, at most
// kLshMaxTable (load <= 0.73); prime, so a double-hashing probe sequence
// visits every slot.
inline int lsh_table_slots(int B) {
  int p = 1;
  while (p < B) p <<= 1;
  int t = std::min(2 * p, kLshMaxTable);
  auto prime = [](int v) {
    if (v < 2) return false;
    for (int i = 2; i * i <= v; ++i)
      if (v % i == 0) return false;
    return true;
  };
  while (!prime(t)) ++t;
  return t;
}
constexpr int kLshBlock = 256;
constexpr int64_t kLshCenterSample = 65536;

// ---------------------------------------------------------------- hashing

constexpr int kHashRows = 64;   // rows a workgroup stages (one per lane)
constexpr int kHashWaves = 4;
// Chains a hash lane runs per table: K rounded up to a multiple of 8.
inline int hash_width(int K) { return (K + 7) / 8 * 8; }

// Codes (and margins) of rows [0, n) for every table. A workgroup stages
// kHashRows rows in LDS with coalesced loads (row stride d + 1 floats: lanes
// on different banks), then wave w hashes tables w, w + 4, ... of them, one
// row per lane; the projections come from wj[t][j][0..KB) (j-major, zero
// padded from K to KB = K rounded up to 8) at uniform addresses, i.e. through
// the scalar cache, so each FMA takes its projection value as a scalar
// operand and no bit needs a branch. With gridDim.y > 1 (few rows, e.g. a
// query batch) the tables are spread over blockIdx.y as well. Bit b of codes[t * ldc + i] is
// s >= off[t * K + b], s the canonical fmaf chain over j of W[t * K + b][j] *
// x[j]; margins, when not null, hold s - off at
// margins[i * mrow + (t * K + b) * mcol] (queries: mrow = L * K, mcol = 1;
// the transposed sample of the offset computation: mrow = 1, mcol = rows).
// codes may be null.
template <int KB>
__global__ __launch_bounds__(kHashRows * kHashWaves) void lsh_hash_kernel(
    const float* __restrict__ x, int64_t ldx, int64_t n, int d, const float* __restrict__ wj,
    const float* __restrict__ off, int L, int K, uint32_t* __restrict__ codes, int64_t ldc,
    float* __restrict__ margins, int64_t mrow, int64_t mcol) {
  extern __shared__ float xs[];  // [kHashRows][d + 1]
  const int64_t r0 = (int64_t)blockIdx.x * kHashRows;
  const int rows = (int)(n - r0 < kHashRows ? n - r0 : kHashRows);
  const int ld = d + 1;
  for (int e = threadIdx.x; e < rows * d; e += blockDim.x) {
    const int r = e / d, j = e - r * d;
    xs[r * ld + j] = x[(r0 + r) * ldx + j];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (lane >= rows) return;
  const float* xr = xs + lane * ld;
  const int64_t i = r0 + lane;
  for (int t = blockIdx.y * kHashWaves + wave; t < L; t += gridDim.y * kHashWaves) {
    const float* wt = wj + (int64_t)t * d * KB;
    float acc[KB];
#pragma unroll
    for (int b = 0; b < KB; ++b) acc[b] = 0.0f;
    for (int j = 0; j < d; ++j) {
      const float xv = xr[j];
      const float* wr = wt + j * KB;
#pragma unroll
      for (int b = 0; b < KB; ++b) acc[b] = fmaf(wr[b], xv, acc[b]);
    }
    const float* o = off + (int64_t)t * K;
    uint32_t code = 0;
#pragma unroll
    for (int b = 0; b < KB; ++b)
      if (b < K && acc[b] >= o[b]) code |= 1u << b;
    if (codes) codes[(int64_t)t * ldc + i] = code;
    if (margins) {
#pragma unroll
      for (int b = 0; b < KB; ++b)
        if (b < K) margins[i * mrow + ((int64_t)t * K + b) * mcol] = acc[b] - o[b];
    }
  }
}

// off[r] = sorted[r * m + m / 2] for r < nr.
__global__ void lsh_median_kernel(const float* __restrict__ sorted, int64_t m, int nr, float* __restrict__ off) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < nr) off[r] = sorted[(int64_t)r * m + m / 2];
}

__global__ void lsh_segment_offsets_kernel(int* __restrict__ seg, int nr, int m) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r <= nr) seg[r] = r * m;
}

__global__ void iota_kernel(int32_t* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = (int32_t)i;
}

// --------------------------------------------------------------- gathering

// Directory over the top H = min(K, 16) bits of the codes: dir[t][h] is the
// first position of table t's sorted codes whose top bits are >= h, for
// h = 0..2^H (dir[t][2^H] = n), so a bucket search starts inside one slot.
constexpr int kDirBits = 16;
__host__ __device__ inline int dir_bits(int K) { return K < kDirBits ? K : kDirBits; }

__device__ inline int64_t lower_bound_u32(const uint32_t* a, int64_t lo, int64_t hi, uint32_t v) {
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__global__ void lsh_directory_kernel(const uint32_t* __restrict__ sorted, int64_t ld, int64_t n, int L, int K,
                                     int32_t* __restrict__ dir) {
  const int H = dir_bits(K);
  const int64_t slots = ((int64_t)1 << H) + 1;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= slots * L) return;
  const int t = (int)(e / slots);
  const int64_t h = e - (int64_t)t * slots;
  const uint64_t first = (uint64_t)h << (K - H);  // smallest code with top bits h
  dir[e] = first >= ((uint64_t)1 << K) ? (int32_t)n
                                       : (int32_t)lower_bound_u32(sorted + (int64_t)t * ld, 0, n, (uint32_t)first);
}

struct LshGatherArgs {
  const int32_t* dir;      // [L][2^H + 1] directory (lsh_directory_kernel)
  const uint32_t* sorted;  // [L][ld] sorted codes of each table
  const int32_t* perm;     // [L][ld] rows in bucket order
  int64_t ld;              // table stride
  int64_t n;               // rows in each table
  const uint32_t* qcodes;  // [L][nq]
  const float* qmarg;      // [nq][L*K]
  int nq, L, K, probes, budget;
  int table;               // de-duplication hash-set slots (> budget)
  uint64_t* cand;          // [nq][budget] words (index only; key ignored by the re-rank):
                           // the distinct rows first, in no particular order, then empty
  unsigned long long* raw_total;
  unsigned long long* distinct_total;
};

// Sort key of a probe candidate bit: |margin| as an ordered bit pattern
// (NaN after +inf), then the bit index.
__device__ inline uint64_t probe_key(float m, int b) {
  return ((uint64_t)__float_as_uint(fabsf(m)) << 5) | (uint64_t)b;
}

__global__ __launch_bounds__(kLshBlock) void lsh_gather_kernel(LshGatherArgs a) {
  __shared__ uint32_t pcode[kLshMaxEntries];
  __shared__ int32_t elo[kLshMaxEntries];
  __shared__ int32_t ecnt[kLshMaxEntries];
  __shared__ int32_t eoff[kLshMaxEntries + 1];
  __shared__ int32_t part[kLshBlock];
  extern __shared__ int32_t tab[];  // [table] hash set of rows (dynamic: sized by the launch)
  __shared__ int32_t sh_distinct;
  const int q = blockIdx.x;
  if (q >= a.nq) return;
  const int L = a.L, K = a.K, P = a.probes, B = a.budget;
  const int E = (P + 1) * L;

  // 1. probe codes of every table, entry e = p * L + t
  for (int t = threadIdx.x; t < L; t += blockDim.x) {
    const uint32_t c = a.qcodes[(int64_t)t * a.nq + q];
    pcode[t] = c;
    const float* m = a.qmarg + (int64_t)q * L * K + (int64_t)t * K;
    uint32_t used = 0;
    for (int p = 1; p <= P; ++p) {
      uint64_t best = ~0ull;
      for (int b = 0; b < K; ++b) {
        if (used >> b & 1u) continue;
        const uint64_t key = probe_key(m[b], b);
        if (key < best) best = key;
      }
      const int bb = (int)(best & 31u);
      used |= 1u << bb;
      pcode[p * L + t] = c ^ (1u << bb);
    }
  }
  __syncthreads();

  // 2. bucket ranges inside the code's directory slot; counts clamped to B
  //    (only the first B entries matter)
  const int H = dir_bits(K);
  const int64_t slots = ((int64_t)1 << H) + 1;
  for (int e = threadIdx.x; e < E; e += blockDim.x) {
    const int t = e % L;
    const uint32_t* s = a.sorted + (int64_t)t * a.ld;
    const uint32_t c = pcode[e];
    const int32_t* dt = a.dir + (int64_t)t * slots;
    const uint32_t h = c >> (K - H);
    const int64_t s0 = dt[h], s1 = dt[h + 1];
    const int64_t lo = lower_bound_u32(s, s0, s1, c);
    const int64_t hi = lower_bound_u32(s, lo, s1, c + 1);
    elo[e] = (int32_t)lo;
    ecnt[e] = (int32_t)(hi - lo < B ? hi - lo : B);
  }
  __syncthreads();

  // 3. exclusive scan of the counts: each thread sums a run of entries, a
  //    scan over the threads' sums, then the runs are rewritten
  const int per = (E + blockDim.x - 1) / blockDim.x;
  const int e0 = threadIdx.x * per, e1 = min(E, e0 + per);
  int run = 0;
  for (int e = e0; e < e1; ++e) run += ecnt[e];
  part[threadIdx.x] = run;
  __syncthreads();
  for (int off = 1; off < (int)blockDim.x; off <<= 1) {
    const int v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  int acc = part[threadIdx.x] - run;  // exclusive prefix of this run
  for (int e = e0; e < e1; ++e) {
    eoff[e] = acc;
    acc += ecnt[e];
  }
  __syncthreads();
  if (threadIdx.x == 0) eoff[E] = part[blockDim.x - 1];
  __syncthreads();
  const int R = min(eoff[E], B);

  // 4. gather the first R entries of the concatenation into an LDS hash set
  //    (open addressing, double hashing over a prime T > R slots, so a free one
  //    is always reached): a row's first insertion appends it to the query's list, so the
  //    list holds each distinct row once, in no particular order (every
  //    consumer orders candidates by (distance, index)); appends are counted
  //    per wave (ballot, one atomic).
  const int T = a.table;
  for (int i = threadIdx.x; i < T; i += blockDim.x) tab[i] = -1;
  if (threadIdx.x == 0) sh_distinct = 0;
  __syncthreads();
  uint64_t* out = a.cand + (int64_t)q * B;
  const int lane = threadIdx.x & 63;
  for (int j0 = 0; j0 < R; j0 += blockDim.x) {  // uniform trip count
    const int j = j0 + threadIdx.x;
    bool fresh = false;
    int32_t row = -1;
    if (j < R) {
      int lo = 0, hi = E;  // last e with eoff[e] <= j
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (eoff[mid] <= j) lo = mid; else hi = mid;
      }
      const int t = lo % L;
      row = a.perm[(int64_t)t * a.ld + elo[lo] + (j - eoff[lo])];
      uint32_t slot = (uint32_t)(((uint64_t)((uint32_t)row * 0x9E3779B1u) * (uint32_t)T) >> 32);
      const uint32_t step =
          1u + (uint32_t)(((uint64_t)((uint32_t)row * 0x85EBCA77u) * (uint32_t)(T > 1 ? T - 1 : 1)) >> 32);
      while (true) {
        const int old = atomicCAS(&tab[slot], -1, row);
        if (old == -1) {
          fresh = true;
          break;
        }
        if (old == row) break;
        slot += step;  // step in [1, T - 1]
        if (slot >= (uint32_t)T) slot -= (uint32_t)T;
      }
    }
    const uint64_t m = __ballot(fresh);
    if (m != 0) {
      const int leader = __ffsll((unsigned long long)m) - 1;
      int base = 0;
      if (lane == leader) base = atomicAdd(&sh_distinct, __popcll(m));
      base = __shfl(base, leader, 64);
      const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
      if (fresh) out[base + rank] = pack_cand(0.0f, row);
    }
  }
  __syncthreads();

  // 5. the rest of the list empty; totals
  const int nd = sh_distinct;
  for (int j = nd + threadIdx.x; j < B; j += blockDim.x) out[j] = kEmptyCand;
  if (threadIdx.x == 0) {
    atomicAdd(a.raw_total, (unsigned long long)R);
    atomicAdd(a.distinct_total, (unsigned long long)nd);
  }
}

// BF16 pre-rank of the distinct candidates (first stage when `prerank` is
// set): per query (one workgroup), rel_b = fmaf(-2, dot_b, ||x||^2), written
// as (rel_b, index) words in candidate order (empty slots stay empty).
// dot_b is the product of bf16(q) and bf16(x) (both RNE, zero padded to ldb =
// d rounded up to 8) in piece/tree order: the P = ldb / 8 pieces of 8 elements
// are spread over G = P rounded up to a power of two lanes (P <= 32: the
// dimension is at most 256), lane p chaining fmaf over the 8 elements of piece p
// from +0 in element order (lanes past P keep +0); the G partial sums are added pairwise,
// s[p] += s[p + m] for m = G/2, ..., 1. On the GPU that sum is a butterfly
// (xor m) that leaves every lane of the group with the same bits, because
// addition commutes. G lanes read one row, 16 B each, so one load instruction
// covers 64 / G rows (1 KB, 8 cache lines, when G = 16): the gather runs at
// the rate of the exact re-rank's in half the bytes, and every lane works.
// Each wave owns a quarter of the candidate list and walks it in chunks of 64
// (one candidate word per lane), issuing the chunk's row loads in batches of
// up to 16 before using them.
__device__ __forceinline__ float bf16_piece_dot(const uint4 xv, const float (&qf)[8], float acc) {
  const uint32_t h[4] = {xv.x, xv.y, xv.z, xv.w};
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const float e = __uint_as_float(t & 1 ? h[t >> 1] & 0xffff0000u : h[t >> 1] << 16);
    acc = fmaf(qf[t], e, acc);
  }
  return acc;
}

__device__ __forceinline__ void bf16_piece_unpack(const uint4 v, float (&f)[8]) {
  const uint32_t h[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int t = 0; t < 8; ++t) f[t] = __uint_as_float(t & 1 ? h[t >> 1] & 0xffff0000u : h[t >> 1] << 16);
}

template <int G>
__global__ __launch_bounds__(kLshBlock) void lsh_prerank_kernel(const uint16_t* __restrict__ xb, int64_t ldb,
                                                                const float* __restrict__ xn, int pieces,
                                                                const uint16_t* __restrict__ qb,
                                                                const uint64_t* __restrict__ cand, int B, int nq,
                                                                uint64_t* __restrict__ out) {
  constexpr int R = 64 / G;             // rows per load instruction
  constexpr int NL = G < 16 ? G : 16;   // loads issued before use
  __shared__ float sdot[kLshBlock];
  const int q = blockIdx.x;
  if (q >= nq) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int p = lane % G, grp = lane / G;
  const bool live = p < pieces;
  const int pc = live ? p : 0;  // lanes past the row re-read piece 0 and keep +0
  float qf[8];
  bf16_piece_unpack(*reinterpret_cast<const uint4*>(qb + (int64_t)q * ldb + 8 * pc), qf);
  const int per = (B + 3) / 4;
  const int c0 = min(B, w * per), c1 = min(B, c0 + per);
  const uint64_t* src = cand + (int64_t)q * B;
  uint64_t* dst = out + (int64_t)q * B;
  float* const wdot = sdot + w * 64;
  for (int cb = c0; cb < c1; cb += 64) {
    const int nc = min(64, c1 - cb);
    const uint64_t word = lane < nc ? src[cb + lane] : kEmptyCand;
    const int32_t myidx = word != kEmptyCand ? cand_idx(word) : -1;
    const uint64_t present = __ballot(myidx >= 0);
    if (present == 0) {
      if (lane < nc) dst[cb + lane] = kEmptyCand;
      continue;
    }
    const int32_t any_row = __shfl(myidx, __ffsll((unsigned long long)present) - 1, 64);
    const int32_t rowi = myidx >= 0 ? myidx : any_row;
    const float xnr = xn[rowi];
#pragma unroll
    for (int sb = 0; sb < G / NL; ++sb) {
      uint4 v[NL];
#pragma unroll
      for (int l = 0; l < NL; ++l) {
        const int64_t ri = __shfl(rowi, (sb * NL + l) * R + grp, 64);
        v[l] = *reinterpret_cast<const uint4*>(xb + ri * ldb + 8 * pc);
      }
#pragma unroll
      for (int l = 0; l < NL; ++l) {
        float acc = bf16_piece_dot(v[l], qf, 0.0f);
        acc = live ? acc : 0.0f;
#pragma unroll
        for (int m = G / 2; m >= 1; m >>= 1) acc += __shfl_xor(acc, m, 64);
        if (p == 0) wdot[(sb * NL + l) * R + grp] = acc;
      }
    }
    __builtin_amdgcn_wave_barrier();
    if (lane < nc) dst[cb + lane] = myidx >= 0 ? pack_cand(fmaf(-2.0f, wdot[lane], xnr), myidx) : kEmptyCand;
    __builtin_amdgcn_wave_barrier();
  }
}

void launch_prerank(const uint16_t* xb, int64_t ldb, const float* xn, const uint16_t* qb, const uint64_t* cand,
                    int B, int nq, uint64_t* out, hipStream_t st) {
  const int pieces = (int)(ldb / 8);  // <= 32: the index's dimension is at most 256
  if (pieces < 1 || pieces > 32) fail(RPLSH_ERR_UNSUPPORTED, "prerank: row of %d pieces (at most 32)", pieces);
  int G = 1;
  while (G < pieces) G <<= 1;
  const dim3 grid((unsigned)nq), block(kLshBlock);
  switch (G) {
    case 1: lsh_prerank_kernel<1><<<grid, block, 0, st>>>(xb, ldb, xn, pieces, qb, cand, B, nq, out); break;
    case 2: lsh_prerank_kernel<2><<<grid, block, 0, st>>>(xb, ldb, xn, pieces, qb, cand, B, nq, out); break;
    case 4: lsh_prerank_kernel<4><<<grid, block, 0, st>>>(xb, ldb, xn, pieces, qb, cand, B, nq, out); break;
    case 8: lsh_prerank_kernel<8><<<grid, block, 0, st>>>(xb, ldb, xn, pieces, qb, cand, B, nq, out); break;
    case 16: lsh_prerank_kernel<16><<<grid, block, 0, st>>>(xb, ldb, xn, pieces, qb, cand, B, nq, out); break;
    default: lsh_prerank_kernel<32><<<grid, block, 0, st>>>(xb, ldb, xn, pieces, qb, cand, B, nq, out); break;
  }
  RPLSH_HIP(hipGetLastError());
}

}  // namespace
}  // namespace rplsh

using namespace rplsh;

struct rplsh_lsh {
  rplsh_resources* res = nullptr;
  int dim = 0;
  rplsh_lsh_config cfg{};
  int64_t n = 0, capacity = 0, ldx = 0;
  float* x = nullptr;         // [capacity][ldx], zero padded
  float* xn = nullptr;        // [capacity]
  uint16_t* xb = nullptr;     // [capacity][ldb] BF16 rows (pre-rank only)
  int64_t ldb = 0;
  float* w = nullptr;         // [L*K][dim]
  float* wj = nullptr;        // [L][dim][KB]: W j-major per table, zero padded to KB (hash kernel)
  float* off = nullptr;       // [L*K] hyperplane offsets (0, or medians set at the first add)
  bool off_set = false;
  uint32_t* codes = nullptr;  // [L][capacity] codes in row order
  uint32_t* sorted = nullptr; // [L][capacity]
  int32_t* dir = nullptr;     // [L][2^min(K,16) + 1] bucket directory
  int32_t* perm = nullptr;    // [L][capacity]
  DevBuf iota, sort_tmp, ctr_dots, ctr_sorted, ctr_seg, qstage, qpad, qn, qbf, scored, topk, ktau, kbound,
      kcount, qcodes, qmarg, cand, idev, ddev, counters, status, relk, fp64;
  rplsh_lsh_stats last{};
};

namespace {

hipStream_t lsh_stream(rplsh_lsh* h) {
  void* s = nullptr;
  if (rplsh_resources_default_stream(h->res, h->cfg.device, &s) != RPLSH_OK)
    fail(RPLSH_ERR_STATE, "default stream: %s", rplsh_last_error());
  return (hipStream_t)s;
}

void lsh_free(rplsh_lsh* h) {
  for (void* p : {(void*)h->x, (void*)h->xn, (void*)h->xb, (void*)h->codes, (void*)h->sorted, (void*)h->perm})
    if (p) RPLSH_HIP(hipFree(p));
  h->x = h->xn = nullptr;
  h->xb = nullptr;
  h->codes = h->sorted = nullptr;
  h->perm = nullptr;
  h->n = h->capacity = 0;
}

// Offsets back to zero (reset, create): with `center` the next add sets them.
void clear_offsets(rplsh_lsh* h) {
  RPLSH_HIP(hipMemset(h->off, 0, (size_t)h->cfg.tables * h->cfg.bits * sizeof(float)));
  h->off_set = false;
}

void launch_hash(const rplsh_lsh* h, const float* rows, int64_t n, uint32_t* codes, int64_t ldc,
                 float* margins, int64_t mrow, int64_t mcol, hipStream_t st) {
  if (n <= 0) return;
  const int L = h->cfg.tables, K = h->cfg.bits;
  const size_t lds = (size_t)kHashRows * (h->dim + 1) * sizeof(float);
  // Few row groups (a query batch): one table per wave over a second grid
  // dimension, so the launch still fills the device.
  const int64_t groups = div_up(n, kHashRows);
  const unsigned gy = groups < 2048 ? (unsigned)div_up(L, kHashWaves) : 1u;
  const dim3 grid((unsigned)groups, gy), block(kHashRows * kHashWaves);
  switch (hash_width(K)) {
    case 8:
      lsh_hash_kernel<8><<<grid, block, lds, st>>>(rows, h->ldx, n, h->dim, h->wj, h->off, L, K, codes, ldc,
                                                   margins, mrow, mcol);
      break;
    case 16:
      lsh_hash_kernel<16><<<grid, block, lds, st>>>(rows, h->ldx, n, h->dim, h->wj, h->off, L, K, codes, ldc,
                                                    margins, mrow, mcol);
      break;
    default:
      lsh_hash_kernel<24><<<grid, block, lds, st>>>(rows, h->ldx, n, h->dim, h->wj, h->off, L, K, codes, ldc,
                                                    margins, mrow, mcol);
      break;
  }
  RPLSH_HIP(hipGetLastError());
}

// With `center`: offsets = medians of the projections over the first
// min(n, kLshCenterSample) rows (off must be zero when this runs).
void set_offsets(rplsh_lsh* h, const float* rows, int64_t n, hipStream_t st) {
  const int nr = h->cfg.tables * h->cfg.bits;
  const int64_t m = std::min<int64_t>(n, kLshCenterSample);
  float* dots = h->ctr_dots.as<float>((size_t)nr * m);
  float* sorted = h->ctr_sorted.as<float>((size_t)nr * m);
  int* seg = h->ctr_seg.as<int>((size_t)nr + 1);
  launch_hash(h, rows, m, nullptr, 0, dots, 1, m, st);  // dots[r * m + i]
  lsh_segment_offsets_kernel<<<(unsigned)div_up(nr + 1, 256), 256, 0, st>>>(seg, nr, (int)m);
  RPLSH_HIP(hipGetLastError());
  size_t tmp = 0;
  RPLSH_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, tmp, dots, sorted, (int)(nr * m), nr, seg,
                                                       seg + 1, 0, 32, st));
  void* tp = h->sort_tmp.ensure(tmp);
  RPLSH_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(tp, tmp, dots, sorted, (int)(nr * m), nr, seg, seg + 1,
                                                       0, 32, st));
  lsh_median_kernel<<<(unsigned)div_up(nr, 256), 256, 0, st>>>(sorted, m, nr, h->off);
  RPLSH_HIP(hipGetLastError());
}

// Queries (host or device, [nq][dim]) as device rows [nq][ldx], zero padded.
float* stage_queries(rplsh_lsh* h, int64_t nq, const float* queries, hipStream_t st) {
  const float* src = queries;
  if (pointer_device(queries) != h->cfg.device) {
    float* t = h->qstage.as<float>((size_t)nq * h->dim);
    RPLSH_HIP(hipMemcpyAsync(t, queries, (size_t)nq * h->dim * sizeof(float), hipMemcpyDefault, st));
    src = t;
  }
  float* qp = h->qpad.as<float>((size_t)nq * h->ldx);
  launch_pad_rows(src, nq, h->dim, h->dim, (int)h->ldx, qp, st);
  return qp;
}

// Sorts every table's (code, row) pairs of rows [0, n) by code, stably.
void build_tables(rplsh_lsh* h, hipStream_t st) {
  const int L = h->cfg.tables;
  int32_t* rows = h->iota.as<int32_t>((size_t)h->n);
  iota_kernel<<<(unsigned)std::min<int64_t>(div_up(h->n, 256), 65536), 256, 0, st>>>(rows, h->n);
  RPLSH_HIP(hipGetLastError());
  size_t tmp = 0;
  RPLSH_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, h->codes, h->sorted, rows, h->perm,
                                               (int)h->n, 0, h->cfg.bits, st));
  void* tp = h->sort_tmp.ensure(tmp);
  for (int t = 0; t < L; ++t) {
    const int64_t o = (int64_t)t * h->capacity;
    size_t tb = tmp;
    RPLSH_HIP(hipcub::DeviceRadixSort::SortPairs(tp, tb, h->codes + o, h->sorted + o, rows, h->perm + o,
                                                 (int)h->n, 0, h->cfg.bits, st));
  }
  const int64_t entries = (((int64_t)1 << dir_bits(h->cfg.bits)) + 1) * L;
  lsh_directory_kernel<<<(unsigned)div_up(entries, 256), 256, 0, st>>>(h->sorted, h->capacity, h->n, L,
                                                                       h->cfg.bits, h->dir);
  RPLSH_HIP(hipGetLastError());
}

}  // namespace

extern "C" {

void rplsh_lsh_config_default(rplsh_lsh_config* cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->device = 0;
  cfg->tables = 16;
  cfg->bits = 16;
  cfg->probes = 0;
  cfg->max_candidates = 1024;
  cfg->seed = 1234;
  cfg->center = 1;
  cfg->prerank = 0;
}

rplsh_status rplsh_lsh_create(rplsh_resources* res, int dim, const rplsh_lsh_config* cfg, rplsh_lsh** out) {
  return guard([&] {
    RPLSH_REQUIRE(out, "out is NULL");
    RPLSH_REQUIRE(res, "resources is NULL");
    RPLSH_REQUIRE(dim > 0 && dim <= 256, "dimension must be in [1, 256] (got %d)", dim);
    rplsh_lsh_config c;
    rplsh_lsh_config_default(&c);
    if (cfg) c = *cfg;
    int count = 0;
    rplsh_device_count(&count);
    RPLSH_REQUIRE(c.device >= 0 && c.device < count, "device %d out of range [0, %d)", c.device, count);
    RPLSH_REQUIRE(c.tables >= 1 && c.bits >= 1 && c.bits <= kLshMaxBits,
                  "tables >= 1 and bits in [1, %d] required", kLshMaxBits);
    RPLSH_REQUIRE(c.probes >= 0 && c.probes <= c.bits, "probes must be in [0, bits]");
    RPLSH_REQUIRE((int64_t)(c.probes + 1) * c.tables <= kLshMaxEntries,
                  "(probes + 1) * tables must be <= %d", kLshMaxEntries);
    RPLSH_REQUIRE(c.max_candidates >= 1 && c.max_candidates <= kLshMaxBudget,
                  "max_candidates must be in [1, %d]", kLshMaxBudget);
    RPLSH_REQUIRE(c.prerank >= 0 && c.prerank <= 64, "prerank must be in [0, 64]");
    RPLSH_REQUIRE(c.max_candidates <= kLshMaxExact || c.prerank > 0,
                  "max_candidates above %d needs a pre-rank (the exact re-rank takes at most %d)",
                  kLshMaxExact, kLshMaxExact);
    if (rplsh_resources_init_device(res, c.device) != RPLSH_OK)
      fail(RPLSH_ERR_STATE, "%s", rplsh_last_error());
    DeviceScope scope(c.device);
    auto* h = new rplsh_lsh();
    h->res = res;
    h->dim = dim;
    h->cfg = c;
    h->ldx = round_up(dim, 4);
    h->ldb = round_up(dim, 8);
    // matrixRandomInitialize (matrix_ops.cu:71-76) with mean 0, stddev 1: the
    // signs are what matter, so the scale is left at 1.
    const int64_t rows = (int64_t)c.tables * c.bits;
    std::vector<float> wh((size_t)rows * dim);
    if (rplsh_gen_gaussian(wh.data(), rows, dim, c.seed, 0) != RPLSH_OK) {
      delete h;
      fail(RPLSH_ERR_STATE, "projection generator: %s", rplsh_last_error());
    }
    try {
      RPLSH_HIP(hipMalloc(&h->w, wh.size() * sizeof(float)));
      RPLSH_HIP(hipMemcpy(h->w, wh.data(), wh.size() * sizeof(float), hipMemcpyHostToDevice));
      const int kb = hash_width(c.bits);
      std::vector<float> wjh((size_t)c.tables * dim * kb, 0.0f);
      for (int t = 0; t < c.tables; ++t)
        for (int b = 0; b < c.bits; ++b)
          for (int j = 0; j < dim; ++j)
            wjh[((size_t)t * dim + j) * kb + b] = wh[((size_t)t * c.bits + b) * dim + j];
      RPLSH_HIP(hipMalloc(&h->wj, wjh.size() * sizeof(float)));
      RPLSH_HIP(hipMemcpy(h->wj, wjh.data(), wjh.size() * sizeof(float), hipMemcpyHostToDevice));
      RPLSH_HIP(hipMalloc(&h->off, (size_t)rows * sizeof(float)));
      clear_offsets(h);
      RPLSH_HIP(hipMalloc(&h->dir, (((size_t)1 << dir_bits(c.bits)) + 1) * c.tables * sizeof(int32_t)));
    } catch (...) {
      if (h->w) (void)hipFree(h->w);
      if (h->wj) (void)hipFree(h->wj);
      if (h->off) (void)hipFree(h->off);
      if (h->dir) (void)hipFree(h->dir);
      delete h;
      throw;
    }
    *out = h;
  });
}

rplsh_status rplsh_lsh_destroy(rplsh_lsh* h) {
  return guard([&] {
    if (!h) return;
    DeviceScope scope(h->cfg.device);
    (void)hipDeviceSynchronize();
    lsh_free(h);
    if (h->w) RPLSH_HIP(hipFree(h->w));
    if (h->wj) RPLSH_HIP(hipFree(h->wj));
    if (h->off) RPLSH_HIP(hipFree(h->off));
    if (h->dir) RPLSH_HIP(hipFree(h->dir));
    delete h;
  });
}

rplsh_status rplsh_lsh_reset(rplsh_lsh* h) {
  return guard([&] {
    RPLSH_REQUIRE(h, "index is NULL");
    DeviceScope scope(h->cfg.device);
    RPLSH_HIP(hipStreamSynchronize(lsh_stream(h)));
    lsh_free(h);
    clear_offsets(h);
  });
}

rplsh_status rplsh_lsh_add(rplsh_lsh* h, int64_t num, const float* data) {
  return guard([&] {
    RPLSH_REQUIRE(h, "index is NULL");
    if (num <= 0) return;
    RPLSH_REQUIRE(data, "data is NULL");
    RPLSH_REQUIRE(h->n + num < (1ll << 31), "index size limited to 2^31 rows");
    DeviceScope scope(h->cfg.device);
    hipStream_t st = lsh_stream(h);
    const int L = h->cfg.tables;
    const int64_t need = h->n + num;
    if (need > h->capacity) {
      const int64_t cap = h->capacity == 0 ? need : std::max(need, h->capacity + h->capacity / 2);
      float *nx = nullptr, *nn = nullptr;
      uint32_t *nc = nullptr, *ns = nullptr;
      int32_t* np = nullptr;
      uint16_t* nb = nullptr;
      if (h->cfg.prerank > 0) RPLSH_HIP(hipMalloc(&nb, (size_t)cap * h->ldb * sizeof(uint16_t)));
      RPLSH_HIP(hipMalloc(&nx, (size_t)cap * h->ldx * sizeof(float)));
      RPLSH_HIP(hipMalloc(&nn, (size_t)cap * sizeof(float)));
      RPLSH_HIP(hipMalloc(&nc, (size_t)cap * L * sizeof(uint32_t)));
      RPLSH_HIP(hipMalloc(&ns, (size_t)cap * L * sizeof(uint32_t)));
      RPLSH_HIP(hipMalloc(&np, (size_t)cap * L * sizeof(int32_t)));
      if (h->n > 0) {
        RPLSH_HIP(hipMemcpyAsync(nx, h->x, (size_t)h->n * h->ldx * sizeof(float), hipMemcpyDeviceToDevice, st));
        RPLSH_HIP(hipMemcpyAsync(nn, h->xn, (size_t)h->n * sizeof(float), hipMemcpyDeviceToDevice, st));
        if (nb) RPLSH_HIP(hipMemcpyAsync(nb, h->xb, (size_t)h->n * h->ldb * sizeof(uint16_t), hipMemcpyDeviceToDevice, st));
        RPLSH_HIP(hipMemcpy2DAsync(nc, cap * sizeof(uint32_t), h->codes, h->capacity * sizeof(uint32_t),
                                   h->n * sizeof(uint32_t), L, hipMemcpyDeviceToDevice, st));
      }
      RPLSH_HIP(hipStreamSynchronize(st));
      const int64_t n_keep = h->n;
      lsh_free(h);
      h->x = nx;
      h->xn = nn;
      h->xb = nb;
      h->codes = nc;
      h->sorted = ns;
      h->perm = np;
      h->capacity = cap;
      h->n = n_keep;
    }
    float* dst = h->x + h->n * h->ldx;
    if (h->ldx == h->dim) {
      RPLSH_HIP(hipMemcpyAsync(dst, data, (size_t)num * h->dim * sizeof(float), hipMemcpyDefault, st));
    } else {
      RPLSH_HIP(hipMemsetAsync(dst, 0, (size_t)num * h->ldx * sizeof(float), st));
      RPLSH_HIP(hipMemcpy2DAsync(dst, h->ldx * sizeof(float), data, h->dim * sizeof(float),
                                 h->dim * sizeof(float), num, hipMemcpyDefault, st));
    }
    launch_l2norm(dst, num, h->dim, h->ldx, h->xn + h->n, 1, st);
    if (h->xb) launch_encode_bf16(dst, num, h->dim, h->ldx, (int)h->ldb, h->xb + h->n * h->ldb, st);
    if (h->cfg.center && !h->off_set) set_offsets(h, dst, num, st);
    h->off_set = true;
    launch_hash(h, dst, num, h->codes + h->n, h->capacity, nullptr, 0, 0, st);
    h->n = need;
    build_tables(h, st);
    RPLSH_HIP(hipStreamSynchronize(st));
  });
}

rplsh_status rplsh_lsh_search(rplsh_lsh* h, int64_t nq, const float* queries, int k, int32_t* indices,
                              float* distances) {
  return guard([&] {
    RPLSH_REQUIRE(h, "index is NULL");
    if (nq <= 0) return;
    RPLSH_REQUIRE(queries && indices && distances, "NULL pointer");
    RPLSH_REQUIRE(k > 0, "k must be > 0 (got %d)", k);
    {
      const int64_t B0 = h->cfg.max_candidates;
      const int64_t kpre0 = h->cfg.prerank > 0 ? std::min<int64_t>((int64_t)h->cfg.prerank * k, B0) : B0;
      if (kpre0 > kLshMaxExact)
        fail(RPLSH_ERR_UNSUPPORTED, "exact re-rank of %lld candidates per query (at most %d: lower prerank * k)",
             (long long)kpre0, kLshMaxExact);
    }
    RPLSH_REQUIRE(nq < (1ll << 31) && (int64_t)nq * k < (1ll << 40), "batch too large");
    DeviceScope scope(h->cfg.device);
    hipStream_t st = lsh_stream(h);
    const int dev = h->cfg.device;
    const int L = h->cfg.tables, K = h->cfg.bits, B = h->cfg.max_candidates;
    rplsh_lsh_stats stats{};
    stats.queries = nq;
    hipEvent_t ev[4];
    for (auto& e : ev) RPLSH_HIP(hipEventCreate(&e));
    RPLSH_HIP(hipEventRecord(ev[0], st));

    const bool i_dev = pointer_device(indices) == dev;
    const bool d_dev = pointer_device(distances) == dev;
    int32_t* id = i_dev ? indices : h->idev.as<int32_t>((size_t)nq * k);
    float* dd = d_dev ? distances : h->ddev.as<float>((size_t)nq * k);
    unsigned long long* ctr = h->counters.as<unsigned long long>(2);
    RPLSH_HIP(hipMemsetAsync(ctr, 0, 2 * sizeof(unsigned long long), st));

    if (h->n == 0) {
      launch_fill_i32(id, nq * k, -1, st);
      launch_fill_f32(dd, nq * k, FLT_MAX, st);
      RPLSH_HIP(hipEventRecord(ev[1], st));
      RPLSH_HIP(hipEventRecord(ev[2], st));
    } else {
      // queries: padded rows (16-byte aligned, zero padding), norms, codes and margins
      float* qp = stage_queries(h, nq, queries, st);
      float* qn = h->qn.as<float>((size_t)nq);
      launch_l2norm(qp, nq, h->dim, h->ldx, qn, 1, st);
      uint32_t* qc = h->qcodes.as<uint32_t>((size_t)nq * L);
      float* qm = h->qmarg.as<float>((size_t)nq * L * K);
      launch_hash(h, qp, nq, qc, nq, qm, (int64_t)L * K, 1, st);
      // candidates
      uint64_t* cand = h->cand.as<uint64_t>((size_t)nq * B);
      LshGatherArgs ga{};
      ga.dir = h->dir;
      ga.sorted = h->sorted;
      ga.perm = h->perm;
      ga.ld = h->capacity;
      ga.n = h->n;
      ga.qcodes = qc;
      ga.qmarg = qm;
      ga.nq = (int)nq;
      ga.L = L;
      ga.K = K;
      ga.probes = h->cfg.probes;
      ga.budget = B;
      ga.cand = cand;
      ga.raw_total = ctr;
      ga.distinct_total = ctr + 1;
      ga.table = lsh_table_slots(B);
      lsh_gather_kernel<<<(unsigned)nq, kLshBlock, (size_t)ga.table * sizeof(int32_t), st>>>(ga);
      RPLSH_HIP(hipGetLastError());
      RPLSH_HIP(hipEventRecord(ev[1], st));
      // exact re-rank of the distinct candidates (no certificate, no FP64 band)
      int* status = h->status.as<int>((size_t)nq);
      float* relk = h->relk.as<float>((size_t)nq);
      int* fp64c = h->fp64.as<int>(1);
      launch_fill_i32(status, nq, 0, st);
      launch_fill_i32(fp64c, 1, 0, st);
      RerankArgs ra{};
      ra.x = h->x;
      ra.ldx = h->ldx;
      ra.xnorm = h->xn;
      ra.d = h->dim;
      ra.q = qp;
      ra.ldq = h->ldx;
      ra.qnorm = qn;
      const int kpre = h->cfg.prerank > 0 ? (int)std::min<int64_t>((int64_t)h->cfg.prerank * k, B) : B;
      if (kpre < B) {
        // BF16 pre-rank of the B candidates, then the kpre smallest words
        uint16_t* qbf = h->qbf.as<uint16_t>((size_t)nq * h->ldb);
        launch_encode_bf16(qp, nq, h->dim, h->ldx, (int)h->ldb, qbf, st);
        uint64_t* scored = h->scored.as<uint64_t>((size_t)nq * B);
        launch_prerank(h->xb, h->ldb, h->xn, qbf, cand, B, (int)nq, scored, st);
        int* kcount = h->kcount.as<int>((size_t)nq);
        float* ktau = h->ktau.as<float>((size_t)nq);
        float* kbound = h->kbound.as<float>((size_t)nq);
        uint64_t* top = h->topk.as<uint64_t>((size_t)nq * kpre);
        launch_fill_i32(kcount, nq, B, st);
        launch_fill_f32(ktau, nq, INFINITY, st);
        SelectArgs sa{};
        sa.cand = scored;
        sa.counts = kcount;
        sa.cap = B;
        sa.nq = (int)nq;
        sa.K = kpre;
        sa.topk_mode = 1;
        sa.tau_in = ktau;
        sa.topk_out = top;
        sa.bound_out = kbound;
        launch_select(sa, st);
        cand = top;
      }
      ra.topk = cand;
      ra.kp = kpre;
      ra.k = k;
      ra.nq = (int)nq;
      ra.exact = 1;
      ra.bound = nullptr;
      ra.fp64_boundary = 0;
      ra.out_idx = id;
      ra.out_dist = dd;
      ra.status = status;
      ra.relk_out = relk;
      ra.fp64_count = fp64c;
      launch_rerank(ra, st);
      RPLSH_HIP(hipEventRecord(ev[2], st));
    }
    if (!i_dev) RPLSH_HIP(hipMemcpyAsync(indices, id, (size_t)nq * k * sizeof(int32_t), hipMemcpyDefault, st));
    if (!d_dev) RPLSH_HIP(hipMemcpyAsync(distances, dd, (size_t)nq * k * sizeof(float), hipMemcpyDefault, st));
    unsigned long long hc[2] = {0, 0};
    RPLSH_HIP(hipMemcpyAsync(hc, ctr, sizeof(hc), hipMemcpyDeviceToHost, st));
    RPLSH_HIP(hipEventRecord(ev[3], st));
    RPLSH_HIP(hipEventSynchronize(ev[3]));
    float ms = 0.0f;
    RPLSH_HIP(hipEventElapsedTime(&ms, ev[0], ev[3]));
    stats.ms_total = ms;
    RPLSH_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    stats.ms_candidates = ms;
    RPLSH_HIP(hipEventElapsedTime(&ms, ev[1], ev[2]));
    stats.ms_rerank = ms;
    stats.raw_candidates = (int64_t)hc[0];
    stats.distinct_candidates = (int64_t)hc[1];
    for (auto& e : ev) (void)hipEventDestroy(e);
    h->last = stats;
  });
}

rplsh_status rplsh_lsh_last_stats(const rplsh_lsh* h, rplsh_lsh_stats* out) {
  return guard([&] {
    RPLSH_REQUIRE(h && out, "NULL argument");
    *out = h->last;
  });
}

rplsh_status rplsh_lsh_num_data(const rplsh_lsh* h, int64_t* n) {
  return guard([&] {
    RPLSH_REQUIRE(h && n, "NULL argument");
    *n = h->n;
  });
}

rplsh_status rplsh_lsh_get_config(const rplsh_lsh* h, rplsh_lsh_config* out) {
  return guard([&] {
    RPLSH_REQUIRE(h && out, "NULL argument");
    *out = h->cfg;
  });
}

rplsh_status rplsh_lsh_projections(const rplsh_lsh* h, float* out) {
  return guard([&] {
    RPLSH_REQUIRE(h && out, "NULL argument");
    DeviceScope scope(h->cfg.device);
    RPLSH_HIP(hipMemcpy(out, h->w, (size_t)h->cfg.tables * h->cfg.bits * h->dim * sizeof(float),
                        hipMemcpyDeviceToHost));
  });
}

rplsh_status rplsh_lsh_offsets(const rplsh_lsh* h, float* out) {
  return guard([&] {
    RPLSH_REQUIRE(h && out, "NULL argument");
    DeviceScope scope(h->cfg.device);
    RPLSH_HIP(hipMemcpy(out, h->off, (size_t)h->cfg.tables * h->cfg.bits * sizeof(float), hipMemcpyDeviceToHost));
  });
}

rplsh_status rplsh_lsh_tables(const rplsh_lsh* h, uint32_t* codes, uint32_t* sorted, int32_t* perm) {
  return guard([&] {
    RPLSH_REQUIRE(h, "index is NULL");
    DeviceScope scope(h->cfg.device);
    const size_t row = (size_t)h->n;
    for (int t = 0; t < h->cfg.tables && row > 0; ++t) {
      const size_t o = (size_t)t * h->capacity;
      if (codes) RPLSH_HIP(hipMemcpy(codes + t * row, h->codes + o, row * 4, hipMemcpyDeviceToHost));
      if (sorted) RPLSH_HIP(hipMemcpy(sorted + t * row, h->sorted + o, row * 4, hipMemcpyDeviceToHost));
      if (perm) RPLSH_HIP(hipMemcpy(perm + t * row, h->perm + o, row * 4, hipMemcpyDeviceToHost));
    }
  });
}

rplsh_status rplsh_lsh_hash_queries(rplsh_lsh* h, int64_t nq, const float* queries, uint32_t* codes,
                                    float* margins) {
  return guard([&] {
    RPLSH_REQUIRE(h && queries && codes, "NULL argument");
    if (nq <= 0) return;
    DeviceScope scope(h->cfg.device);
    hipStream_t st = lsh_stream(h);
    const int L = h->cfg.tables, K = h->cfg.bits;
    float* qp = stage_queries(h, nq, queries, st);
    uint32_t* qc = h->qcodes.as<uint32_t>((size_t)nq * L);
    float* qm = h->qmarg.as<float>((size_t)nq * L * K);
    launch_hash(h, qp, nq, qc, nq, qm, (int64_t)L * K, 1, st);
    RPLSH_HIP(hipMemcpyAsync(codes, qc, (size_t)nq * L * sizeof(uint32_t), hipMemcpyDefault, st));
    if (margins)
      RPLSH_HIP(hipMemcpyAsync(margins, qm, (size_t)nq * L * K * sizeof(float), hipMemcpyDefault, st));
    RPLSH_HIP(hipStreamSynchronize(st));
  });
}

}  // extern "C"
