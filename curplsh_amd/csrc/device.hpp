// Device-side shared definitions: candidate encoding, MFMA vector types and
// the host-visible launch-argument structs of the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace rplsh {

// A candidate is one 64-bit word: the order-preserving bit pattern of its
// float key in the high half and its row index in the low half. Comparing
// words as unsigned integers is the (key, index) lexicographic order the
// search defines for ties ("lower index first").
__host__ __device__ inline uint32_t key_bits(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__host__ __device__ inline float key_from_bits(uint32_t u) {
  u = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

__host__ __device__ inline uint64_t pack_cand(float key, int32_t idx) {
  return ((uint64_t)key_bits(key) << 32) | (uint32_t)idx;
}

__host__ __device__ inline float cand_key(uint64_t c) { return key_from_bits((uint32_t)(c >> 32)); }
__host__ __device__ inline int32_t cand_idx(uint64_t c) { return (int32_t)(uint32_t)(c & 0xffffffffu); }

constexpr uint64_t kEmptyCand = ~0ull;

// Per-query status bits written by the select / re-rank kernels.
enum StatusBits : int {
  kStatusOverflow = 1,  // more survivors than the candidate capacity
  kStatusShort = 2,     // fewer survivors than the rank asked for
  kStatusCertFail = 4,  // plain path: candidate margin not certified
};

// Arguments of the fused contraction + threshold filter.
struct FilterArgs {
  const void* x;          // database rows, element type of the kernel
  int64_t ldx;            // row stride in elements
  int dvalid_bytes;       // bytes of a row holding data; the rest of the row is zero
  const float* xnorm;     // [n] squared norms (FP32, canonical)
  const float* xscale;    // [n] per-row scale (FP8 only)
  const float* xnh;       // packed ring kernels: -||x||^2 / 2 per row; xnh[xnh_inf] = -inf
  int64_t xnh_inf;
  int64_t m;              // logical rows of this pass
  int64_t stride;         // physical row = logical row * stride
  int64_t rows_per_slice; // logical rows per workgroup slice (multiple of 64)
  const void* q;          // encoded queries [nq_pad][row bytes], zero padded
  const float* qscale;    // [nq] per-row scale (FP8 only)
  const float* tau;       // [nq] threshold on the approximate relative distance
  int nq;
  const int* nq_dev;      // if set: the number of live queries (<= nq), read on device
  int num_qtiles;
  int num_slices;
  int cap;                // candidate slots per query
  int dense;              // packed kernels: every row survives (tau = +inf); row j's word goes to
                          // slot j, counts are preset to m by the caller
  int* counts;            // [nq] survivors (may exceed cap: overflow)
  uint64_t* cand;         // [nq][cap]
  // LDS kernels, full level after a carried sample level: tiles u (32 rows)
  // with u < rank_tiles and tile_rank[u] < skip_tiles were scored by that
  // level and are skipped (nullptr: none).
  const int32_t* tile_rank;
  int64_t rank_tiles;
  int64_t skip_tiles;
  // First logical row (a multiple of 32): a sample level after a carried one
  // scores rows [row_begin, m) only; the slices cover m - row_begin rows.
  int64_t row_begin;
  // Cluster-residual copies (LDS kernels with the residual flag, resid.hip):
  // x holds r = x - c in tiles of one centroid, xnorm holds -||r||^2 / 2
  // (-inf: padding row), xscale the scales of r; slice s covers tiles
  // [slice_t0[s], slice_t0[s + 1]) of centroid slice_cl[s]; the queries are
  // FP32 (qf32, ld ldq, d elements) and are encoded relative to the slice's
  // centroid in the kernel; rid maps a row to the index row it holds.
  const int32_t* rid;
  const int32_t* slice_t0;
  const int32_t* slice_cl;
  const float* cent;
  const float* qf32;
  int ldq;
  int d;
};

// Sample copy of a packed index: whole 32-row tiles, sample tile s holding
// tile (s * step) mod tiles of the index (step coprime with tiles, the number
// of whole tiles), so every prefix is a spread-out set of distinct tiles.
// Index row of sample row r:
__host__ __device__ inline int64_t sample_source_row(int64_t r, int64_t step, int64_t tiles) {
  return (int64_t)(((uint64_t)(r >> 5) * (uint64_t)step) % (uint64_t)tiles) * 32 + (r & 31);
}

// Carry of a sample level's survivors into the next level's lists.
struct CarryArgs {
  uint64_t* cand;        // [nq][cap]: the sample level's lists, compacted in place
  int* counts;           // [nq]: in the sample level's counts, out the carried counts
  int cap;
  int nq;
  const int* nq_dev;     // if set: live queries (<= nq), read on device
  const float* tau;      // [nq] the next level's threshold
  int64_t row_limit;     // rows carried: sample rows below it (the next level skips them)
  int64_t step, tiles;   // step > 0: map sample rows to index rows (sample_source_row)
  int* status;           // [nq] kStatusOverflow when a list overflowed
};

struct SelectArgs {
  const uint64_t* cand;  // [nq][cap]
  const int* counts;     // [nq]
  int cap;
  int nq;
  const int* nq_dev;     // if set: live queries (<= nq), read on device
  int K;                 // rank (threshold mode) or list length (top-K mode)
  int topk_mode;         // 0: tau_out = key of K-th smallest; 1: sorted K smallest
  const float* tau_in;   // [nq] threshold the candidates were filtered with
  float* tau_out;        // [nq]
  uint64_t* topk_out;    // [nq][K]
  float* bound_out;      // [nq] key every non-selected row is >= to
  int* status;           // [nq]
};

struct RerankArgs {
  const float* x;        // FP32 database [n][ldx]
  int64_t ldx;
  const float* xnorm;    // [n]
  int d;
  const float* q;        // FP32 queries [nq][ldq]
  int64_t ldq;
  const float* qnorm;    // [nq]
  const uint64_t* topk;  // [nq][kp] approximate candidates
  int kp;
  int k;
  int nq;
  const int* nq_dev;     // if set: live queries (<= nq), read on device
  int exact;             // add ||q||^2
  const float* bound;    // [nq] certification bound (plain path) or nullptr
  float eps_coeff;       // plain: certified margin = eps_coeff * (||q||^2 + 2 xn_max)
  float band_coeff;      // FP64 band half-width = band_coeff * (||q||^2 + 2 xn_max) + ulp terms
  float xn_max;
  int fp64_boundary;
  int32_t* out_idx;      // [nq][k]
  float* out_dist;       // [nq][k]
  int* status;           // [nq]
  float* relk_out;       // [nq] k-th canonical relative distance (for a retry)
  int* fp64_count;       // total boundary candidates re-ranked in FP64
};

// Persistent serving kernel (persist.hip). Query preparation of one batch:
// canonical FP32 norms, the encoded copy for the candidate contraction
// (BF16, or e4m3 + power-of-two row scale), threshold +inf, status 0, the
// first level's count preset, and the batch counters cleared.
struct PrepArgs {
  const float* q;   // FP32 queries [nq][d] (device)
  int nq;
  int d;
  int dpad;
  int fp8;          // 0: BF16, 1: FP8 e4m3
  float* qnorm;
  void* qenc;       // [nq][dpad]
  float* qscale;    // FP8 only
  float* tau;
  int* status;
  int* counts;
  int count_init;   // counts[q] before the first level (its row count when dense)
  int* ctr;         // BatchCounters, four ints
};

// The re-rank of a batch plus what its failed-query list and survivor total
// need (collect_failed and sum_counts of the launched pipeline).
struct RerankTailArgs {
  RerankArgs ra;
  const int* counts;   // last level's survivors per query
  int* flist;          // failed queries
  int* failed;         // BatchCounters::failed
  long long* survivors;
};

// Longest candidate list the persistent kernel selects (a wave per query).
constexpr int kPersistMaxKp = 512;

// A phase of one batch in a persistent program: `units` consecutive work
// units from `begin`.
enum PersistKind : int { kPkPrep = 0, kPkFilter = 1, kPkSelect = 2, kPkRerank = 3 };
struct PersistPhase {
  int kind;
  int begin;        // first unit
  int units;
  int arg;          // byte offset of its argument record in the program
  int count_reset;  // select phases: counts[q] after the select (-1: unchanged)
  int pad[3];
};
// Program layout: PersistHeader, PersistPhase[nphases], argument records
// (16-byte aligned).
struct PersistHeader {
  int nphases;
  int total_units;
  int pad[2];
};

}  // namespace rplsh
