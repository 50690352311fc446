// Host launchers of the device kernels (internal C++ interface between the
// translation units of libcurplsh_amd; nothing here crosses the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.hpp"

namespace rplsh {

// norm.hip
void launch_l2norm(const float* x, int64_t n, int d, int64_t ldx, float* out, int squared,
                   hipStream_t st);
void launch_encode_bf16(const float* x, int64_t n, int d, int64_t ldx, int ld, uint16_t* out,
                        hipStream_t st);
void launch_encode_e4m3(const float* x, int64_t n, int d, int64_t ldx, int ld, uint8_t* out,
                        float* scales, hipStream_t st);
void launch_fill_f32(float* p, int64_t n, float v, hipStream_t st);
void launch_fill_i32(int* p, int64_t n, int v, hipStream_t st);
void launch_pad_rows(const float* x, int64_t n, int d, int64_t ldx, int ld, float* out,
                     hipStream_t st);
// out must be zero (or a smaller non-negative value) before the call.
void launch_max_reduce(const float* v, int64_t n, float* out, hipStream_t st);

// filter.hip: element kinds of the candidate contraction.
enum class Elem { F32 = 0, BF16 = 1, FP8 = 2 };

// Padded row length (elements) the filter kernels use for dimension d, or 0
// if d is not supported: up to 256 for every element kind (register-resident
// query fragments), above that FP32 only (K-looped kernel) up to kMaxKLoopDim.
constexpr int kMaxKLoopDim = 4096;
int filter_dpad(Elem e, int d);
// Queries per workgroup of the filter kernel for (e, dpad).
int filter_block_queries(Elem e, int dpad);
void launch_filter(Elem e, int dpad, const FilterArgs& a, hipStream_t st);

// filter_packed.hip: the BF16 / FP8 candidate copies in 32-row MFMA tiles.
int64_t packed_tile_bytes(Elem e, int dpad);
// Variants of the packed filter (chosen by the autotuner): LDS-staged with
// wide or narrow query tiles per wave, or direct-to-register loads.
enum FilterVariant : int {
  kFilterLdsWide = 0,
  kFilterLdsNarrow = 1,
  kFilterDirect = 2,
  kFilterRingWide = 3,
  kFilterRingNarrow = 4,
  kFilterLdsWide8 = 5,  // LDS wide with 8-wave workgroups (one per CU)
};
int filter_packed_block_queries(Elem e, int dpad, int variant);
void launch_filter_packed(Elem e, int dpad, int variant, const FilterArgs& a, hipStream_t st);
// Sample copy (sm rows, a multiple of 32, at most 32 * tiles) of an index
// with `tiles` whole 32-row tiles: sample row r holds index row
// sample_source_row(r, step, tiles). snorm[r] = xn[src], snh[r] = -xn[src] / 2,
// sscale[r] = xscale[src] (if sscale), snh[nh_inf] = -inf.
void launch_sample_rows(const float* xn, const float* xscale, int64_t tiles, int64_t sm, int64_t step,
                        float* snorm, float* snh, int64_t nh_inf, float* sscale, hipStream_t st);
// rank[u] = s for the index tile u = (s * step) mod tiles held by sample tile
// s < stiles; INT32_MAX for every other u < rank_tiles.
void launch_tile_ranks(int64_t tiles, int64_t stiles, int64_t step, int32_t* rank, int64_t rank_tiles,
                       hipStream_t st);
// Compacts each query's list to the survivors of the next level's test
// (acc = -key/2 >= max(-tau/2, -FLT_MAX)) from sample rows below row_limit,
// their ids mapped to index rows when step > 0; counts become the carried
// counts.
void launch_carry(const CarryArgs& a, hipStream_t st);
// nh[i] = -xn[i] / 2 for i in [from, n); nh[inf_at] = -inf.
void launch_neg_half_norms(const float* xn, int64_t from, int64_t n, float* nh, int64_t inf_at,
                           hipStream_t st);
void launch_encode_sample(Elem e, int dpad, const float* x, int64_t tiles, int d, int64_t ldx, int64_t sm,
                          int64_t step, void* out, const float* scales, hipStream_t st);
// Encodes rows [row_from, n) (re-encoding the partial tile before row_from);
// FP8 also writes the per-row power-of-two scales of rows [row_from, n).
void launch_encode_packed(Elem e, int dpad, const float* x, int64_t n, int d, int64_t ldx,
                          int64_t row_from, void* out, float* scales, hipStream_t st);

// A level over a cluster-residual copy (the LDS variants, wide or narrow, with
// the residual fields of FilterArgs; any other variant runs the wide one).
int filter_resid_block_queries(Elem e, int dpad, int variant);
void launch_filter_resid(Elem e, int dpad, int variant, const FilterArgs& a, hipStream_t st);

// select.hip
void launch_select(const SelectArgs& a, hipStream_t st);
void launch_rerank(const RerankArgs& a, hipStream_t st);
// Compacts the ids of live queries (< *nq_dev when given) with a non-zero
// status into list, count in *n.
void launch_collect_failed(const int* status, int nq, const int* nq_dev, int* list, int* n,
                           hipStream_t st);
// Sums counts[0..nq) into *out (int64).
void launch_sum_counts(const int* counts, int nq, int cap, long long* out, hipStream_t st);
// Row gather / scatter for re-running a subset of queries.
void launch_gather_rows(const float* src, int64_t ld, const int* rows, int nrows, int d,
                        float* dst, hipStream_t st);
void launch_scatter_results(const int32_t* si, const float* sd, const int* rows, int nrows, int k,
                            int32_t* di, float* dd, hipStream_t st);

}  // namespace rplsh
