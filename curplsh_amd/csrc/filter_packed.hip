// Fused contraction + threshold filter over an MFMA-tile-packed database
// (BF16 and FP8 candidate copies). Same arithmetic and filter as
// filter_kernel (filter.hip); the index stores its candidate copy packed by
// 32-row tiles: within a tile, K-chunk c is a contiguous block of 64
// lane-pieces, lane l's piece holding row (l % 32), elements
// [c*KC + (l/32)*KC/2, +KC/2) -- exactly the A operand of
// v_mfma_f32_32x32x16_bf16 (KC = 16) or of v_mfma_scale_f32_32x32x64_f8f6f4
// (KC = 64), so one coalesced 1 KB (BF16) / 2 KB (FP8) read is an operand.
// Each wave holds TN query tiles (32 queries each) as B operands in registers
// for the whole launch, so every A operand feeds TN MFMAs. Three data
// movements:
//
//   * filter_lds_kernel (the default, wide or narrow query tiles): the
//     workgroup streams its slice in stages of up to 4 tiles through LDS,
//     double buffered, one barrier per stage; survivors go to a queue of the
//     wave's own in LDS.
//   * filter_ring_kernel: the same tiles copied by LDS-DMA into a 3-slot ring.
//   * filter_packed_kernel (direct): operands straight from global memory into
//     registers, no LDS staging and no workgroup barrier; the waves of a
//     workgroup share rows through L1.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <cfloat>
#include <cstdint>

#include "common.hpp"
#include "device.hpp"
#include "encode.hpp"
#include "kernels.hpp"
#include "packed_mma.hpp"

namespace rplsh {

namespace {

constexpr int kBlock = 256;

// Direct variant: one wave per SIMD, up to TN = 8 query tiles per wave. Chunk
// c of the next tile is loaded into the operand register of chunk c as soon
// as the current tile's MFMAs on it are issued; -||x||^2/2 of the next tile
// is prefetched at the top of the iteration. The four waves of a workgroup
// take different queries and the same rows, so a tile fetched by one wave is
// an L1 hit for the others.
template <Elem E, int DPAD, int TN>
__global__ __launch_bounds__(kBlock, 1) void filter_packed_kernel(FilterArgs a) {
  using G = Geometry<E, DPAD>;
  constexpr int LB = G::LB;
  constexpr int NC = G::NC;
  constexpr int BQ = 4 * TN * 32;

  const int per_xcd = (int)(gridDim.x >> 3);
  const int logical = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
  if (logical >= a.num_qtiles * a.num_slices) return;
  const int qtile = logical % a.num_qtiles;
  const int slice = logical / a.num_qtiles;
  const int nq = a.nq_dev ? min(*a.nq_dev, a.nq) : a.nq;
  if (qtile * BQ >= nq) return;
  const int64_t r_begin = a.row_begin + (int64_t)slice * a.rows_per_slice;
  int64_t r_end = r_begin + a.rows_per_slice;
  if (r_end > a.m) r_end = a.m;
  if (r_begin >= r_end) return;

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int h = lane >> 5, col = lane & 31;

  Frag<LB> qf[TN][NC];
  float thr[TN];
  int qsc[TN];
  int qrow[TN];
#pragma unroll
  for (int nt = 0; nt < TN; ++nt) {
    const int qr = qtile * BQ + (wave * TN + nt) * 32 + col;
    qrow[nt] = qr;
    if (qr < nq) {
      const char* qp = reinterpret_cast<const char*>(a.q) + (int64_t)qr * G::ROWB;
#pragma unroll
      for (int c = 0; c < NC; ++c) qf[nt][c] = load_frag<LB>(qp + 2 * LB * c + LB * h);
      thr[nt] = fmaxf(-0.5f * a.tau[qr], -FLT_MAX);
      qsc[nt] = E == Elem::FP8 ? e8m0_of(a.qscale[qr]) : 127;
    } else {
#pragma unroll
      for (int c = 0; c < NC; ++c) qf[nt][c] = zero_frag<LB>();
      thr[nt] = INFINITY;
      qsc[nt] = 127;
    }
  }

  const char* xb = reinterpret_cast<const char*>(a.x);
  const int64_t stride = a.stride;
  const int64_t m = a.m;

  // Address of this lane's A operand piece of the 32-row tile starting at
  // logical row j0 (lane row j0 + col), or nullptr past the end.
  auto tile_ptr = [&](int64_t j0) -> const char* {
    const int64_t j = j0 + col;
    if (j >= m) return nullptr;
    const int64_t p = j * stride;
    return xb + (p >> 5) * G::TILE + ((p & 31) + 32 * h) * LB;
  };
  auto tile_scale = [&](int64_t j0) -> int {
    const int64_t j = j0 + col;
    return (E == Elem::FP8 && j < m) ? e8m0_of(a.xscale[j * stride]) : 127;
  };
  // -||x||^2/2 of this lane's output rows j0 + 8*(i/4) + 4*h + (i%4).
  auto load_cinit = [&](int64_t j0) -> f32x16 {
    f32x16 ci;
    if (stride == 1 && j0 + 32 <= m) {
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const f32x4 v = *reinterpret_cast<const f32x4*>(a.xnorm + j0 + 8 * g + 4 * h);
        ci[4 * g + 0] = -0.5f * v[0];
        ci[4 * g + 1] = -0.5f * v[1];
        ci[4 * g + 2] = -0.5f * v[2];
        ci[4 * g + 3] = -0.5f * v[3];
      }
    } else {
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int64_t r = j0 + 8 * (i >> 2) + 4 * h + (i & 3);
        ci[i] = r < m ? -0.5f * a.xnorm[r * stride] : -INFINITY;
      }
    }
    return ci;
  };

  // Single operand buffer: chunk c of the next tile is loaded into af[c] as
  // soon as the current tile's MFMAs on af[c] have been issued.
  Frag<LB> af[NC];
  {
    const char* p0 = tile_ptr(r_begin);
#pragma unroll
    for (int c = 0; c < NC; ++c) af[c] = p0 ? load_frag<LB>(p0 + c * G::CHUNK) : zero_frag<LB>();
  }
  f32x16 ci = load_cinit(r_begin);
  int xs = tile_scale(r_begin);

  for (int64_t j0 = r_begin; j0 < r_end; j0 += 32) {
    const bool has_next = j0 + 32 < r_end;
    const char* pn = has_next ? tile_ptr(j0 + 32) : nullptr;
    f32x16 nci = ci;
    int nxs = xs;
    if (has_next) {
      nci = load_cinit(j0 + 32);
      nxs = tile_scale(j0 + 32);
    }

    f32x16 acc[TN];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
#pragma unroll
      for (int nt = 0; nt < TN; ++nt)
        acc[nt] = mfma_chunk<E>(af[c], qf[nt][c], c == 0 ? ci : acc[nt], xs, qsc[nt]);
      if (has_next) af[c] = pn ? load_frag<LB>(pn + c * G::CHUNK) : zero_frag<LB>();
    }

    float mx[TN];
    bool hit = false;
#pragma unroll
    for (int nt = 0; nt < TN; ++nt) {
      float v = acc[nt][0];
#pragma unroll
      for (int i = 1; i < 16; ++i) v = fmaxf(v, acc[nt][i]);
      mx[nt] = v;
      hit |= v >= thr[nt];
    }
    if (__any(hit)) {
#pragma unroll
      for (int nt = 0; nt < TN; ++nt) {
        if (mx[nt] >= thr[nt]) {
          const int qr = qrow[nt];
          int* cnt = a.counts + qr;
          uint64_t* dst = a.cand + (int64_t)qr * a.cap;
#pragma unroll
          for (int i = 0; i < 16; ++i) {
            const float v = acc[nt][i];
            if (v >= thr[nt]) {
              const int64_t row = j0 + 8 * (i >> 2) + 4 * h + (i & 3);
              const int pos = atomicAdd(cnt, 1);
              if (pos < a.cap) dst[pos] = pack_cand(-2.0f * v, (int32_t)(row * stride));
            }
          }
        }
      }
    }

    ci = nci;
    xs = nxs;
  }
}

// LDS-staged variants (wide / narrow query tiles): the workgroup's 4 waves
// take different queries and share each stage of SM packed 32-row tiles,
// copied linearly into LDS (one coalesced 16 B piece per thread and step;
// gathered row by row on strided levels), double buffered: the global loads
// of stage s+1 are issued before stage s is computed and stored to LDS after
// it, one barrier per stage. Each wave works one 32-row tile at a time with TN
// query tiles, so an A operand read from LDS (consecutive lanes, consecutive
// 16 B: conflict-free) feeds TN MFMAs.
// Survivors go to a queue of the wave's own in LDS, at positions given by
// ballot prefix counts from a wave-uniform fill level (no atomics, no LDS
// round trip in the tile loop); the wave moves it to the per-query lists
// (global atomics) when the next values would not fit, and at the end.
constexpr int kSeg = 256;  // entries per wave queue (LDS variants)
// Word 3 of a raw buffer resource on gfx9-family parts (gfx950): 32-bit
// data format, no swizzle; with stride 0, an offset at or past num_records
// reads as zero.
constexpr int kBufferWord3 = 0x00020000;
// Ring variants: survivors are appended to one of two LDS queues (by stage
// parity) with LDS atomics. After the barrier that ends stage s, the queue of
// stage s is flushed to the per-query lists if more than half full: nobody
// appends to it during stage s+1, so every thread reads the same tail and
// takes the same branch, and it is not reused before the barrier that ends
// stage s+1.
constexpr int kQueue = 512;  // entries per survivor queue (power of two)
constexpr int kHitSlots = 8;  // lanes a wave stages at a time in emit_hits

// Slow path of the filters: emits every value >= its query's threshold in the
// wave's current tile (row0 = its first row), query tile by query tile. The
// lanes that hold one stage their 16 accumulator values in the wave's LDS
// slots hb[kHitSlots][16], kHitSlots lanes at a time; the wave then serves
// four staged lanes per step, 16 lanes each: lane j of group g compares value
// j of staged lane L_g, a ballot gives the hits, one LDS atomic on `tail`
// reserves their queue positions and put() stores them. No per-value branch:
// a sparse tile (the full-database level) costs a few instructions per
// hitting lane instead of a compare-and-branch per accumulator value. Value i
// of lane L is row row0 + 8 (i >> 2) + 4 (L >> 5) + (i & 3) of the query in
// qrow of lane L (the 32x32 MFMA accumulator layout).
template <int TN, typename Put>
__device__ __forceinline__ void emit_hits(const f32x16 (&acc)[TN], const float (&mx)[TN],
                                          const float (&thr)[TN], const int (&qrow)[TN], int64_t row0,
                                          int lane, float* hb, unsigned* tail, Put&& put) {
  const int g = lane >> 4, j = lane & 15;
#pragma unroll
  for (int nt = 0; nt < TN; ++nt) {
    uint64_t B = __ballot(mx[nt] >= thr[nt]);
    while (B != 0) {
      const unsigned rank =
          __builtin_amdgcn_mbcnt_hi((uint32_t)(B >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)B, 0u));
      if ((B >> lane & 1ull) && rank < (unsigned)kHitSlots) {
        f32x4* d = reinterpret_cast<f32x4*>(hb + rank * 16);
        d[0] = f32x4{acc[nt][0], acc[nt][1], acc[nt][2], acc[nt][3]};
        d[1] = f32x4{acc[nt][4], acc[nt][5], acc[nt][6], acc[nt][7]};
        d[2] = f32x4{acc[nt][8], acc[nt][9], acc[nt][10], acc[nt][11]};
        d[3] = f32x4{acc[nt][12], acc[nt][13], acc[nt][14], acc[nt][15]};
      }
      __builtin_amdgcn_wave_barrier();
      const int staged = __popcll(B) < kHitSlots ? __popcll(B) : kHitSlots;
      for (int s0 = 0; s0 < staged; s0 += 4) {
        // staged lanes s0 .. s0 + 3: the lowest set bits of B, one per group
        int lsel = 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int L = B != 0 ? __ffsll((unsigned long long)B) - 1 : 0;  // wave-uniform
          if (u == g) lsel = L;
          if (s0 + u < staged) B &= B - 1;
        }
        const bool live = s0 + g < staged;
        const float thr_l = __shfl(thr[nt], lsel);
        const int qr_l = __shfl(qrow[nt], lsel);
        const float v = live ? hb[(s0 + g) * 16 + j] : -INFINITY;
        const bool hit = v >= thr_l;
        const uint64_t H = __ballot(hit);
        unsigned base = 0u;
        if (lane == 0) base = atomicAdd(tail, (unsigned)__popcll(H));
        base = __builtin_amdgcn_readfirstlane(base);
        if (hit) {
          const unsigned t =
              base + __builtin_amdgcn_mbcnt_hi((uint32_t)(H >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)H, 0u));
          put(t, qr_l, v, row0 + 8 * (j >> 2) + 4 * (lsel >> 5) + (j & 3));
        }
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
}

// Query operands of a workgroup over a cluster-residual slice of centroid
// `cl`: s = q - c for query tile nt of each wave (queries q0 + (wave * TN +
// nt) * 32 + col), encoded as a query row (BF16: round to nearest even; FP8:
// a power-of-two scale per query from max |s|, which keeps every value within
// +-448, then e4m3 by the hardware conversion, round to nearest even), and
// cq = q.c - ||c||^2/2 (FP32, each lane its half of the row in element
// order, the halves added in a fixed order). A lane's fragment elements are
// exactly its half of its query's row (elements c * KC + h * KC/2 + [0,
// KC/2) of chunk c): it loads those elements of q (all loads issued before
// use), takes c from the LDS copy of the centroid, writes the encoded bytes
// to its own place in an LDS row buffer and reads its fragments back from
// there, so no lane waits for another. The LDS area (the staging buffers,
// free before the stage loop, or a buffer of the kernel's own) holds the
// centroid (DPAD floats) and 32 encoded rows per wave. Every thread of the
// workgroup calls it.
template <Elem E, int DPAD, int TN, int NW>
__device__ __forceinline__ void resid_query_operands(const FilterArgs& a, int q0, int cl, int nq, int wave,
                                                     int lane, char* area,
                                                     Frag<PackTraits<E>::kLaneBytes> (&qf)[TN][Geometry<E, DPAD>::NC],
                                                     float (&thr)[TN], int (&qsc)[TN], int (&qrow)[TN],
                                                     float (&cq)[TN]) {
  using G = Geometry<E, DPAD>;
  constexpr int LB = G::LB;
  constexpr int NC = G::NC;
  constexpr int KC = PackTraits<E>::kChunkElems;
  constexpr int HALF = KC / 2;
  constexpr int NV = HALF / 4;  // 4-element groups per chunk and lane
  const int h = lane >> 5, col = lane & 31;
  const int d = a.d;
  float* cen = reinterpret_cast<float*>(area);
  char* erow = area + DPAD * 4 + (wave * 32 + col) * G::ROWB;
  const float* cr = a.cent + (int64_t)cl * d;
  for (int k = threadIdx.x; k < DPAD; k += NW * 64) cen[k] = k < d ? cr[k] : 0.0f;
  __syncthreads();
  const bool vec = (d & 3) == 0 && (a.ldq & 3) == 0;  // 16-byte aligned rows
#pragma unroll
  for (int nt = 0; nt < TN; ++nt) {
    const int qr = q0 + (wave * TN + nt) * 32 + col;
    qrow[nt] = qr;
    const bool live = qr < nq;
    const float* qp = a.qf32 + (int64_t)(live ? qr : 0) * a.ldq;
    f32x4 qv[NC][NV];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const int k0 = c * KC + h * HALF + 4 * v;
        if (vec) {
          qv[c][v] = k0 < d ? *reinterpret_cast<const f32x4*>(qp + k0) : f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        } else {
#pragma unroll
          for (int u = 0; u < 4; ++u) qv[c][v][u] = k0 + u < d ? qp[k0 + u] : 0.0f;
        }
      }
    float mx = 0.0f, dot = 0.0f;
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const int k0 = c * KC + h * HALF + 4 * v;
        const f32x4 cv = *reinterpret_cast<const f32x4*>(cen + k0);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          mx = fmaxf(mx, fabsf(qv[c][v][u] - cv[u]));  // zero past d (both are)
          dot = fmaf(cv[u], qv[c][v][u] - 0.5f * cv[u], dot);
        }
      }
    mx = fmaxf(mx, __shfl_xor(mx, 32, 64));
    const float other = __shfl_xor(dot, 32, 64);
    dot = h == 0 ? dot + other : other + dot;
    const float sc = E == Elem::FP8 ? e4m3_row_scale(mx) : 1.0f;
    const float inv = 1.0f / sc;  // a power of two: exact
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const int k0 = c * KC + h * HALF + 4 * v;
        const f32x4 sv = qv[c][v] - *reinterpret_cast<const f32x4*>(cen + k0);
        if constexpr (E == Elem::FP8) {
          int w = __builtin_amdgcn_cvt_pk_fp8_f32(sv[0] * inv, sv[1] * inv, 0, false);
          w = __builtin_amdgcn_cvt_pk_fp8_f32(sv[2] * inv, sv[3] * inv, w, true);
          *reinterpret_cast<int*>(erow + k0) = w;
        } else {
          const uint32_t w0 = (uint32_t)bf16_rne(sv[0]) | ((uint32_t)bf16_rne(sv[1]) << 16);
          const uint32_t w1 = (uint32_t)bf16_rne(sv[2]) | ((uint32_t)bf16_rne(sv[3]) << 16);
          *reinterpret_cast<uint32_t*>(erow + 2 * k0) = w0;
          *reinterpret_cast<uint32_t*>(erow + 2 * k0 + 4) = w1;
        }
      }
    // The lane reads back only the bytes it wrote.
#pragma unroll
    for (int c = 0; c < NC; ++c) qf[nt][c] = live ? load_frag<LB>(erow + 2 * LB * c + LB * h) : zero_frag<LB>();
    thr[nt] = live ? fmaxf(-0.5f * a.tau[qr], -FLT_MAX) : INFINITY;
    cq[nt] = live ? dot : 0.0f;
    qsc[nt] = E == Elem::FP8 ? e8m0_of(sc) : 127;
  }
  __syncthreads();  // the area becomes the staging buffers
}

template <Elem E, int DPAD, int TN, bool WIDE, int NW>
struct StageShape {
  using G = Geometry<E, DPAD>;
  // Stage tiles: the double-buffered stage within ~48 KB (wide, two 4-wave
  // workgroups per CU), ~36 KB (narrow, three) or ~64 KB (one 8-wave
  // workgroup per CU), at most 4 tiles. Larger stages halve the barriers and
  // stage set-up per tile; a wide BF16 kernel at dpad 96 with 4 query tiles
  // spills a few 64-bit addresses of the strided-stage path then, and is
  // still faster (10M x 96: 16.3 -> 15.4 ms per full level; 2M x 128:
  // 5.50 -> 4.97 ms).
  static constexpr int kBudget = NW == 8 ? 64 * 1024 : (WIDE ? 48 * 1024 : 36 * 1024);
  static constexpr int kMaxSM = 4;  // more tiles per stage need more staging registers
  static constexpr int kFit = kBudget / (2 * G::TILE);
  static constexpr int SM = kFit > kMaxSM ? kMaxSM : (kFit < 1 ? 1 : kFit);
  static constexpr int SR = 32 * SM;  // rows per stage
  static constexpr int STAGE = SM * G::TILE;
  static constexpr int LDS = 2 * STAGE + 2 * SR * 8 + NW * kSeg * 12;
  static constexpr int WAVES_PER_SIMD = (NW == 4 && !WIDE && 3 * LDS <= 160 * 1024) ? 3 : 2;
};

// RESID: a level over a cluster-residual copy (resid.hip): the slice is a
// tile range of one centroid c, the query operands are s = q - c encoded in
// the kernel (FP8: a power-of-two scale per query and centroid), the C
// operand adds q.c - ||c||^2/2 to the copy's -||r||^2/2, and survivors carry
// the index row the residual row holds.
// DENSE: the dense first pass (tau = +inf for every query, a.dense set), which
// writes every row's word to its own slot. It is an instantiation of its own:
// compiled into the same kernel as the filtering loop, its code made that
// loop's setup spill registers to scratch memory (up to 184 bytes per lane).
template <Elem E, int DPAD, int TN, bool WIDE, int NW = 4, bool RESID = false, bool DENSE = false>
__global__ __launch_bounds__(NW * 64, (StageShape<E, DPAD, TN, WIDE, NW>::WAVES_PER_SIMD)) void filter_lds_kernel(
    FilterArgs a) {
  using G = Geometry<E, DPAD>;
  using SS = StageShape<E, DPAD, TN, WIDE, NW>;
  constexpr int NT = NW * 64;  // threads per workgroup
  constexpr int LB = G::LB;
  constexpr int NC = G::NC;
  constexpr int BQ = NW * TN * 32;
  constexpr int SM = SS::SM;
  constexpr int SR = SS::SR;
  constexpr int STAGE = SS::STAGE;
  constexpr int UNITS = STAGE / 16;
  constexpr int UPT = (UNITS + NT - 1) / NT;
  constexpr int TILE_UNITS = G::TILE / 16;
  constexpr int CHUNK_UNITS = G::CHUNK / 16;
  constexpr int LANE_UNITS = LB / 16;
  static_assert(SR <= NT, "one row norm per thread");
  static_assert(!(DENSE && RESID), "residual levels are never dense");

  __shared__ __attribute__((aligned(16))) char sx[2][STAGE];
  __shared__ __attribute__((aligned(16))) float sn[2][SR];  // -||x||^2/2, -inf past the end
  __shared__ int se[2][SR];                                 // E8M0 row scales (FP8)
  // Survivor queue of wave w: qword/qquery[w][t], t < its fill level.
  __shared__ uint64_t qword[NW][kSeg];
  __shared__ int qquery[NW][kSeg];

  const int per_xcd = (int)(gridDim.x >> 3);
  const int logical = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
  if (logical >= a.num_qtiles * a.num_slices) return;
  const int qtile = logical % a.num_qtiles;
  const int slice = logical / a.num_qtiles;
  const int nq = a.nq_dev ? min(*a.nq_dev, a.nq) : a.nq;
  if (qtile * BQ >= nq) return;
  int64_t r_begin, r_end;
  if constexpr (RESID) {
    r_begin = 32 * (int64_t)a.slice_t0[slice];
    r_end = 32 * (int64_t)a.slice_t0[slice + 1];
  } else {
    r_begin = a.row_begin + (int64_t)slice * a.rows_per_slice;
    r_end = r_begin + a.rows_per_slice;
    if (r_end > a.m) r_end = a.m;
  }
  if (r_begin >= r_end) return;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int h = lane >> 5, col = lane & 31;

  Frag<LB> qf[TN][NC];
  float thr[TN];
  int qsc[TN];
  int qrow[TN];
  float cq[TN];  // RESID: q.c - ||c||^2/2
  if constexpr (RESID) {
    // The centroid and every wave's encoded rows: in the staging buffers, or
    // (wide rows, one-tile stages) in a buffer of their own.
    constexpr int NEED = DPAD * 4 + NW * 32 * G::ROWB;
    constexpr bool OWN = NEED > 2 * STAGE;
    __shared__ __attribute__((aligned(16))) char rarea[OWN ? NEED : 16];
    char* area = OWN ? rarea : &sx[0][0];
    resid_query_operands<E, DPAD, TN, NW>(a, qtile * BQ, a.slice_cl[slice], nq, wave, lane, area, qf, thr,
                                          qsc, qrow, cq);
  } else {
#pragma unroll
    for (int nt = 0; nt < TN; ++nt) {
      const int qr = qtile * BQ + (wave * TN + nt) * 32 + col;
      qrow[nt] = qr;
      cq[nt] = 0.0f;
      if (qr < nq) {
        const char* qp = reinterpret_cast<const char*>(a.q) + (int64_t)qr * G::ROWB;
#pragma unroll
        for (int c = 0; c < NC; ++c) qf[nt][c] = load_frag<LB>(qp + 2 * LB * c + LB * h);
        thr[nt] = fmaxf(-0.5f * a.tau[qr], -FLT_MAX);
        qsc[nt] = E == Elem::FP8 ? e8m0_of(a.qscale[qr]) : 127;
      } else {
#pragma unroll
        for (int c = 0; c < NC; ++c) qf[nt][c] = zero_frag<LB>();
        thr[nt] = INFINITY;
        qsc[nt] = 127;
      }
    }
  }

  const char* xb = reinterpret_cast<const char*>(a.x);
  const int64_t stride = a.stride;
  // Stage loads: the data (pre), the row norm and FP8 scale of stage row tid
  // (pn, pe: raw, turned into the C operand and the E8M0 scale by sstore)
  // and, in lane t < SM of every wave, the rank of stage tile t (grk; sstore
  // turns the ranks into a wave-uniform skip mask). The norm, scale and rank
  // loads are unconditional and their values untouched until sstore: an
  // operation on a loaded value, or a branch merging it, in gload would wait
  // for every load of the stage right after their issue.
  i32x4 pre[UPT];
  float pn = 0.0f;
  float pe = 1.0f;
  int64_t pj0 = 0;  // first row of the loaded stage
  int32_t grk = INT32_MAX;
  unsigned gskip = 0u;  // bit t: tile t of the stored stage is skipped
  static_assert(SM <= 32, "one skip bit per stage tile");
  const int32_t* rank_src = a.tile_rank ? a.tile_rank : a.counts;  // read only with tile_rank
  const int64_t rank_last = a.tile_rank ? a.rank_tiles - 1 : 0;

  // Row and tile indices of the per-stage loads: 32-bit (an index holds fewer
  // than 2^31 rows), except over residual copies, whose padded rows may not.
  using RowT = std::conditional_t<RESID, int64_t, int32_t>;
  const RowT r_end_t = (RowT)r_end;
  const RowT stride_t = (RowT)stride;
  const RowT rank_last_t = (RowT)rank_last;
  const int tiles_total = (int)((a.m + 31) >> 5);  // the copy's tiles (the last may be partial)
  auto gload = [&](int64_t j0) {
    pj0 = j0;
    {
      // Tiles past rank_tiles hold no rows (clamped).
      RowT u = (RowT)(j0 >> 5) + (RowT)(lane < SM ? lane : 0);
      u = u < rank_last_t ? u : rank_last_t;
      grk = rank_src[u];
    }
    {
      RowT j = (RowT)j0 + (RowT)(tid < SR ? tid : 0);
      j = (j < r_end_t ? j : r_end_t - 1) * stride_t;
      pn = a.xnorm[j];
      if (E == Elem::FP8) pe = a.xscale[j];
    }
    if (DPAD <= 128 && stride == 1) {
      // Stride 1: j0 is a multiple of 32, so the stage is SM consecutive whole
      // tiles, one contiguous run of STAGE bytes, read through a buffer
      // resource over the copy's bytes from the stage start: no per-load
      // 64-bit address arithmetic, and pieces past the copy's last tile read
      // as zeros (rows past r_end take the C operand -inf, so whatever they
      // hold never passes a threshold; dense levels skip them). Wider rows
      // keep the plain loads below: at dpad 256 the resource's registers
      // cost a spill in the stage loop (2M x 256: +3 %).
      const int t0 = (int)(j0 >> 5);
      const int tl = tiles_total - t0;  // > 0
      const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
          (void*)(xb + (int64_t)t0 * G::TILE), (short)0, tl >= SM ? STAGE : tl * G::TILE, kBufferWord3);
#pragma unroll
      for (int t = 0; t < UPT; ++t) {
        const int u = tid + NT * t;
        i32x4 v = {0, 0, 0, 0};
        if (UNITS % NT == 0 || u < UNITS)
          v = __builtin_bit_cast(i32x4, __builtin_amdgcn_raw_buffer_load_b128(rs, u * 16, 0, 0));
        pre[t] = v;
      }
      return;
    }
    if (DPAD > 128 && stride == 1 && j0 + SR <= r_end) {
      // Whole stage at stride 1 (wide rows): one contiguous run of STAGE bytes.
      const char* src = xb + (j0 >> 5) * G::TILE;
#pragma unroll
      for (int t = 0; t < UPT; ++t) {
        const int u = tid + NT * t;
        i32x4 v = {0, 0, 0, 0};
        if (UNITS % NT == 0 || u < UNITS) v = *reinterpret_cast<const i32x4*>(src + u * 16);
        pre[t] = v;
      }
      return;
    }
    // General rows (stride > 1, or the last stage of wide rows). The lane's
    // unit offsets are derived from an opaque copy of tid, so that the
    // compiler recomputes them here instead of keeping UPT 64-bit pointers
    // live across the whole stage loop (which spilled one to scratch).
    int tid_o = tid;
    asm volatile("" : "+v"(tid_o));
#pragma unroll
    for (int t = 0; t < UPT; ++t) {
      const int u = tid_o + NT * t;
      i32x4 v = {0, 0, 0, 0};
      if (UNITS % NT == 0 || u < UNITS) {
        const int mt = u / TILE_UNITS;
        const int w = u - mt * TILE_UNITS;
        const int c = w / CHUNK_UNITS;
        const int l = (w - c * CHUNK_UNITS) / LANE_UNITS;
        const int sub = w % LANE_UNITS;
        const int64_t j = j0 + mt * 32 + (l & 31);
        if (j < r_end) {
          const int64_t p = j * stride;
          v = *reinterpret_cast<const i32x4*>(xb + (p >> 5) * G::TILE + c * G::CHUNK +
                                              ((p & 31) + 32 * (l >> 5)) * LB + sub * 16);
        }
      }
      pre[t] = v;
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int t = 0; t < UPT; ++t) {
      const int u = tid + NT * t;
      if (UNITS % NT == 0 || u < UNITS) *reinterpret_cast<i32x4*>(&sx[buf][u * 16]) = pre[t];
    }
    if (tid < SR) {
      // -||x||^2/2 (residual copies hold it already); -inf past the end
      const bool in = (RowT)pj0 + (RowT)tid < r_end_t;
      sn[buf][tid] = in ? (RESID ? pn : -0.5f * pn) : -INFINITY;
      se[buf][tid] = E == Elem::FP8 && in ? e8m0_of(pe) : 127;
    }
    gskip = (unsigned)__ballot(lane < SM && a.tile_rank != nullptr && grk < a.skip_tiles);
  };

  // RESID: queue words carry the residual copy's row; the index row it holds
  // (rid) is looked up when the word leaves for the global list, so the emit
  // path inside the tile loop issues no global load (a load there would make
  // the next wait for memory cover the next stage's prefetch too).
  auto to_index_row = [&](uint64_t w) -> uint64_t {
    if constexpr (RESID) return (w & 0xffffffff00000000ull) | (uint32_t)a.rid[cand_idx(w)];
    return w;
  };
  // Fill level of this wave's queue (wave-uniform: a survivor's position is
  // a ballot prefix count, no atomics).
  unsigned fill = 0u;
  // This wave: moves its queue to the per-query lists.
  auto wave_flush = [&]() {
    for (unsigned t = lane; t < fill; t += 64) {
      const int qr = qquery[wave][t];
      const int pos = atomicAdd(a.counts + qr, 1);
      if (pos < a.cap) a.cand[(int64_t)qr * a.cap + pos] = to_index_row(qword[wave][t]);
    }
    fill = 0u;
  };

  // The stage loop.
  auto run = [&]() {
    int64_t j0 = r_begin;
    gload(j0);
    sstore(0);
    __syncthreads();
    int buf = 0;
    for (; j0 < r_end; j0 += SR) {
      const bool has_next = j0 + SR < r_end;
      const unsigned skip = DENSE ? 0u : gskip;  // this stage's, set when it was stored
      if (has_next) gload(j0 + SR);
      // One 32-row tile per iteration; not unrolled (code size).
#pragma unroll 1
      for (int mt = 0; mt < SM; ++mt) {
        if (skip & (1u << mt)) continue;  // scored by the carried sample level
        f32x16 ci;
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          const f32x4 v = *reinterpret_cast<const f32x4*>(&sn[buf][mt * 32 + 8 * g + 4 * h]);
          ci[4 * g + 0] = v[0];
          ci[4 * g + 1] = v[1];
          ci[4 * g + 2] = v[2];
          ci[4 * g + 3] = v[3];
        }
        const int xsc = E == Elem::FP8 ? se[buf][mt * 32 + col] : 127;
        const char* ap = &sx[buf][mt * G::TILE + lane * LB];
        f32x16 acc[TN];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          const Frag<LB> af = load_frag<LB>(ap + c * G::CHUNK);
#pragma unroll
          for (int nt = 0; nt < TN; ++nt) {
            f32x16 c0 = ci;
            if constexpr (RESID) {
#pragma unroll
              for (int i = 0; i < 16; ++i) c0[i] += cq[nt];
            }
            acc[nt] = mfma_chunk<E>(af, qf[nt][c], c == 0 ? c0 : acc[nt], xsc, qsc[nt]);
          }
        }
        if (DENSE) {
#pragma unroll
          for (int nt = 0; nt < TN; ++nt) {
#pragma unroll
            for (int i = 0; i < 16; ++i) {
              const int64_t row = j0 + mt * 32 + 8 * (i >> 2) + 4 * h + (i & 3);
              if (row < r_end && qrow[nt] < nq)
                a.cand[(int64_t)qrow[nt] * a.cap + row] =
                    pack_cand(-2.0f * acc[nt][i], RESID ? a.rid[row] : (int32_t)(row * stride));
            }
          }
          continue;
        }
        // Largest value per lane and query tile (NaN-propagating maximum:
        // v_maximum3, no canonicalizing moves; a NaN sends the lane to the
        // value-by-value test below, where it never passes).
        float mx[TN];
        bool hit = false;
#pragma unroll
        for (int nt = 0; nt < TN; ++nt) {
          float v = acc[nt][0];
#pragma unroll
          for (int i = 1; i < 16; ++i) v = __builtin_elementwise_maximum(v, acc[nt][i]);
          mx[nt] = v;
          hit |= !(v < thr[nt]);
        }
        if (__any(hit)) {
          // Emit every value >= its query's threshold, one accumulator index
          // at a time: value i of lane L is row 8 (i >> 2) + 4 (L >> 5) +
          // (i & 3) of the tile for the query in qrow of lane L (the 32x32
          // MFMA accumulator layout); a ballot gives the lanes and their
          // queue positions. Four indices (at most 4 x 64 = kSeg values) at
          // a time, after making room for them.
          const int32_t id0 = (int32_t)((j0 + mt * 32 + 4 * h) * stride);
#pragma unroll
          for (int nt = 0; nt < TN; ++nt) {
            if (!__any(!(mx[nt] < thr[nt]))) continue;
#pragma unroll
            for (int g = 0; g < 4; ++g) {
              // Most hitting query tiles hold one or two values: skip a group
              // of four on its maximum (three instructions) first.
              const float gm = __builtin_elementwise_maximum(
                  __builtin_elementwise_maximum(acc[nt][4 * g], acc[nt][4 * g + 1]),
                  __builtin_elementwise_maximum(acc[nt][4 * g + 2], acc[nt][4 * g + 3]));
              if (!__any(!(gm < thr[nt]))) continue;
              uint64_t B[4];
              unsigned cnt = 0u;
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                B[u] = __ballot(acc[nt][4 * g + u] >= thr[nt]);
                cnt += (unsigned)__popcll(B[u]);
              }
              if (cnt == 0u) continue;
              if (fill + cnt > (unsigned)kSeg) wave_flush();
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                if (B[u] == 0) continue;
                if (B[u] >> lane & 1ull) {
                  const unsigned t = fill + __builtin_amdgcn_mbcnt_hi((uint32_t)(B[u] >> 32),
                                                                      __builtin_amdgcn_mbcnt_lo((uint32_t)B[u], 0u));
                  const float v = acc[nt][4 * g + u];
                  // row 8 g + u of the lane's half: (int32) row * stride, in
                  // wrapping 32-bit arithmetic
                  qword[wave][t] = pack_cand(-2.0f * v, id0 + (int32_t)((8 * g + u) * stride));
                  qquery[wave][t] = qrow[nt];
                }
                fill += (unsigned)__popcll(B[u]);
              }
            }
          }
        }
      }
      if (has_next) sstore(buf ^ 1);
      __syncthreads();
      buf ^= 1;
    }
  };
  run();
  if constexpr (!DENSE) wave_flush();
}

// Ring variants (wide / narrow query tiles): the same tiles, waves, queries
// and arithmetic as the LDS variants, different data movement. Stages are
// copied global -> LDS by the LDS-DMA path (global_load_lds_dwordx4: no
// staging registers, no VALU) into a ring of kRing slots, so the copies of
// stages s+1 and s+2 are in flight while stage s is computed. One barrier per
// stage; before it each wave waits (vmcnt) only for its own copies of stage s.
// The C operand -||x||^2/2 is copied in beside the tile (-inf for rows past
// the end), FP8 row scales as floats.
constexpr int kRing = 3;

template <Elem E, int DPAD, bool WIDE>
struct RingShape {
  using G = Geometry<E, DPAD>;
  static constexpr int kFit = (WIDE ? 56 * 1024 : 36 * 1024) / (kRing * G::TILE);
  static constexpr int SM = kFit > 4 ? 4 : (kFit < 1 ? 1 : kFit);  // tiles per stage
  static constexpr int SR = 32 * SM;
  static constexpr int STAGE = SM * G::TILE;
  static constexpr int DATA_OPS = STAGE / 1024;  // one op: 64 lanes x 16 B
  static constexpr int ROW_OPS = (SR + 63) / 64;  // ops per per-row array
  static constexpr int OPS = DATA_OPS + ROW_OPS * (E == Elem::FP8 ? 2 : 1);
  static constexpr int LDS = kRing * STAGE + kRing * SR * 8 + 2 * kQueue * 12 + 4 * kHitSlots * 64;
  static constexpr int WAVES_PER_SIMD = (!WIDE && 3 * LDS <= 160 * 1024) ? 3 : 2;
  static_assert(STAGE % 1024 == 0, "stage is whole 1 KB copies");
};

// s_waitcnt vmcnt(N) lgkmcnt(0) (gfx9 encoding; expcnt not waited on).
template <int N>
__device__ inline void wait_vm_lgkm() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4));
}

// Workgroup barrier without the release fence of __syncthreads (which would
// wait for every copy in flight); the caller waits for what must be complete.
__device__ inline void stage_barrier() { asm volatile("s_barrier" ::: "memory"); }

// Never launched for a dense pass (the LDS kernel's dense instantiation runs it).
template <Elem E, int DPAD, int TN, bool WIDE>
__global__ __launch_bounds__(kBlock, (RingShape<E, DPAD, WIDE>::WAVES_PER_SIMD)) void filter_ring_kernel(
    FilterArgs a) {
  using G = Geometry<E, DPAD>;
  using RS = RingShape<E, DPAD, WIDE>;
  constexpr int LB = G::LB;
  constexpr int NC = G::NC;
  constexpr int BQ = 4 * TN * 32;
  constexpr int SM = RS::SM;
  constexpr int SR = RS::SR;
  constexpr int TILE_UNITS = G::TILE / 16;
  constexpr int CHUNK_UNITS = G::CHUNK / 16;
  constexpr int LANE_UNITS = LB / 16;
  typedef __attribute__((address_space(3))) void lds_void;

  __shared__ __attribute__((aligned(16))) char sx[kRing][RS::STAGE];
  __shared__ __attribute__((aligned(16))) float sn[kRing][SR];
  __shared__ __attribute__((aligned(16))) float se[kRing][E == Elem::FP8 ? SR : 1];
  __shared__ uint64_t qword[2][kQueue];
  __shared__ int qquery[2][kQueue];
  __shared__ unsigned qtail[2];
  __shared__ __attribute__((aligned(16))) float hitbuf[kBlock / 64][kHitSlots * 16];

  const int per_xcd = (int)(gridDim.x >> 3);
  const int logical = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
  if (logical >= a.num_qtiles * a.num_slices) return;
  const int qtile = logical % a.num_qtiles;
  const int slice = logical / a.num_qtiles;
  const int nq = a.nq_dev ? min(*a.nq_dev, a.nq) : a.nq;
  if (qtile * BQ >= nq) return;
  const int64_t r_begin = a.row_begin + (int64_t)slice * a.rows_per_slice;
  int64_t r_end = r_begin + a.rows_per_slice;
  if (r_end > a.m) r_end = a.m;
  if (r_begin >= r_end) return;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int h = lane >> 5, col = lane & 31;

  Frag<LB> qf[TN][NC];
  float thr[TN];
  int qsc[TN];
  int qrow[TN];
#pragma unroll
  for (int nt = 0; nt < TN; ++nt) {
    const int qr = qtile * BQ + (wave * TN + nt) * 32 + col;
    qrow[nt] = qr;
    if (qr < nq) {
      const char* qp = reinterpret_cast<const char*>(a.q) + (int64_t)qr * G::ROWB;
#pragma unroll
      for (int c = 0; c < NC; ++c) qf[nt][c] = load_frag<LB>(qp + 2 * LB * c + LB * h);
      thr[nt] = fmaxf(-0.5f * a.tau[qr], -FLT_MAX);
      qsc[nt] = E == Elem::FP8 ? e8m0_of(a.qscale[qr]) : 127;
    } else {
#pragma unroll
      for (int c = 0; c < NC; ++c) qf[nt][c] = zero_frag<LB>();
      thr[nt] = INFINITY;
      qsc[nt] = 127;
    }
  }

  const char* xb = reinterpret_cast<const char*>(a.x);
  const int64_t stride = a.stride;
  const int n_st = (int)((r_end - r_begin + SR - 1) / SR);

  // Copies stage [j0, j0 + SR) into ring slot `slot`. Wave w issues ops w,
  // w + 4, ...: every wave issues at least OPS / 4 per stage.
  auto issue = [&](int slot, int64_t j0) {
    const bool whole = stride == 1 && j0 + SR <= r_end;
    for (int o = wave; o < RS::OPS; o += 4) {
      if (o < RS::DATA_OPS) {
        const int u = o * 64 + lane;
        const char* src;
        if (whole) {
          src = xb + (j0 >> 5) * G::TILE + (int64_t)u * 16;
        } else {
          const int mt = u / TILE_UNITS;
          const int w = u - mt * TILE_UNITS;
          const int c = w / CHUNK_UNITS;
          const int l = (w - c * CHUNK_UNITS) / LANE_UNITS;
          const int sub = w % LANE_UNITS;
          int64_t j = j0 + mt * 32 + (l & 31);
          if (j >= r_end) j = r_end - 1;  // any valid row: its C operand is -inf
          const int64_t p = j * stride;
          src = xb + (p >> 5) * G::TILE + c * G::CHUNK + ((p & 31) + 32 * (l >> 5)) * LB + sub * 16;
        }
        __builtin_amdgcn_global_load_lds(src, (lds_void*)&sx[slot][o * 1024], 16, 0, 0);
      } else {
        const int r = o - RS::DATA_OPS;
        const int part = r % RS::ROW_OPS;
        const int row = part * 64 + lane;
        if (row < SR) {
          const int64_t j = j0 + row;
          if (r < RS::ROW_OPS) {
            const float* src = j < r_end ? a.xnh + j * stride : a.xnh + a.xnh_inf;
            __builtin_amdgcn_global_load_lds(src, (lds_void*)&sn[slot][part * 64], 4, 0, 0);
          } else {
            const float* src = a.xscale + (j < r_end ? j : r_end - 1) * stride;
            __builtin_amdgcn_global_load_lds(src, (lds_void*)&se[slot][part * 64], 4, 0, 0);
          }
        }
      }
    }
  };

  unsigned head[2] = {0u, 0u};
  auto put = [&](int par, unsigned t, int qr, float v, int64_t row) {
    const uint64_t w = pack_cand(-2.0f * v, (int32_t)(row * stride));
    if (t - head[par] < (unsigned)kQueue) {
      qword[par][t & (kQueue - 1)] = w;
      qquery[par][t & (kQueue - 1)] = qr;
    } else {
      const int pos = atomicAdd(a.counts + qr, 1);
      if (pos < a.cap) a.cand[(int64_t)qr * a.cap + pos] = w;
    }
  };
  auto flush = [&](int par, unsigned tail) {
    const unsigned n = min(tail - head[par], (unsigned)kQueue);
    for (unsigned i = tid; i < n; i += kBlock) {
      const unsigned t = (head[par] + i) & (kQueue - 1);
      const int qr = qquery[par][t];
      const int pos = atomicAdd(a.counts + qr, 1);
      if (pos < a.cap) a.cand[(int64_t)qr * a.cap + pos] = qword[par][t];
    }
    head[par] = tail;
  };

  if (tid < 2) qtail[tid] = 0u;
  auto run = [&]() {
#pragma unroll
    for (int i = 0; i < kRing - 1; ++i)
      if (i < n_st) issue(i, r_begin + (int64_t)i * SR);
    int slot = 0;
    for (int s = 0; s < n_st; ++s) {
      const int64_t j0 = r_begin + (int64_t)s * SR;
      // This wave's copies of stage s are done when at most those of the
      // later stages already issued are outstanding.
      if (s + 1 < n_st) {
        wait_vm_lgkm<RS::OPS / 4>();
      } else {
        wait_vm_lgkm<0>();
      }
      stage_barrier();
      if (s + kRing - 1 < n_st) issue(slot == 0 ? kRing - 1 : slot - 1, j0 + (int64_t)(kRing - 1) * SR);
      if (s > 0) {
        const int par = (s - 1) & 1;
        const unsigned tail = qtail[par];
        if (tail - head[par] > (unsigned)(kQueue / 2)) flush(par, tail);
      }
      const int par = s & 1;
#pragma unroll 1
      for (int mt = 0; mt < SM; ++mt) {
        f32x16 ci;
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          const f32x4 v = *reinterpret_cast<const f32x4*>(&sn[slot][mt * 32 + 8 * g + 4 * h]);
          ci[4 * g + 0] = v[0];
          ci[4 * g + 1] = v[1];
          ci[4 * g + 2] = v[2];
          ci[4 * g + 3] = v[3];
        }
        const int xsc = E == Elem::FP8 ? e8m0_of(se[slot][E == Elem::FP8 ? mt * 32 + col : 0]) : 127;
        const char* ap = &sx[slot][mt * G::TILE + lane * LB];
        f32x16 acc[TN];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          const Frag<LB> af = load_frag<LB>(ap + c * G::CHUNK);
#pragma unroll
          for (int nt = 0; nt < TN; ++nt)
            acc[nt] = mfma_chunk<E>(af, qf[nt][c], c == 0 ? ci : acc[nt], xsc, qsc[nt]);
        }
        float mx[TN];
        bool hit = false;
#pragma unroll
        for (int nt = 0; nt < TN; ++nt) {
          float v = acc[nt][0];
#pragma unroll
          for (int i = 1; i < 16; ++i) v = fmaxf(v, acc[nt][i]);
          mx[nt] = v;
          hit |= v >= thr[nt];
        }
        if (__any(hit))
          emit_hits<TN>(acc, mx, thr, qrow, j0 + mt * 32, lane, hitbuf[wave], &qtail[par],
                        [&](unsigned t, int qr, float v, int64_t row) { put(par, t, qr, v, row); });
      }
      slot = slot == kRing - 1 ? 0 : slot + 1;
    }
  };
  run();
  __syncthreads();
  flush(0, qtail[0]);
  flush(1, qtail[1]);
}

// Packed encoders: one thread per 16-byte output piece.
template <Elem E, int DPAD>
__global__ void encode_packed_kernel(const float* __restrict__ x, int64_t n, int d, int64_t ldx,
                                     int64_t tile_from, int64_t tile_to, char* __restrict__ out,
                                     const float* __restrict__ scales, int64_t map_step,
                                     int64_t map_n) {
  using G = Geometry<E, DPAD>;
  constexpr int LB = G::LB;
  constexpr int PIECES = G::TILE / 16;  // 16-byte pieces per tile
  const int64_t total = (tile_to - tile_from) * PIECES;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = tile_from + e / PIECES;
    const int w = (int)(e % PIECES);               // piece within the tile
    const int c = w / (G::CHUNK / 16);             // K-chunk
    const int l = (w % (G::CHUNK / 16)) / (LB / 16);  // lane
    const int sub = w % (LB / 16);                 // 16-byte part of the lane piece
    const int64_t row = t * 32 + (l & 31);
    const int k0 = c * PackTraits<E>::kChunkElems + (l >> 5) * (PackTraits<E>::kChunkElems / 2) +
                   sub * (16 / PackTraits<E>::kBytes);
    i32x4 piece = {0, 0, 0, 0};
    if (row < n) {
      // Source row: row itself, or its index row for a sample copy (map_n =
      // the index's whole tiles).
      const int64_t src = map_step > 0 ? sample_source_row(row, map_step, map_n) : row;
      if (E == Elem::BF16) {
        uint16_t v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = k0 + i < d ? bf16_rne(x[src * ldx + k0 + i]) : 0;
        __builtin_memcpy(&piece, v, 16);
      } else {
        const float s = scales[row];
        uint8_t v[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = k0 + i < d ? e4m3_rne_sat(x[src * ldx + k0 + i] / s) : 0;
        __builtin_memcpy(&piece, v, 16);
      }
    }
    *reinterpret_cast<i32x4*>(out + t * G::TILE + (int64_t)w * 16) = piece;
  }
}

// Per-row power-of-two FP8 scale (same rule as the row-major encoder).
__global__ __launch_bounds__(256) void e4m3_scales_kernel(const float* __restrict__ x, int64_t n,
                                                          int d, int64_t ldx,
                                                          float* __restrict__ scales) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * 4 + wave;
  if (row >= n) return;
  float mx = 0.0f;
  for (int j = lane; j < d; j += 64) mx = fmaxf(mx, fabsf(x[row * ldx + j]));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
  if (lane == 0) scales[row] = e4m3_row_scale(mx);
}

// Query tiles per wave of the direct variant: as many as keep the query
// operands within ~160 registers (one wave per SIMD), at most 7.
template <Elem E, int DPAD>
constexpr int direct_tn() {
  using G = Geometry<E, DPAD>;
  constexpr int per_tile = G::NC * G::LB / 4;
  constexpr int tn = 160 / per_tile;
  return tn > 7 ? 7 : (tn < 1 ? 1 : tn);
}

// Query tiles per wave of the LDS variants: "wide" (two waves per SIMD)
// keeps the query operands within ~80 registers (at most 4 tiles), "narrow"
// (three waves per SIMD, ~168 registers a wave) one tile fewer for BF16 and
// two fewer for FP8 (its three-tile narrow kernel spills). More query tiles
// per wave spread each tile's fixed instructions (operand and C-operand LDS
// reads, skip test, loop) over more MFMAs: FP8 at d = 96 (dpad 128), wide,
// 4 tiles instead of 3: full-database launch 12.05 -> 10.90 ms (same box).
template <Elem E, int DPAD, bool WIDE>
constexpr int lds_tn() {
  using G = Geometry<E, DPAD>;
  constexpr int per_tile = G::NC * G::LB / 4;
  constexpr int cap = 4;
  constexpr int wide = 80 / per_tile > cap ? cap : (80 / per_tile < 1 ? 1 : 80 / per_tile);
  constexpr int fewer = E == Elem::FP8 ? 2 : 1;
  constexpr int narrow = wide - fewer < 1 ? 1 : wide - fewer;
  return WIDE ? wide : narrow;
}

// Query tiles per wave of the non-residual LDS wide kernel: BF16 holds up to
// ~96 registers of query operands there and at least 2 tiles (4 at dpad <= 96,
// 3 at 128, 2 from 160 to 256), which spreads each tile's fixed instructions (operand and
// C-operand reads, skip test, loop) over more MFMAs. Measured on one box
// (10k queries, full-database launch): 10M x 96 15.42 -> 15.19 ms, 2M x 192
// 9.70 -> 7.79 ms, 2M x 256 12.34 -> 9.48 ms, 2M x 128 unchanged; 3 tiles at
// dpad 160 spill and were slower. Everything else as lds_tn.
template <Elem E, int DPAD>
constexpr int wide_tn() {
  using G = Geometry<E, DPAD>;
  constexpr int per_tile = G::NC * G::LB / 4;
  constexpr int t96 = 96 / per_tile > 4 ? 4 : (96 / per_tile < 2 ? 2 : 96 / per_tile);
  return E == Elem::BF16 ? t96 : lds_tn<E, DPAD, true>();
}

template <Elem E, int DPAD>
int tn_of(int variant) {
  switch (variant) {
    case kFilterDirect: return direct_tn<E, DPAD>();
    case kFilterLdsNarrow:
    case kFilterRingNarrow: return lds_tn<E, DPAD, false>();
    case kFilterLdsWide: return wide_tn<E, DPAD>();
    default: return lds_tn<E, DPAD, true>();
  }
}

template <Elem E, int DPAD>
void launch_packed_one(int variant, const FilterArgs& a, hipStream_t st) {
  const int total = a.num_qtiles * a.num_slices;
  const unsigned grid = (unsigned)round_up(total, 8);
  if (a.dense) {
    RPLSH_REQUIRE(variant == kFilterLdsNarrow, "dense pass: the narrow LDS kernel's tiling expected");
    filter_lds_kernel<E, DPAD, lds_tn<E, DPAD, false>(), false, 4, false, true><<<grid, kBlock, 0, st>>>(a);
    RPLSH_HIP(hipGetLastError());
    return;
  }
  switch (variant) {
    case kFilterDirect:
      filter_packed_kernel<E, DPAD, direct_tn<E, DPAD>()><<<grid, kBlock, 0, st>>>(a);
      break;
    case kFilterLdsNarrow:
      filter_lds_kernel<E, DPAD, lds_tn<E, DPAD, false>(), false><<<grid, kBlock, 0, st>>>(a);
      break;
    case kFilterLdsWide8:
      filter_lds_kernel<E, DPAD, lds_tn<E, DPAD, true>(), true, 8><<<grid, 512, 0, st>>>(a);
      break;
    case kFilterRingWide:
      RPLSH_REQUIRE(a.xnh, "ring filter: C-operand array missing");
      filter_ring_kernel<E, DPAD, lds_tn<E, DPAD, true>(), true><<<grid, kBlock, 0, st>>>(a);
      break;
    case kFilterRingNarrow:
      RPLSH_REQUIRE(a.xnh, "ring filter: C-operand array missing");
      filter_ring_kernel<E, DPAD, lds_tn<E, DPAD, false>(), false><<<grid, kBlock, 0, st>>>(a);
      break;
    default:
      filter_lds_kernel<E, DPAD, wide_tn<E, DPAD>(), true><<<grid, kBlock, 0, st>>>(a);
      break;
  }
  RPLSH_HIP(hipGetLastError());
}

template <Elem E, int DPAD>
void encode_one(const float* x, int64_t n, int d, int64_t ldx, int64_t row_from, char* out,
                float* scales, hipStream_t st) {
  using G = Geometry<E, DPAD>;
  if (E == Elem::FP8 && n > row_from) {
    e4m3_scales_kernel<<<(unsigned)div_up(n - row_from, 4), 256, 0, st>>>(
        x + row_from * ldx, n - row_from, d, ldx, scales + row_from);
    RPLSH_HIP(hipGetLastError());
  }
  const int64_t t0 = row_from / 32, t1 = div_up(n, 32);
  if (t1 <= t0) return;
  const int64_t work = (t1 - t0) * (G::TILE / 16);
  int64_t g = div_up(work, 256);
  if (g > 65536) g = 65536;
  encode_packed_kernel<E, DPAD><<<(unsigned)g, 256, 0, st>>>(x, n, d, ldx, t0, t1, out, scales, 0, 0);
  RPLSH_HIP(hipGetLastError());
}

// Sample copy: row r < sm holds index row sample_source_row(r, step, tiles).
template <Elem E, int DPAD>
void encode_sample_one(const float* x, int64_t tiles, int d, int64_t ldx, int64_t sm, int64_t step,
                       char* out, const float* scales, hipStream_t st) {
  using G = Geometry<E, DPAD>;
  const int64_t t1 = div_up(sm, 32);
  const int64_t work = t1 * (G::TILE / 16);
  int64_t g = div_up(work, 256);
  if (g > 65536) g = 65536;
  encode_packed_kernel<E, DPAD><<<(unsigned)g, 256, 0, st>>>(x, sm, d, ldx, 0, t1, out, scales, step, tiles);
  RPLSH_HIP(hipGetLastError());
}

// Calls f.template operator()<E, DPAD>() for the instantiated (E, dpad).
template <typename F>
auto dispatch(Elem e, int dpad, F&& f) {
  if (e == Elem::BF16) {
    switch (dpad) {
      case 32: return f.template operator()<Elem::BF16, 32>();
      case 64: return f.template operator()<Elem::BF16, 64>();
      case 96: return f.template operator()<Elem::BF16, 96>();
      case 128: return f.template operator()<Elem::BF16, 128>();
      case 160: return f.template operator()<Elem::BF16, 160>();
      case 192: return f.template operator()<Elem::BF16, 192>();
      case 256: return f.template operator()<Elem::BF16, 256>();
      default: break;
    }
  } else if (e == Elem::FP8) {
    switch (dpad) {
      case 64: return f.template operator()<Elem::FP8, 64>();
      case 128: return f.template operator()<Elem::FP8, 128>();
      case 256: return f.template operator()<Elem::FP8, 256>();
      default: break;
    }
  }
  fail(RPLSH_ERR_UNSUPPORTED, "packed kernels: element kind / dimension %d not instantiated", dpad);
}

__global__ void sample_rows_kernel(const float* __restrict__ xn, const float* __restrict__ xscale,
                                   int64_t tiles, int64_t sm, int64_t step, float* __restrict__ snorm,
                                   float* __restrict__ snh, int64_t nh_inf,
                                   float* __restrict__ sscale) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < sm) {
    const int64_t src = sample_source_row(r, step, tiles);
    snorm[r] = xn[src];
    snh[r] = -0.5f * xn[src];
    if (sscale) sscale[r] = xscale[src];
  }
  if (r == 0) snh[nh_inf] = -INFINITY;
}

// rank[u] = (u * step_inv) mod tiles -- the sample tile holding index tile u --
// when below stiles, else INT32_MAX (also for u >= tiles: the partial tile).
__global__ void tile_ranks_kernel(int64_t tiles, int64_t stiles, int64_t step_inv, int32_t* __restrict__ rank,
                                  int64_t rank_tiles) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= rank_tiles) return;
  int32_t v = INT32_MAX;
  if (u < tiles) {
    const int64_t s = (int64_t)(((uint64_t)u * (uint64_t)step_inv) % (uint64_t)tiles);
    if (s < stiles) v = (int32_t)s;
  }
  rank[u] = v;
}

// One wave per query: keeps, in list order, the words whose row lies below
// row_limit and that pass the next level's own test on acc = -key/2 (the
// filter wrote key = -2 acc, so both halvings are exact), with the sample row
// replaced by its index row when step > 0 (into the full level). Reads of a
// chunk precede its writes, and writes never pass the chunk being read.
__global__ __launch_bounds__(256) void carry_kernel(CarryArgs a) {
  const int q = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (q >= a.nq || (a.nq_dev && q >= *a.nq_dev)) return;
  const int count = a.counts[q];
  const int n = count < a.cap ? count : a.cap;
  const float thr = fmaxf(-0.5f * a.tau[q], -FLT_MAX);
  const int64_t rlim = a.row_limit;
  uint64_t* list = a.cand + (int64_t)q * a.cap;
  int kept = 0;
  for (int i0 = 0; i0 < n; i0 += 64) {
    const int i = i0 + lane;
    uint64_t w = i < n ? list[i] : kEmptyCand;
    bool keep = false;
    if (w != kEmptyCand) {
      const int32_t r = cand_idx(w);
      keep = -0.5f * cand_key(w) >= thr && r < rlim;
      if (keep && a.step > 0) w = pack_cand(cand_key(w), (int32_t)sample_source_row(r, a.step, a.tiles));
    }
    const uint64_t m = __ballot(keep);
    const int pos = kept + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    if (keep) list[pos] = w;
    kept += __popcll(m);
  }
  if (lane == 0) {
    a.counts[q] = kept;
    if (count > a.cap && a.status) atomicOr(a.status + q, (int)kStatusOverflow);
  }
}

__global__ void neg_half_norms_kernel(const float* __restrict__ xn, int64_t from, int64_t n,
                                      float* __restrict__ nh, int64_t inf_at) {
  const int64_t i = from + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) nh[i] = -0.5f * xn[i];
  if (i == from) nh[inf_at] = -INFINITY;
}


}  // namespace

void launch_sample_rows(const float* xn, const float* xscale, int64_t tiles, int64_t sm, int64_t step,
                        float* snorm, float* snh, int64_t nh_inf, float* sscale, hipStream_t st) {
  RPLSH_REQUIRE(sm > 0 && sm % 32 == 0 && tiles > 0 && sm <= 32 * tiles && step > 0 && sm <= nh_inf,
                "sample rows: bad arguments");
  sample_rows_kernel<<<(unsigned)div_up(sm, (int64_t)256), 256, 0, st>>>(xn, xscale, tiles, sm, step,
                                                                        snorm, snh, nh_inf, sscale);
  RPLSH_HIP(hipGetLastError());
}

void launch_tile_ranks(int64_t tiles, int64_t stiles, int64_t step, int32_t* rank, int64_t rank_tiles,
                       hipStream_t st) {
  RPLSH_REQUIRE(tiles > 0 && stiles >= 0 && stiles <= tiles && rank_tiles >= tiles && tiles < INT32_MAX,
                "tile ranks: bad arguments");
  // step^-1 mod tiles (extended Euclid; step is coprime with tiles)
  int64_t r0 = tiles, r1 = step % tiles, t0 = 0, t1 = 1;
  while (r1 != 0) {
    const int64_t qq = r0 / r1;
    int64_t tmp = r0 - qq * r1;
    r0 = r1;
    r1 = tmp;
    tmp = t0 - qq * t1;
    t0 = t1;
    t1 = tmp;
  }
  RPLSH_REQUIRE(r0 == 1 || tiles == 1, "tile ranks: step not coprime with the tile count");
  const int64_t inv = ((t0 % tiles) + tiles) % tiles;
  tile_ranks_kernel<<<(unsigned)div_up(rank_tiles, (int64_t)256), 256, 0, st>>>(tiles, stiles, inv, rank,
                                                                                rank_tiles);
  RPLSH_HIP(hipGetLastError());
}

void launch_carry(const CarryArgs& a, hipStream_t st) {
  if (a.nq <= 0) return;
  RPLSH_REQUIRE(a.cand && a.counts && a.tau && a.row_limit > 0 && (a.step == 0 || a.tiles > 0),
                "carry: bad arguments");
  carry_kernel<<<(unsigned)div_up(a.nq, 4), 256, 0, st>>>(a);
  RPLSH_HIP(hipGetLastError());
}

void launch_neg_half_norms(const float* xn, int64_t from, int64_t n, float* nh, int64_t inf_at,
                           hipStream_t st) {
  RPLSH_REQUIRE(from >= 0 && n <= inf_at && nh, "half norms: bad range");
  const int64_t cnt = std::max<int64_t>(n - from, 1);
  neg_half_norms_kernel<<<(unsigned)div_up(cnt, (int64_t)256), 256, 0, st>>>(xn, from, n, nh, inf_at);
  RPLSH_HIP(hipGetLastError());
}


int64_t packed_tile_bytes(Elem e, int dpad) {
  return e == Elem::BF16 ? (int64_t)32 * dpad * 2 : (int64_t)32 * dpad;
}

namespace {

struct TnOf {
  int variant;
  template <Elem E, int DPAD>
  int operator()() const {
    return tn_of<E, DPAD>(variant);
  }
};

struct LaunchFilter {
  int variant;
  const FilterArgs& a;
  hipStream_t st;
  template <Elem E, int DPAD>
  int operator()() const {
    launch_packed_one<E, DPAD>(variant, a, st);
    return 0;
  }
};

struct ResidTn {
  bool narrow;
  template <Elem E, int DPAD>
  int operator()() const {
    return narrow ? lds_tn<E, DPAD, false>() : lds_tn<E, DPAD, true>();
  }
};

struct LaunchResid {
  bool narrow;
  const FilterArgs& a;
  hipStream_t st;
  template <Elem E, int DPAD>
  int operator()() const {
    const int64_t total = (int64_t)a.num_qtiles * a.num_slices;
    RPLSH_REQUIRE(total < (1ll << 31), "residual filter grid too large");
    const unsigned grid = (unsigned)round_up(total, 8);
    if (narrow)
      filter_lds_kernel<E, DPAD, lds_tn<E, DPAD, false>(), false, 4, true><<<grid, kBlock, 0, st>>>(a);
    else
      filter_lds_kernel<E, DPAD, lds_tn<E, DPAD, true>(), true, 4, true><<<grid, kBlock, 0, st>>>(a);
    RPLSH_HIP(hipGetLastError());
    return 0;
  }
};

struct EncodeSample {
  const float* x;
  int64_t tiles;
  int d;
  int64_t ldx, sm, step;
  char* out;
  const float* scales;
  hipStream_t st;
  template <Elem E, int DPAD>
  int operator()() const {
    encode_sample_one<E, DPAD>(x, tiles, d, ldx, sm, step, out, scales, st);
    return 0;
  }
};

struct EncodePacked {
  const float* x;
  int64_t n;
  int d;
  int64_t ldx, row_from;
  char* out;
  float* scales;
  hipStream_t st;
  template <Elem E, int DPAD>
  int operator()() const {
    encode_one<E, DPAD>(x, n, d, ldx, row_from, out, scales, st);
    return 0;
  }
};

}  // namespace

int filter_packed_block_queries(Elem e, int dpad, int variant) {
  return (variant == kFilterLdsWide8 ? 8 : 4) * 32 * dispatch(e, dpad, TnOf{variant});
}

void launch_filter_packed(Elem e, int dpad, int variant, const FilterArgs& a, hipStream_t st) {
  dispatch(e, dpad, LaunchFilter{variant, a, st});
}

int filter_resid_block_queries(Elem e, int dpad, int variant) {
  const bool narrow = variant == kFilterLdsNarrow || variant == kFilterRingNarrow;
  return 4 * 32 * dispatch(e, dpad, ResidTn{narrow});
}

void launch_filter_resid(Elem e, int dpad, int variant, const FilterArgs& a, hipStream_t st) {
  if (a.nq <= 0 || a.num_slices <= 0) return;
  const bool narrow = variant == kFilterLdsNarrow || variant == kFilterRingNarrow;
  dispatch(e, dpad, LaunchResid{narrow, a, st});
}

void launch_encode_packed(Elem e, int dpad, const float* x, int64_t n, int d, int64_t ldx,
                          int64_t row_from, void* out, float* scales, hipStream_t st) {
  dispatch(e, dpad, EncodePacked{x, n, d, ldx, row_from, static_cast<char*>(out), scales, st});
}

void launch_encode_sample(Elem e, int dpad, const float* x, int64_t tiles, int d, int64_t ldx, int64_t sm,
                          int64_t step, void* out, const float* scales, hipStream_t st) {
  RPLSH_REQUIRE(sm % 32 == 0 && sm <= 32 * tiles, "sample copy: bad size");
  dispatch(e, dpad, EncodeSample{x, tiles, d, ldx, sm, step, static_cast<char*>(out), scales, st});
}

}  // namespace rplsh
