// Fused candidate contraction + threshold filter (the hot kernel).
//
// Replaces, in one pass over the database, the reference's per-tile
// cublasSgemm (src/internal/matrix_ops.cu:56-58 via search.cu:176-178), the
// fused base-norm add + k-select (select.cu:189-224) and its global distance
// tile: the Bq x Bn distance matrix never exists in HBM.
//
// Work decomposition
//   workgroup  = 4 waves; 4*TN*32 queries (TN query tiles of 32 per wave) x one
//                slice of database rows. Query fragments stay in registers for
//                the whole slice; the slice streams through LDS in chunks of 64
//                rows (double buffered, one barrier per chunk).
//   MFMA       = 32x32 output tiles, database rows on M (A operand), queries on
//                N (B operand): lane l owns query column l%32 and 16 rows.
//   accumulate = the first MFMA of every tile takes C = -||x||^2/2 of its rows,
//                so the accumulator ends as  q.x - ||x||^2/2  and the
//                approximate relative distance is  key = -2 * acc.
//   filter     = key <= tau_q  <=>  acc >= -tau_q/2: one max-reduction of the
//                lane's 16 values per tile and a compare; survivors (rare) are
//                appended to the query's candidate list with a global atomic.
//   grid       = (query tiles x slices), renumbered so consecutive workgroups
//                on one XCD (hardware dispatch is round-robin over the 8 XCDs)
//                take all query tiles of the same slice: the slice is read
//                from HBM once per XCD and re-served from that XCD's L2.
//
// Element kinds
//   F32  : v_mfma_f32_32x32x2_f32, 4 per 32-byte chunk (lane half h supplies
//          elements 8c+4h..8c+4h+3 of chunk c; both operands use the same
//          assignment, so the contraction is unchanged).
//   BF16 : v_mfma_f32_32x32x16_bf16, one per 32-byte chunk.
//   FP8  : v_mfma_scale_f32_32x32x64_f8f6f4 (e4m3), one per 64-byte chunk; the
//          per-row power-of-two scales go in as E8M0 scale operands, so the
//          accumulator is already in real units.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstdint>

#include "common.hpp"
#include "device.hpp"
#include "kernels.hpp"

namespace rplsh {

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kBlock = 256;
constexpr int kRowsPerChunk = 64;

template <Elem E>
struct ElemTraits;
template <>
struct ElemTraits<Elem::F32> {
  static constexpr int kBytes = 4;
  static constexpr int kLaneBytes = 16;
};
template <>
struct ElemTraits<Elem::BF16> {
  static constexpr int kBytes = 2;
  static constexpr int kLaneBytes = 16;
};
template <>
struct ElemTraits<Elem::FP8> {
  static constexpr int kBytes = 1;
  static constexpr int kLaneBytes = 32;
};

template <int LB>
struct Frag;
template <>
struct Frag<16> {
  i32x4 v;
};
template <>
struct Frag<32> {
  i32x4 v0, v1;
};

template <int LB>
__device__ inline Frag<LB> load_frag(const char* p);
template <>
__device__ inline Frag<16> load_frag<16>(const char* p) {
  Frag<16> f;
  f.v = *reinterpret_cast<const i32x4*>(p);
  return f;
}
template <>
__device__ inline Frag<32> load_frag<32>(const char* p) {
  Frag<32> f;
  f.v0 = *reinterpret_cast<const i32x4*>(p);
  f.v1 = *reinterpret_cast<const i32x4*>(p + 16);
  return f;
}

template <int LB>
__device__ inline Frag<LB> zero_frag() {
  Frag<LB> f;
  __builtin_memset(&f, 0, sizeof(f));
  return f;
}

// acc += A(rows) * B(queries) over one K-chunk.
template <Elem E>
__device__ inline f32x16 mfma_chunk(const Frag<ElemTraits<E>::kLaneBytes>& a,
                                    const Frag<ElemTraits<E>::kLaneBytes>& b, f32x16 c, int sa,
                                    int sb);

template <>
__device__ inline f32x16 mfma_chunk<Elem::F32>(const Frag<16>& a, const Frag<16>& b, f32x16 c,
                                               int, int) {
  const f32x4 av = __builtin_bit_cast(f32x4, a.v);
  const f32x4 bv = __builtin_bit_cast(f32x4, b.v);
  c = __builtin_amdgcn_mfma_f32_32x32x2f32(av[0], bv[0], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_32x32x2f32(av[1], bv[1], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_32x32x2f32(av[2], bv[2], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_32x32x2f32(av[3], bv[3], c, 0, 0, 0);
  return c;
}

template <>
__device__ inline f32x16 mfma_chunk<Elem::BF16>(const Frag<16>& a, const Frag<16>& b, f32x16 c,
                                                int, int) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a.v),
                                                  __builtin_bit_cast(bf16x8, b.v), c, 0, 0, 0);
}

template <>
__device__ inline f32x16 mfma_chunk<Elem::FP8>(const Frag<32>& a, const Frag<32>& b, f32x16 c,
                                               int sa, int sb) {
  i32x8 av, bv;
  av.lo = a.v0;
  av.hi = a.v1;
  bv.lo = b.v0;
  bv.hi = b.v1;
  // cbsz = blgp = 0: both operands e4m3; byte 0 of sa/sb is the E8M0 scale.
  return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, bv, c, 0, 0, 0, sa, 0, sb);
}

// E8M0 exponent byte of a power-of-two scale.
__device__ inline int e8m0_of(float s) { return (int)((__float_as_uint(s) >> 23) & 0xff); }

template <Elem E, int DPAD, int TN>
__global__ __launch_bounds__(kBlock, 2) void filter_kernel(FilterArgs a) {
  using T = ElemTraits<E>;
  constexpr int LB = T::kLaneBytes;
  constexpr int CB = 2 * LB;                    // bytes per MFMA K-chunk of a row
  constexpr int ROWB = DPAD * T::kBytes;        // bytes per row
  constexpr int NC = ROWB / CB;                 // K-chunks per row
  constexpr int RS = ROWB + 16;                 // LDS row stride (bank spread)
  constexpr int BM = kRowsPerChunk;
  constexpr int BQ = 4 * TN * 32;
  constexpr int UNITS = BM * ROWB / 16;         // 16-byte units per chunk
  constexpr int UPT = (UNITS + kBlock - 1) / kBlock;
  constexpr int UROW = ROWB / 16;               // units per row
  static_assert(ROWB % CB == 0, "row must be whole K-chunks");

  __shared__ __attribute__((aligned(16))) char sx[2][BM * RS];
  __shared__ __attribute__((aligned(16))) float sn[2][BM];
  __shared__ int se[2][BM];

  // XCD-aware renumbering (see header).
  const int per_xcd = (int)(gridDim.x >> 3);
  const int logical = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
  if (logical >= a.num_qtiles * a.num_slices) return;
  const int qtile = logical % a.num_qtiles;
  const int slice = logical / a.num_qtiles;
  const int nq = a.nq_dev ? min(*a.nq_dev, a.nq) : a.nq;
  if (qtile * BQ >= nq) return;  // whole workgroup leaves together

  const int64_t r_begin = (int64_t)slice * a.rows_per_slice;
  int64_t r_end = r_begin + a.rows_per_slice;
  if (r_end > a.m) r_end = a.m;
  if (r_begin >= r_end) return;  // whole workgroup leaves together

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int h = lane >> 5, col = lane & 31;

  // Query fragments and thresholds of this lane's TN query columns.
  Frag<LB> qf[TN][NC];
  float thr[TN];
  int qsc[TN];
  int qrow[TN];
#pragma unroll
  for (int nt = 0; nt < TN; ++nt) {
    const int qr = qtile * BQ + (wave * TN + nt) * 32 + col;
    qrow[nt] = qr;
    if (qr < nq) {
      const char* qp = reinterpret_cast<const char*>(a.q) + (int64_t)qr * ROWB;
#pragma unroll
      for (int c = 0; c < NC; ++c) qf[nt][c] = load_frag<LB>(qp + CB * c + LB * h);
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
  const int64_t row_bytes_g = a.ldx * T::kBytes;
  i32x4 pre[UPT];
  float pn = 0.0f;
  int pe = 127;

  auto gload = [&](int64_t c0) {
#pragma unroll
    for (int t = 0; t < UPT; ++t) {
      const int u = tid + kBlock * t;
      i32x4 v = {0, 0, 0, 0};
      if (UNITS % kBlock == 0 || u < UNITS) {
        const int r = u / UROW;
        const int cb = (u - r * UROW) * 16;
        const int64_t j = c0 + r;
        if (j < r_end && cb < a.dvalid_bytes)
          v = *reinterpret_cast<const i32x4*>(xb + j * a.stride * row_bytes_g + cb);
      }
      pre[t] = v;
    }
    if (tid < BM) {
      const int64_t j = c0 + tid;
      if (j < r_end) {
        pn = -0.5f * a.xnorm[j * a.stride];
        if (E == Elem::FP8) pe = e8m0_of(a.xscale[j * a.stride]);
      } else {
        pn = -INFINITY;
        pe = 127;
      }
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int t = 0; t < UPT; ++t) {
      const int u = tid + kBlock * t;
      if (UNITS % kBlock == 0 || u < UNITS) {
        const int r = u / UROW;
        const int cb = (u - r * UROW) * 16;
        *reinterpret_cast<i32x4*>(&sx[buf][r * RS + cb]) = pre[t];
      }
    }
    if (tid < BM) {
      sn[buf][tid] = pn;
      se[buf][tid] = pe;
    }
  };

  int64_t c0 = r_begin;
  gload(c0);
  sstore(0);
  __syncthreads();
  int buf = 0;

  for (; c0 < r_end; c0 += BM) {
    const bool has_next = c0 + BM < r_end;
    if (has_next) gload(c0 + BM);

    f32x16 acc[2][TN];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      // C-init rows of this lane: mt*32 + 8*(i/4) + 4*h + (i%4), i = 0..15.
      f32x16 ci;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const f32x4 v = *reinterpret_cast<const f32x4*>(&sn[buf][mt * 32 + 8 * j + 4 * h]);
        ci[4 * j + 0] = v[0];
        ci[4 * j + 1] = v[1];
        ci[4 * j + 2] = v[2];
        ci[4 * j + 3] = v[3];
      }
      const int xsc = E == Elem::FP8 ? se[buf][mt * 32 + col] : 127;
      const char* ap = &sx[buf][(mt * 32 + col) * RS + LB * h];
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const Frag<LB> af = load_frag<LB>(ap + CB * c);
#pragma unroll
        for (int nt = 0; nt < TN; ++nt)
          acc[mt][nt] = mfma_chunk<E>(af, qf[nt][c], c == 0 ? ci : acc[mt][nt], xsc, qsc[nt]);
      }
    }

    // Filter: survivors are rows with acc >= thr (key = -2 acc <= tau).
    float mx[2][TN];
    bool hit = false;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int nt = 0; nt < TN; ++nt) {
        float m = acc[mt][nt][0];
#pragma unroll
        for (int i = 1; i < 16; ++i) m = fmaxf(m, acc[mt][nt][i]);
        mx[mt][nt] = m;
        hit |= m >= thr[nt];
      }
    if (__any(hit)) {
#pragma unroll
      for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int nt = 0; nt < TN; ++nt) {
          if (mx[mt][nt] >= thr[nt]) {
            const int qr = qrow[nt];
            int* cnt = a.counts + qr;
            uint64_t* dst = a.cand + (int64_t)qr * a.cap;
#pragma unroll
            for (int i = 0; i < 16; ++i) {
              const float v = acc[mt][nt][i];
              if (v >= thr[nt]) {
                const int64_t row = c0 + mt * 32 + 8 * (i >> 2) + 4 * h + (i & 3);
                const int pos = atomicAdd(cnt, 1);
                if (pos < a.cap) dst[pos] = pack_cand(-2.0f * v, (int32_t)(row * a.stride));
              }
            }
          }
        }
    }

    if (has_next) sstore(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
}

// K-looped FP32 filter for dimensions above the register-resident kernels'
// range (dpad > 256, a multiple of kKPanel). Neither operand stays in
// registers: each workgroup takes kLoopBQ queries x one slice of rows, and
// walks the slice in chunks of kLoopBM rows; for every chunk the contraction
// runs over the padded dimension in panels of kKPanel elements, each panel of
// rows AND of queries staged through LDS (double buffered: the next panel's
// loads are in flight in registers while the current one is multiplied). The
// first MFMA of every chunk takes C = -||x||^2/2, as in filter_kernel, and the
// epilogue (max-reduce, compare, survivor append) is the same.
// Wave w owns query tile w (32 queries) and all four 32-row tiles of a chunk.
constexpr int kKPanel = 32;   // elements per panel (128 bytes of FP32)
constexpr int kLoopBM = 128;  // rows per chunk
constexpr int kLoopBQ = 128;  // queries per workgroup

__global__ __launch_bounds__(kBlock, 2) void filter_kloop_kernel(FilterArgs a, int dpad) {
  constexpr int PB = kKPanel * 4;          // bytes of a row panel
  constexpr int RS = PB + 16;              // LDS row stride
  constexpr int NCP = PB / 32;             // 32-byte K-chunks per panel
  constexpr int UROW = PB / 16;            // 16-byte units per row panel
  constexpr int XU = kLoopBM * UROW / kBlock;  // row units per thread
  constexpr int QU = kLoopBQ * UROW / kBlock;  // query units per thread
  static_assert(kLoopBM * UROW % kBlock == 0 && kLoopBQ * UROW % kBlock == 0, "units per thread");

  __shared__ __attribute__((aligned(16))) char sx[2][kLoopBM * RS];
  __shared__ __attribute__((aligned(16))) char sq[2][kLoopBQ * RS];
  __shared__ __attribute__((aligned(16))) float sn[kLoopBM];

  const int per_xcd = (int)(gridDim.x >> 3);
  const int logical = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
  if (logical >= a.num_qtiles * a.num_slices) return;
  const int qtile = logical % a.num_qtiles;
  const int slice = logical / a.num_qtiles;
  const int nq = a.nq_dev ? min(*a.nq_dev, a.nq) : a.nq;
  if (qtile * kLoopBQ >= nq) return;  // whole workgroup leaves together
  const int64_t r_begin = (int64_t)slice * a.rows_per_slice;
  int64_t r_end = r_begin + a.rows_per_slice;
  if (r_end > a.m) r_end = a.m;
  if (r_begin >= r_end) return;  // whole workgroup leaves together

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int h = lane >> 5, col = lane & 31;
  const int qr = qtile * kLoopBQ + wave * 32 + col;  // this lane's query column
  const float thr = qr < nq ? fmaxf(-0.5f * a.tau[qr], -FLT_MAX) : INFINITY;

  const char* xb = reinterpret_cast<const char*>(a.x);
  const char* qb = reinterpret_cast<const char*>(a.q);
  const int64_t xrow_bytes = a.ldx * 4;
  const int64_t qrow_bytes = (int64_t)dpad * 4;
  const int npanels = dpad / kKPanel;
  i32x4 px[XU], pq[QU];

  // Registers <- panel p of chunk c0 (rows past r_end and bytes past the
  // row's data read as zero; queries past nq as zero).
  auto gload = [&](int64_t c0, int p) {
#pragma unroll
    for (int t = 0; t < XU; ++t) {
      const int u = tid + kBlock * t;
      const int r = u / UROW;
      const int cb = p * PB + (u - r * UROW) * 16;
      const int64_t j = c0 + r;
      i32x4 v = {0, 0, 0, 0};
      if (j < r_end && cb < a.dvalid_bytes)
        v = *reinterpret_cast<const i32x4*>(xb + j * a.stride * xrow_bytes + cb);
      px[t] = v;
    }
#pragma unroll
    for (int t = 0; t < QU; ++t) {
      const int u = tid + kBlock * t;
      const int r = u / UROW;
      const int cb = p * PB + (u - r * UROW) * 16;
      const int qq = qtile * kLoopBQ + r;
      i32x4 v = {0, 0, 0, 0};
      if (qq < nq) v = *reinterpret_cast<const i32x4*>(qb + (int64_t)qq * qrow_bytes + cb);
      pq[t] = v;
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int t = 0; t < XU; ++t) {
      const int u = tid + kBlock * t;
      const int r = u / UROW;
      *reinterpret_cast<i32x4*>(&sx[buf][r * RS + (u - r * UROW) * 16]) = px[t];
    }
#pragma unroll
    for (int t = 0; t < QU; ++t) {
      const int u = tid + kBlock * t;
      const int r = u / UROW;
      *reinterpret_cast<i32x4*>(&sq[buf][r * RS + (u - r * UROW) * 16]) = pq[t];
    }
  };

  int64_t c0 = r_begin;
  int p = 0;
  gload(c0, 0);
  sstore(0);
  int buf = 0;
  f32x16 acc[4];
  for (;;) {
    if (p == 0) {
      // Norms of this chunk (the previous chunk's epilogue read them before
      // the barrier that ended it).
      if (tid < kLoopBM) {
        const int64_t j = c0 + tid;
        sn[tid] = j < r_end ? -0.5f * a.xnorm[j * a.stride] : -INFINITY;
      }
    }
    __syncthreads();
    // Prefetch the next panel (of this chunk or the next one).
    int64_t nc0 = c0;
    int np = p + 1;
    if (np == npanels) {
      np = 0;
      nc0 = c0 + kLoopBM;
    }
    const bool has_next = nc0 < r_end;
    if (has_next) gload(nc0, np);
    if (p == 0) {
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const f32x4 v = *reinterpret_cast<const f32x4*>(&sn[mt * 32 + 8 * j + 4 * h]);
          acc[mt][4 * j + 0] = v[0];
          acc[mt][4 * j + 1] = v[1];
          acc[mt][4 * j + 2] = v[2];
          acc[mt][4 * j + 3] = v[3];
        }
    }
    const char* bp = &sq[buf][(wave * 32 + col) * RS + 16 * h];
#pragma unroll
    for (int c = 0; c < NCP; ++c) {
      const Frag<16> bf = load_frag<16>(bp + 32 * c);
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
        const Frag<16> af = load_frag<16>(&sx[buf][(mt * 32 + col) * RS + 16 * h + 32 * c]);
        acc[mt] = mfma_chunk<Elem::F32>(af, bf, acc[mt], 127, 127);
      }
    }
    if (p + 1 == npanels) {
      // Epilogue of the chunk: survivors are rows with acc >= thr.
      float mx[4];
      bool hit = false;
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
        float m = acc[mt][0];
#pragma unroll
        for (int i = 1; i < 16; ++i) m = fmaxf(m, acc[mt][i]);
        mx[mt] = m;
        hit |= m >= thr;
      }
      if (__any(hit)) {
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
          if (mx[mt] >= thr) {
            int* cnt = a.counts + qr;
            uint64_t* dst = a.cand + (int64_t)qr * a.cap;
#pragma unroll
            for (int i = 0; i < 16; ++i) {
              const float v = acc[mt][i];
              if (v >= thr) {
                const int64_t row = c0 + mt * 32 + 8 * (i >> 2) + 4 * h + (i & 3);
                const int pos = atomicAdd(cnt, 1);
                if (pos < a.cap) dst[pos] = pack_cand(-2.0f * v, (int32_t)(row * a.stride));
              }
            }
          }
        }
      }
    }
    if (!has_next) break;
    // Everyone is done reading buf ^ 1 (the barrier at the top of this step),
    // so the next panel goes there; the barrier at the top of the next step
    // publishes it.
    sstore(buf ^ 1);
    buf ^= 1;
    c0 = nc0;
    p = np;
  }
}

template <Elem E, int DPAD>
constexpr int tn_for() {
  return DPAD * ElemTraits<E>::kBytes <= 384 ? 2 : 1;
}

template <Elem E, int DPAD>
void launch_one(const FilterArgs& a, hipStream_t st) {
  constexpr int TN = tn_for<E, DPAD>();
  const int total = a.num_qtiles * a.num_slices;
  const unsigned grid = (unsigned)round_up(total, 8);
  filter_kernel<E, DPAD, TN><<<grid, kBlock, 0, st>>>(a);
  RPLSH_HIP(hipGetLastError());
}

void launch_kloop(int dpad, const FilterArgs& a, hipStream_t st) {
  RPLSH_REQUIRE(dpad % kKPanel == 0 && dpad > 256, "filter: K-looped dimension %d", dpad);
  const int total = a.num_qtiles * a.num_slices;
  const unsigned grid = (unsigned)round_up(total, 8);
  filter_kloop_kernel<<<grid, kBlock, 0, st>>>(a, dpad);
  RPLSH_HIP(hipGetLastError());
}

// Plain path, small batches (at most kSmallMaxQ queries): a streaming FP32
// contraction on the vector ALUs. The FP32 MFMA computes 32 query columns per
// instruction, so for a handful of queries it spends 32x the needed work at a
// quarter of the BF16 rate and the kernel is not HBM-bound; here every
// product is one the batch needs. Eight lanes share a row (lane part p takes
// columns 4p + 32i, 16-byte loads: one load instruction covers 8 rows x 128
// contiguous bytes), the queries (zero padded to dpad) sit in LDS and are read
// as broadcasts, the eight partial sums are added across the lanes, and
// lane p tests and emits queries j = p, p + 8. The key is an approximation in
// a different summation order than the MFMA kernels' (-||x||^2/2 added last),
// within the same forward-error bound the certificate uses, so the re-rank
// and certificate keep the results exact.
constexpr int kSmallMaxQ = 16;

template <int NQ>
__global__ __launch_bounds__(256) void filter_small_kernel(FilterArgs a, int dpad) {
  extern __shared__ __attribute__((aligned(16))) float sq[];  // [NQ][dpad]
  const int nq = a.nq_dev ? min(*a.nq_dev, a.nq) : a.nq;    // <= NQ
  const float* qg = static_cast<const float*>(a.q);
  for (int e = threadIdx.x; e < NQ * dpad; e += blockDim.x) sq[e] = e / dpad < nq ? qg[e] : 0.0f;
  __syncthreads();
  float thr[NQ];
#pragma unroll
  for (int j = 0; j < NQ; ++j) thr[j] = j < nq ? fmaxf(-0.5f * a.tau[j], -FLT_MAX) : INFINITY;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int p = lane & 7, rr = lane >> 3;
  const int ni = dpad / 32;
  const float* xb = static_cast<const float*>(a.x);
  const int64_t groups = (a.m - a.row_begin + 7) / 8;
  for (int64_t g = (int64_t)blockIdx.x * 4 + wave; g < groups; g += (int64_t)gridDim.x * 4) {
    const int64_t r = a.row_begin + g * 8 + rr;  // logical row
    const bool valid = r < a.m;
    const int64_t rp = (valid ? r : a.m - 1) * a.stride;  // physical row (any valid one past the end)
    const float* xr = xb + rp * a.ldx;
    float acc[NQ];
#pragma unroll
    for (int j = 0; j < NQ; ++j) acc[j] = 0.0f;
    for (int i0 = 0; i0 < ni; i0 += 8) {
      f32x4 xv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int col = 4 * p + 32 * (i0 + u);
        xv[u] = i0 + u < ni && col < a.ldx ? *reinterpret_cast<const f32x4*>(xr + col) : f32x4{0.0f, 0.0f, 0.0f, 0.0f};
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if (i0 + u >= ni) break;  // uniform
        const int col = 4 * p + 32 * (i0 + u);
#pragma unroll
        for (int j = 0; j < NQ; ++j) {
          const f32x4 qv = *reinterpret_cast<const f32x4*>(sq + j * dpad + col);
          acc[j] = fmaf(xv[u][0], qv[0], acc[j]);
          acc[j] = fmaf(xv[u][1], qv[1], acc[j]);
          acc[j] = fmaf(xv[u][2], qv[2], acc[j]);
          acc[j] = fmaf(xv[u][3], qv[3], acc[j]);
        }
      }
    }
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
      acc[j] += __shfl_xor(acc[j], 1, 64);
      acc[j] += __shfl_xor(acc[j], 2, 64);
      acc[j] += __shfl_xor(acc[j], 4, 64);
    }
    if (!valid) continue;
    const float nh = -0.5f * a.xnorm[rp];
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
      if (j % 8 != p || j >= nq) continue;
      const float v = nh + acc[j];
      if (v >= thr[j]) {
        const int pos = atomicAdd(a.counts + j, 1);
        if (pos < a.cap) a.cand[(int64_t)j * a.cap + pos] = pack_cand(-2.0f * v, (int32_t)rp);
      }
    }
  }
}

template <int NQ>
void launch_small_one(int dpad, const FilterArgs& a, hipStream_t st) {
  const int64_t groups = div_up(a.m - a.row_begin, 8);
  const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(div_up(groups, 4), 2048));
  filter_small_kernel<NQ><<<grid, 256, (size_t)NQ * dpad * 4, st>>>(a, dpad);
  RPLSH_HIP(hipGetLastError());
}

// The small-batch kernel takes the launch when the batch has at most
// kSmallMaxQ queries and their rows fit in 64 KB of LDS.
bool launch_small(int dpad, const FilterArgs& a, hipStream_t st) {
  if (a.nq > kSmallMaxQ || a.dense || a.tile_rank) return false;
  const int nq_cap = a.nq <= 1 ? 1 : a.nq <= 2 ? 2 : a.nq <= 4 ? 4 : a.nq <= 8 ? 8 : 16;
  if ((size_t)nq_cap * dpad * 4 > 65536) return false;
  switch (nq_cap) {
    case 1: launch_small_one<1>(dpad, a, st); break;
    case 2: launch_small_one<2>(dpad, a, st); break;
    case 4: launch_small_one<4>(dpad, a, st); break;
    case 8: launch_small_one<8>(dpad, a, st); break;
    default: launch_small_one<16>(dpad, a, st); break;
  }
  return true;
}

template <Elem E>
void launch_elem(int dpad, const FilterArgs& a, hipStream_t st) {
  if constexpr (E == Elem::F32) {
    if (launch_small(dpad, a, st)) return;
    if (dpad > 256) return launch_kloop(dpad, a, st);
  }
  switch (dpad) {
    case 64: return launch_one<E, 64>(a, st);
    case 128: return launch_one<E, 128>(a, st);
    case 256: return launch_one<E, 256>(a, st);
    default: break;
  }
  if constexpr (E != Elem::FP8) {
    switch (dpad) {
      case 32: return launch_one<E, 32>(a, st);
      case 96: return launch_one<E, 96>(a, st);
      case 160: return launch_one<E, 160>(a, st);
      case 192: return launch_one<E, 192>(a, st);
      default: break;
    }
  }
  fail(RPLSH_ERR_UNSUPPORTED, "filter: padded dimension %d not instantiated", dpad);
}

}  // namespace

int filter_dpad(Elem e, int d) {
  if (d <= 0) return 0;
  if (d > 256) {
    // FP32 only: the K-looped kernel, dimension padded to whole panels.
    if (e != Elem::F32 || d > kMaxKLoopDim) return 0;
    return (int)round_up(d, kKPanel);
  }
  if (e == Elem::FP8) {
    // 64-element K-chunks.
    if (d <= 64) return 64;
    if (d <= 128) return 128;
    return 256;
  }
  static const int sizes[] = {32, 64, 96, 128, 160, 192, 256};
  for (int s : sizes)
    if (d <= s) return s;
  return 0;
}

int filter_block_queries(Elem e, int dpad) {
  if (e == Elem::F32 && dpad > 256) return kLoopBQ;
  const int bytes = e == Elem::F32 ? 4 : (e == Elem::BF16 ? 2 : 1);
  return 4 * 32 * (dpad * bytes <= 384 ? 2 : 1);
}

void launch_filter(Elem e, int dpad, const FilterArgs& a, hipStream_t st) {
  switch (e) {
    case Elem::F32: return launch_elem<Elem::F32>(dpad, a, st);
    case Elem::BF16: return launch_elem<Elem::BF16>(dpad, a, st);
    case Elem::FP8: return launch_elem<Elem::FP8>(dpad, a, st);
  }
}

}  // namespace rplsh
