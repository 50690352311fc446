// Cluster-residual candidate copy for tightly clustered data (BF16 / FP8).
//
// A BF16 or FP8 copy of the raw rows cannot order neighbours inside a tight
// cluster: the rounding error of q.x grows with |q| |x|, which is of the
// order of the cluster's position, while neighbour distances are of the
// order of its spread. The residual copy encodes each row relative to a
// centroid instead, and the query relative to the same centroid:
//
//   x = c + r,  s = q - c,
//   ||x||^2 - 2 q.x = ||r||^2 + (||c||^2 - 2 q.c) - 2 s.r,
//
// where ||r||^2 and q.c - ||c||^2/2 are exact FP32 terms (the MFMA's C
// operand) and only s.r goes through the low-precision product. Both s and r
// are of the order of the spread when q is near c, which is where the
// ordering matters; far from c the error grows but those rows are far above
// any query's threshold.
//
// Build (add time): centroids by farthest-point seeding on a sample of the
// rows (every cluster of a well-separated mixture gets one before any gets
// two), refined by Lloyd steps on the sample; every row is assigned to its
// nearest centroid by the plain FP32 search; rows are reordered so that each
// 32-row tile holds rows of one centroid (partial tiles padded with rows whose
// C operand is -inf), each centroid's tiles split into slices of at most
// kSliceTiles. The copy holds the residuals in the packed tile layout
// (filter_packed.hip) with their own per-row scales (FP8) and -||r||^2/2.
//
// Besides the copy of every row, copies of every 8th, 64th, ... row of the
// centroid order (systematic samples: each centroid keeps its share) serve
// the sample levels, so that every level's keys -- and so the thresholds
// handed from level to level -- come from the same residual arithmetic.
//
// Search: every level runs the LDS-staged packed filter in its residual mode
// (filter_packed.hip) over one copy, one workgroup per (query block, slice):
// it forms s = q - c for its queries (FP8: a power-of-two scale per query and
// centroid), adds q.c - ||c||^2/2 to the C operand and emits the candidates
// with their original row numbers. The exact re-rank then reads the FP32
// rows as for every other form.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "encode.hpp"
#include "engine.hpp"
#include "kernels.hpp"
#include "packed_mma.hpp"

namespace rplsh {

namespace {

constexpr int kSliceTiles = 256;  // tiles per slice (one centroid each)

// ---- build kernels ------------------------------------------------------

// One wave per sample row: mind[i] = min(mind[i], ||s_i - c||^2); each block
// leaves the largest mind (lowest row on ties) of its rows in bval / bidx.
__global__ __launch_bounds__(256) void kcenter_update_kernel(const float* __restrict__ s, int64_t S, int d,
                                                             const float* __restrict__ c, float* __restrict__ mind,
                                                             float* __restrict__ bval, int* __restrict__ bidx) {
  __shared__ float wv[4];
  __shared__ int wi[4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  float best = -1.0f;
  int besti = INT32_MAX;
  for (int64_t i = (int64_t)blockIdx.x * 4 + w; i < S; i += (int64_t)gridDim.x * 4) {
    float acc = 0.0f;
    for (int j = lane; j < d; j += 64) {
      const float t = s[i * d + j] - c[j];
      acc = fmaf(t, t, acc);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    const float m = fminf(mind[i], acc);
    if (lane == 0) mind[i] = m;
    if (m > best) {  // rows ascend within a wave: the first maximum is the lowest row
      best = m;
      besti = (int)i;
    }
  }
  if (lane == 0) {
    wv[w] = best;
    wi[w] = besti;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float b = wv[0];
    int bi = wi[0];
    for (int u = 1; u < 4; ++u)
      if (wv[u] > b || (wv[u] == b && wi[u] < bi)) {
        b = wv[u];
        bi = wi[u];
      }
    bval[blockIdx.x] = b;
    bidx[blockIdx.x] = bi;
  }
}

// One block: the farthest row over all blocks (lowest row on ties) becomes
// centroid `out` (a copy of that sample row).
__global__ __launch_bounds__(256) void kcenter_pick_kernel(const float* __restrict__ bval,
                                                           const int* __restrict__ bidx, int nb,
                                                           const float* __restrict__ s, int64_t S, int d,
                                                           float* __restrict__ out) {
  __shared__ float sv[256];
  __shared__ int si[256];
  float b = -2.0f;
  int bi = INT32_MAX;
  for (int u = threadIdx.x; u < nb; u += blockDim.x)
    if (bval[u] > b || (bval[u] == b && bidx[u] < bi)) {
      b = bval[u];
      bi = bidx[u];
    }
  sv[threadIdx.x] = b;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      const float v2 = sv[threadIdx.x + o];
      const int i2 = si[threadIdx.x + o];
      if (v2 > sv[threadIdx.x] || (v2 == sv[threadIdx.x] && i2 < si[threadIdx.x])) {
        sv[threadIdx.x] = v2;
        si[threadIdx.x] = i2;
      }
    }
    __syncthreads();
  }
  const int row = si[0] >= 0 && si[0] < S ? si[0] : 0;  // no row compared (NaN distances): row 0
  for (int j = threadIdx.x; j < d; j += blockDim.x) out[j] = s[(int64_t)row * d + j];
}

// One wave per padded row: -||r||^2/2 (-inf for padding rows) and, for FP8,
// the power-of-two scale of r = x[rid[p]] - cent[tcl[p / 32]].
__global__ __launch_bounds__(256) void resid_rows_kernel(const float* __restrict__ x, int64_t ldx, int d,
                                                         const float* __restrict__ cent,
                                                         const int32_t* __restrict__ rid,
                                                         const int32_t* __restrict__ tcl, int64_t npad,
                                                         float* __restrict__ rnh, float* __restrict__ rscale) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int64_t p = (int64_t)blockIdx.x * 4 + w; p < npad; p += (int64_t)gridDim.x * 4) {
    const int32_t src = rid[p];
    if (src < 0) {
      if (lane == 0) {
        rnh[p] = -INFINITY;
        if (rscale) rscale[p] = 1.0f;
      }
      continue;
    }
    const float* xr = x + (int64_t)src * ldx;
    const float* cr = cent + (int64_t)tcl[p >> 5] * d;
    float acc = 0.0f, mx = 0.0f;
    for (int j = lane; j < d; j += 64) {
      const float t = xr[j] - cr[j];
      acc = fmaf(t, t, acc);
      mx = fmaxf(mx, fabsf(t));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      acc += __shfl_xor(acc, o, 64);
      mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    }
    if (lane == 0) {
      rnh[p] = -0.5f * acc;
      if (rscale) rscale[p] = e4m3_row_scale(mx);
    }
  }
}

// Residual rows in the packed tile layout: one thread per 16-byte piece (the
// piece order of encode_packed_kernel).
template <Elem E, int DPAD>
__global__ void resid_pack_kernel(const float* __restrict__ x, int64_t ldx, int d, const float* __restrict__ cent,
                                  const int32_t* __restrict__ rid, const int32_t* __restrict__ tcl,
                                  int64_t tiles, const float* __restrict__ rscale, char* __restrict__ out) {
  using G = Geometry<E, DPAD>;
  constexpr int LB = G::LB;
  constexpr int PIECES = G::TILE / 16;
  const int64_t total = tiles * PIECES;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = e / PIECES;
    const int w = (int)(e % PIECES);
    const int c = w / (G::CHUNK / 16);
    const int l = (w % (G::CHUNK / 16)) / (LB / 16);
    const int sub = w % (LB / 16);
    const int64_t row = t * 32 + (l & 31);
    const int k0 = c * PackTraits<E>::kChunkElems + (l >> 5) * (PackTraits<E>::kChunkElems / 2) +
                   sub * (16 / PackTraits<E>::kBytes);
    i32x4 piece = {0, 0, 0, 0};
    const int32_t src = rid[row];
    if (src >= 0) {
      const float* xr = x + (int64_t)src * ldx;
      const float* cr = cent + (int64_t)tcl[t] * d;
      if (E == Elem::BF16) {
        uint16_t v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = k0 + i < d ? bf16_rne(xr[k0 + i] - cr[k0 + i]) : 0;
        __builtin_memcpy(&piece, v, 16);
      } else {
        const float s = rscale[row];
        uint8_t v[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = k0 + i < d ? e4m3_rne_sat((xr[k0 + i] - cr[k0 + i]) / s) : 0;
        __builtin_memcpy(&piece, v, 16);
      }
    }
    *reinterpret_cast<i32x4*>(out + t * G::TILE + (int64_t)w * 16) = piece;
  }
}

template <typename F>
auto resid_dispatch(Elem e, int dpad, F&& f) {
  if (e == Elem::BF16) {
    switch (dpad) {
      case 32: return f.template operator()<Elem::BF16, 32>();
      case 64: return f.template operator()<Elem::BF16, 64>();
      case 96: return f.template operator()<Elem::BF16, 96>();
      case 128: return f.template operator()<Elem::BF16, 128>();
      case 160: return f.template operator()<Elem::BF16, 160>();
      case 192: return f.template operator()<Elem::BF16, 192>();
      case 256: return f.template operator()<Elem::BF16, 256>();
    }
  } else if (e == Elem::FP8) {
    switch (dpad) {
      case 64: return f.template operator()<Elem::FP8, 64>();
      case 128: return f.template operator()<Elem::FP8, 128>();
      case 256: return f.template operator()<Elem::FP8, 256>();
    }
  }
  fail(RPLSH_ERR_UNSUPPORTED, "residual copy: element kind %d, padded dimension %d", (int)e, dpad);
}

unsigned grid_of(int64_t work, int64_t per_block) {
  int64_t g = div_up(std::max<int64_t>(work, 1), per_block);
  return (unsigned)std::min<int64_t>(g, 65536);
}

struct ResidPack {
  const rplsh_bf* b;
  const rplsh_bf::ResidCopy* c;
  hipStream_t st;
  template <Elem E, int DPAD>
  int operator()() const {
    using G = Geometry<E, DPAD>;
    const int64_t tiles = c->npad / 32;
    resid_pack_kernel<E, DPAD><<<grid_of(tiles * (G::TILE / 16), 256), 256, 0, st>>>(
        b->x, b->ldx, b->dim, b->rcent, c->rid, c->tcl, tiles, c->scale, static_cast<char*>(c->enc));
    RPLSH_HIP(hipGetLastError());
    return 0;
  }
};

}  // namespace

// ---- host build ----------------------------------------------------------

void resid_free(rplsh_bf* b) {
  for (auto& c : b->rcopies) {
    void* bufs[] = {c.rid, c.tcl, c.slice_t0, c.slice_cl, c.enc, c.scale, c.nh};
    for (void* p : bufs)
      if (p) (void)hipFree(p);
  }
  b->rcopies.clear();
  if (b->rcent) (void)hipFree(b->rcent);
  b->rcent = nullptr;
  b->rk = 0;
  b->rn = 0;
  b->rprec = -1;
}

namespace {

// Centroids of b's rows (K of them, dense [K][d] FP32 on device) by
// farthest-point seeding on a sample, then Lloyd steps on the sample.
void build_centroids(rplsh_bf* b, int K, float* cent, hipStream_t st, L2Scratch& sc) {
  const int d = b->dim;
  const int64_t n = b->n;
  const int64_t S = std::min<int64_t>(n, std::max<int64_t>(64ll * K, 65536));
  // Sample rows: evenly spread, deterministic.
  std::vector<int32_t> rows((size_t)S);
  for (int64_t i = 0; i < S; ++i) rows[(size_t)i] = (int32_t)((__int128)i * n / S);
  DevBuf drows, samp, mind, bval, bidx, aidx, adist;
  int32_t* dr = drows.as<int32_t>((size_t)S);
  RPLSH_HIP(hipMemcpyAsync(dr, rows.data(), (size_t)S * 4, hipMemcpyHostToDevice, st));
  float* ds = samp.as<float>((size_t)S * d);
  launch_gather_rows(b->x, b->ldx, dr, (int)S, d, ds, st);

  // Farthest-point seeding: centroid 0 is sample row 0.
  float* md = mind.as<float>((size_t)S);
  launch_fill_f32(md, S, INFINITY, st);
  RPLSH_HIP(hipMemcpyAsync(cent, ds, (size_t)d * 4, hipMemcpyDeviceToDevice, st));
  const int nb = (int)std::min<int64_t>(1024, div_up(S, 4));
  float* bv = bval.as<float>((size_t)nb);
  int* bi = bidx.as<int>((size_t)nb);
  for (int k = 1; k < K; ++k) {
    kcenter_update_kernel<<<nb, 256, 0, st>>>(ds, S, d, cent + (int64_t)(k - 1) * d, md, bv, bi);
    kcenter_pick_kernel<<<1, 256, 0, st>>>(bv, bi, nb, ds, S, d, cent + (int64_t)k * d);
  }
  RPLSH_HIP(hipGetLastError());

  // Lloyd steps on the sample (means in double on the host: order fixed).
  constexpr int kLloyd = 2;
  std::vector<float> hs((size_t)S * d), hc((size_t)K * d);
  RPLSH_HIP(hipMemcpyAsync(hs.data(), ds, hs.size() * 4, hipMemcpyDeviceToHost, st));
  RPLSH_HIP(hipMemcpyAsync(hc.data(), cent, hc.size() * 4, hipMemcpyDeviceToHost, st));
  RPLSH_HIP(hipStreamSynchronize(st));
  int32_t* ai = aidx.as<int32_t>((size_t)S);
  float* ad = adist.as<float>((size_t)S);
  std::vector<int32_t> ha((size_t)S);
  std::vector<double> sum((size_t)K * d);
  std::vector<int64_t> cnt((size_t)K);
  for (int it = 0; it < kLloyd; ++it) {
    search_l2_device(cent, K, d, nullptr, ds, S, 1, ai, ad, 0, st, sc);
    RPLSH_HIP(hipMemcpy(ha.data(), ai, (size_t)S * 4, hipMemcpyDeviceToHost));
    std::fill(sum.begin(), sum.end(), 0.0);
    std::fill(cnt.begin(), cnt.end(), 0);
    for (int64_t i = 0; i < S; ++i) {
      const int c = ha[(size_t)i];
      if (c < 0 || c >= K) continue;
      cnt[(size_t)c] += 1;
      const float* r = hs.data() + (size_t)i * d;
      double* sp = sum.data() + (size_t)c * d;
      for (int j = 0; j < d; ++j) sp[j] += r[j];
    }
    for (int c = 0; c < K; ++c)
      if (cnt[(size_t)c] > 0)
        for (int j = 0; j < d; ++j)
          hc[(size_t)c * d + j] = (float)(sum[(size_t)c * d + j] / (double)cnt[(size_t)c]);
    RPLSH_HIP(hipMemcpyAsync(cent, hc.data(), hc.size() * 4, hipMemcpyHostToDevice, st));
  }
  RPLSH_HIP(hipStreamSynchronize(st));
}

}  // namespace

// Sampling strides of the copies: 1 (every row), then 8, 64, ... while a
// copy keeps at least kMinSampleRows rows.
constexpr int64_t kResidStrideStep = 8;
constexpr int64_t kMinSampleRows = 2048;

void resid_build(rplsh_bf* b, hipStream_t st) {
  const Elem e = b->cfg.precision == RPLSH_PREC_FP8 ? Elem::FP8 : Elem::BF16;
  const int d = b->dim;
  const int dpad = b->dpad;
  const int64_t n = b->n;
  const int K = (int)std::min<int64_t>(b->cfg.residual_clusters, std::max<int64_t>(1, n / 32));
  RPLSH_REQUIRE(K > 0 && n > 0, "residual copy: no rows or no centroids");

  if (b->rn != n || b->rk != K) {
    // Centroids and the assignment of every row.
    resid_free(b);
    RPLSH_HIP(hipMalloc((void**)&b->rcent, (size_t)K * d * sizeof(float)));
    L2Scratch sc;
    sc.wss.resize(4);
    build_centroids(b, K, b->rcent, st, sc);
    constexpr int64_t kBatch = 65536;
    std::vector<int32_t> assign((size_t)n);
    {
      DevBuf qb, ai, ad;
      int32_t* dai = ai.as<int32_t>((size_t)kBatch);
      float* dad = ad.as<float>((size_t)kBatch);
      for (int64_t r0 = 0; r0 < n; r0 += kBatch) {
        const int64_t m = std::min(kBatch, n - r0);
        const float* qp = b->x + r0 * b->ldx;
        if (b->ldx != d) {
          float* t = qb.as<float>((size_t)kBatch * d);
          launch_pad_rows(qp, m, d, b->ldx, d, t, st);
          qp = t;
        }
        search_l2_device(b->rcent, K, d, nullptr, qp, m, 1, dai, dad, 0, st, sc);
        RPLSH_HIP(hipMemcpy(assign.data() + r0, dai, (size_t)m * 4, hipMemcpyDeviceToHost));
      }
    }
    // Rows in centroid order (stable: original order within a centroid).
    std::vector<int64_t> first((size_t)K + 1, 0);
    for (int64_t r = 0; r < n; ++r) {
      const int c = assign[(size_t)r];
      RPLSH_REQUIRE(c >= 0 && c < K, "residual copy: row %lld has no centroid", (long long)r);
      first[(size_t)c + 1] += 1;
    }
    for (int c = 0; c < K; ++c) first[(size_t)c + 1] += first[(size_t)c];
    std::vector<int32_t> order((size_t)n);
    {
      std::vector<int64_t> fill(first.begin(), first.end() - 1);
      for (int64_t r = 0; r < n; ++r) order[(size_t)fill[(size_t)assign[(size_t)r]]++] = (int32_t)r;
    }
    // One copy per stride: positions p = 0, S, 2S, ... of the centroid order
    // (a systematic sample, so each centroid keeps its share of the rows),
    // packed into whole tiles of one centroid, slices of <= kSliceTiles.
    for (int64_t S = 1; b->rcopies.size() < (size_t)DbView::kMaxResidCopies; S *= kResidStrideStep) {
      if (S > 1 && n / S < kMinSampleRows) break;
      rplsh_bf::ResidCopy cp;
      cp.stride = S;
      std::vector<int32_t> rid, tcl, s_t0, s_cl;
      for (int c = 0; c < K; ++c) {
        const int64_t p0 = div_up(first[(size_t)c], S) * S;  // first sampled position in this centroid
        const size_t tile_begin = tcl.size();
        for (int64_t p = p0; p < first[(size_t)c + 1]; p += S) rid.push_back(order[(size_t)p]);
        while (rid.size() % 32) rid.push_back(-1);
        while (tcl.size() * 32 < rid.size()) tcl.push_back(c);
        for (size_t t = tile_begin; t < tcl.size(); t += kSliceTiles) {
          s_t0.push_back((int32_t)t);
          s_cl.push_back(c);
        }
      }
      s_t0.push_back((int32_t)tcl.size());
      cp.rows = 0;
      for (int32_t v : rid) cp.rows += v >= 0;
      cp.npad = (int64_t)rid.size();
      cp.slices = (int64_t)s_cl.size();
      RPLSH_REQUIRE(cp.npad > 0 && cp.npad < (1ll << 31), "residual copy: bad size");
      RPLSH_HIP(hipMalloc((void**)&cp.rid, rid.size() * 4));
      RPLSH_HIP(hipMalloc((void**)&cp.tcl, tcl.size() * 4));
      RPLSH_HIP(hipMalloc((void**)&cp.slice_t0, s_t0.size() * 4));
      RPLSH_HIP(hipMalloc((void**)&cp.slice_cl, s_cl.size() * 4));
      RPLSH_HIP(hipMalloc((void**)&cp.nh, rid.size() * 4));
      b->rcopies.push_back(cp);  // owned from here on (freed by resid_free)
      RPLSH_HIP(hipMemcpy(cp.rid, rid.data(), rid.size() * 4, hipMemcpyHostToDevice));
      RPLSH_HIP(hipMemcpy(cp.tcl, tcl.data(), tcl.size() * 4, hipMemcpyHostToDevice));
      RPLSH_HIP(hipMemcpy(cp.slice_t0, s_t0.data(), s_t0.size() * 4, hipMemcpyHostToDevice));
      RPLSH_HIP(hipMemcpy(cp.slice_cl, s_cl.data(), s_cl.size() * 4, hipMemcpyHostToDevice));
    }
    b->rk = K;
    b->rn = n;
    b->rprec = -1;
  }

  if (b->rprec != b->cfg.precision) {
    // Residual rows of every copy for the current element kind.
    for (auto& cp : b->rcopies) {
      if (cp.enc) (void)hipFree(cp.enc);
      if (cp.scale) (void)hipFree(cp.scale);
      cp.enc = nullptr;
      cp.scale = nullptr;
      RPLSH_HIP(hipMalloc(&cp.enc, (size_t)(cp.npad / 32 * packed_tile_bytes(e, dpad))));
      if (e == Elem::FP8) RPLSH_HIP(hipMalloc((void**)&cp.scale, (size_t)cp.npad * 4));
      resid_rows_kernel<<<grid_of(cp.npad, 4), 256, 0, st>>>(b->x, b->ldx, d, b->rcent, cp.rid, cp.tcl, cp.npad,
                                                               cp.nh, cp.scale);
      RPLSH_HIP(hipGetLastError());
      resid_dispatch(e, dpad, ResidPack{b, &cp, st});
    }
    RPLSH_HIP(hipStreamSynchronize(st));
    b->rprec = b->cfg.precision;
  }
}

}  // namespace rplsh
