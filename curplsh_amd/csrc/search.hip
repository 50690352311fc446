// Search orchestration and the IndexBF / searchL2Distance entry points.
//
// Reference shape (src/internal/search.cu:79-216): query norms, tiled
// cuBLAS GEMMs into a global distance buffer, a select per tile, +||q||^2,
// index offsets and a cross-tile merge -- about 260 launches and ~1 GiB of
// temporaries per call at 1M x 128, 10k queries.
//
// Here a search is a short, fixed pipeline per query batch:
//   1. canonical FP32 query norms; queries encoded for the candidate
//      contraction (FP32 padded copy, BF16, or FP8 e4m3 + power-of-two scale);
//   2. filter levels: the fused MFMA filter runs over strided samples of the
//      database with growing density and then over all rows; after each
//      sample level a radix select turns the rho-th smallest survivor into the
//      next level's per-query threshold tau. The last level keeps every row
//      with approximate relative distance <= tau (about c k' rows per query)
//      and selects the k' best by (approximate key, index). This is exact
//      whenever the last level keeps between k' and the capacity rows: a row
//      the filter drops has key > tau >= every kept key. Packed copies read
//      the sample levels from a sample copy of whole tiles; the last sample
//      level's survivors that pass the final threshold are carried into the
//      full level's lists and the full level skips those tiles, so it keeps
//      the same rows as a scan of every tile (LDS filter variants);
//   3. re-rank: canonical FP32 distances of the k' candidates, top k;
//      precision forms re-order the boundary band in FP64; the plain path
//      certifies that the MFMA error margin cannot hide a better row;
//   4. the rare queries whose last level kept too few / too many rows, or
//      whose certificate failed, are gathered and re-run with a wider margin.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "engine.hpp"
#include "kernels.hpp"
#include "rplsh.h"

namespace rplsh {

namespace {

constexpr float kUnitRoundoff = 5.9604645e-08f;  // 2^-24

}  // namespace

struct Level {
  int64_t stride;
  int64_t m;
  int rank_next;  // rho of the threshold for the next level; 0 on the last
};

// Rows the last level aims to keep per query: c_factor * k', but at least
// kMinTarget, so that a small k' (k = 1: k' = 4) still leaves the last sample
// level a stride of several rows and a threshold rank of at least 8; with
// 3 * 4 = 12 rows the last sample collapsed to rank 1 and about a quarter of
// the queries kept fewer than k' rows and were re-run.
constexpr int64_t kMinTarget = 64;
static int64_t survivor_target(int kp, int c_factor) {
  return std::max<int64_t>((int64_t)c_factor * kp, kMinTarget);
}

// Rank of the last level's threshold within its sample. Automatic
// (rho_final <= 0): larger ranks give a tighter survivor count for the same
// risk but need a denser last sample (stride c*k'/rho); keep that stride >= 16
// with rho in [8, 32].
static int final_rank(int kp, int c_factor, int rho_final) {
  return rho_final > 0 ? rho_final
                       : (int)std::min<int64_t>(32, std::max<int64_t>(8, survivor_target(kp, c_factor) / 16));
}

// Probability that a query keeps fewer than k' rows at the full level: the
// threshold is the rho-th smallest key of a sample of stride s = c*k'/rho
// (c = the survivor target over k'), so
// the full level keeps about s * Gamma(rho, 1) rows, fewer than k' when
// Gamma(rho, 1) < rho / c, i.e. P(Poisson(rho / c) >= rho).
static double short_probability(int rho, double c_factor) {
  if (c_factor <= 1.0) return 1.0;
  const double lam = (double)rho / c_factor;
  double term = std::exp(-lam), below = 0.0;
  for (int i = 0; i < rho; ++i) {
    below += term;
    term *= lam / (i + 1);
  }
  return std::max(0.0, 1.0 - below);
}

// Filter levels for n rows and k' candidates (see header).
static std::vector<Level> plan_levels(int64_t n, int kp, int c_factor, int rho_final) {
  const int64_t kM0 = 1024;   // rows of the first (unfiltered) sample
  const int kRhoFinal = final_rank(kp, c_factor, rho_final);
  const int kRhoMid = 16;     // rank of intermediate thresholds
  const double kGrowth = 16;  // density step between sample levels
  std::vector<Level> L;
  const int64_t s0 = std::max<int64_t>(1, n / kM0);
  if (s0 == 1) {
    L.push_back({1, n, 0});
    return L;
  }
  const int64_t target = survivor_target(kp, c_factor);
  const int64_t s_pf = std::max<int64_t>(1, target / kRhoFinal);
  std::vector<int64_t> strides;
  if (s0 <= s_pf) {
    strides.push_back(s0);
  } else {
    const int nl = (int)std::ceil(std::log((double)s0 / (double)s_pf) / std::log(kGrowth));
    for (int i = 0; i <= nl; ++i) {
      double s = (double)s0 * std::pow((double)s_pf / (double)s0, (double)i / nl);
      int64_t si = std::max<int64_t>(1, (int64_t)std::llround(s));
      if (strides.empty() || si < strides.back()) strides.push_back(si);
    }
  }
  for (size_t i = 0; i < strides.size(); ++i) {
    const int64_t s = strides[i];
    if (s == 1) break;
    const int64_t m = div_up(n, s);
    int rank;
    if (i + 1 < strides.size() && strides[i + 1] > 1) {
      rank = kRhoMid;
    } else {
      rank = (int)std::min<int64_t>(m, div_up(target, s));
    }
    L.push_back({s, m, std::max(1, rank)});
  }
  // A level's threshold of rank r at stride s lets about r * s / s' rows of the
  // next sample (stride s') through; keep that at least twice the next
  // level's own rank, or that rank is not reached and the full level keeps
  // fewer rows than planned (two sample levels 2.3x apart, rank 64 last:
  // ~37 rows reached it, and 30 % of the queries of k' = 4000 at 300K rows
  // kept fewer than k' and were re-run).
  for (size_t i = L.size(); i-- > 1;) {
    const Level& nx = L[i];
    Level& cur = L[i - 1];
    if (nx.stride <= 1) continue;  // the full level: no rank of its own
    const int64_t need = div_up(2 * (int64_t)nx.rank_next * nx.stride, cur.stride);
    cur.rank_next = (int)std::min<int64_t>(cur.m, std::max<int64_t>(cur.rank_next, need));
  }
  L.push_back({1, n, 0});
  return L;
}

static int64_t level_cap(const std::vector<Level>& L, int kp, int c_factor) {
  int64_t cap = 4096;
  cap = std::max<int64_t>(cap, L.front().stride > 1 ? L.front().m : 0);
  cap = std::max<int64_t>(cap, 4 * survivor_target(kp, c_factor));
  if (L.size() == 1) cap = std::max<int64_t>(cap, L.front().m);
  return cap;
}

// Levels over the cluster-residual copies: the sparsest copy first (every
// row kept: the threshold is +inf), then denser sample copies down to the
// last one -- the sparsest whose threshold rank target / stride is at least
// the final rank -- skipping copies while a threshold of rank kRhoMid still
// leaves at most kRhoMid * kResidMaxRatio survivors at the next level taken,
// then the copy of every row. (Each level costs an encoding of the queries per
// centroid besides its rows; a wider step costs survivor emission instead:
// measured on 20M x 128, 1,024 centres, a step of 8 between copies is the
// cheaper of the two.)
// Sample levels of at most kF32SampleRows rows score every stride-th FP32 row
// exactly (the plain path's filter) instead of a residual copy: their keys are
// as consistent with the residual ones as the residual copies' (both close to
// the exact key), and a level that small costs its per-centroid query
// encoding, not its rows, on a residual copy.
struct ResidLevel {
  int copy;            // residual copy (f32_stride == 0)
  int64_t f32_stride;  // > 0: every f32_stride-th FP32 row instead
  int rank_next;       // rank of the threshold for the next level; 0 on the last
};
constexpr int kResidRhoMid = 16;
constexpr int64_t kResidMaxRatio = 8;
constexpr int64_t kF32SampleRows = 65536;
static std::vector<ResidLevel> resid_levels(const DbView& db, int kp, int c_factor, int rho_final) {
  const int64_t target = survivor_target(kp, c_factor);
  const int rf = final_rank(kp, c_factor, rho_final);
  int last = 0;
  for (int j = db.nrcopies - 1; j >= 1; --j)
    if (db.rcopy[j].stride * rf <= target) {
      last = j;
      break;
    }
  std::vector<ResidLevel> L;
  if (last > 0) {
    std::vector<int> chain{db.nrcopies - 1};
    while (chain.back() > last) {
      // The sparsest copy within kResidMaxRatio of the current one, at least `last`.
      int nxt = last;
      for (int j = last; j < chain.back(); ++j)
        if (db.rcopy[chain.back()].stride / db.rcopy[j].stride <= kResidMaxRatio) {
          nxt = j;
          break;
        }
      chain.push_back(nxt);
    }
    for (size_t i = 0; i < chain.size(); ++i) {
      const auto& c = db.rcopy[chain[i]];
      const int rank = chain[i] > last ? kResidRhoMid : (int)std::min<int64_t>(c.rows, div_up(target, c.stride));
      const int64_t f32 = div_up(db.n, c.stride) <= kF32SampleRows ? c.stride : 0;
      L.push_back({chain[i], f32, std::max(1, rank)});
    }
  }
  L.push_back({0, 0, 0});
  return L;
}

// Candidate slots per query of a batch (every level's survivors fit, the
// first level's all of its rows).
static int64_t batch_cap(const DbView& db, int kp, const SearchParams& p) {
  if (db.nrcopies > 0) {
    const std::vector<ResidLevel> L = resid_levels(db, kp, p.c_factor, p.rho);
    int64_t cap = std::max<int64_t>(4096, 4 * survivor_target(kp, p.c_factor));
    cap = std::max<int64_t>(cap, 2 * kResidRhoMid * kResidMaxRatio);
    const ResidLevel& f = L.front();
    return std::max<int64_t>(cap, f.f32_stride > 0 ? div_up(db.n, f.f32_stride) : db.rcopy[f.copy].rows);
  }
  return level_cap(plan_levels(db.n, kp, p.c_factor, p.rho), kp, p.c_factor);
}

// One filter pass. A sample level (stride > 1) of a packed index whose sample
// copy holds at least lv.m rows reads the first lv.m rows of that copy
// instead of every stride-th row (its candidates only set a threshold, their
// row numbers are never used). `dense`: tau is +inf for every query (first
// pass), so every row survives and is written to its own slot.
// `skip_tiles` > 0 (full level after a carried sample level): the tiles that
// level scored are skipped; `row_begin` > 0 (sample level after a carried
// one): rows below it are not scored. Either way the lists already hold the
// carried survivors.
// The rows a level reads: the index (every stride-th row) or, for a sample
// level of a packed index, the first lv.m rows of its sample copy.
static void filter_source(FilterArgs& fa, const DbView& db, const Level& lv) {
  fa.x = db.xenc;
  fa.ldx = db.ldenc;
  fa.dvalid_bytes = db.dvalid_bytes;
  fa.xnorm = db.xnorm;
  fa.xscale = db.xscale;
  fa.xnh = db.xnh;
  fa.xnh_inf = db.xnh_inf;
  fa.m = lv.m;
  fa.stride = lv.stride;
  if (db.packed && lv.stride > 1 && db.srows >= lv.m) {
    fa.x = db.senc;
    fa.xscale = db.sscale;
    fa.xnorm = db.snorm;
    fa.xnh = db.snh;
    fa.xnh_inf = db.snh_inf;
    fa.stride = 1;
  }
}

static void run_filter(const DbView& db, const Level& lv, bool dense, const void* qenc,
                       const float* qscale, const float* tau, int nq, const int* nq_dev, int cap,
                       int* counts, uint64_t* cand, int64_t target_wgs, int variant,
                       int64_t skip_tiles, int64_t row_begin, hipStream_t st, bool l2_slices = false) {
  FilterArgs fa{};
  filter_source(fa, db, lv);
  fa.dense = dense && db.packed && variant != kFilterDirect && lv.m <= cap ? 1 : 0;
  // The dense pass runs in the narrow LDS kernel's dense instantiation,
  // whatever the level's variant: its words do not depend on the tiling.
  if (fa.dense) variant = kFilterLdsNarrow;
  fa.q = qenc;
  fa.qscale = qscale;
  fa.tau = tau;
  fa.nq = nq;
  fa.nq_dev = nq_dev;
  const int bq = db.packed ? filter_packed_block_queries(db.elem, db.dpad, variant)
                           : filter_block_queries(db.elem, db.dpad);
  fa.num_qtiles = (int)div_up(nq, bq);
  RPLSH_REQUIRE(row_begin >= 0 && row_begin % 32 == 0 && row_begin < lv.m && (row_begin == 0 || !dense),
                "filter: bad first row");
  fa.row_begin = row_begin;
  const int64_t rows = lv.m - row_begin;
  int64_t slices = std::max<int64_t>(1, div_up(target_wgs, fa.num_qtiles));
  if (l2_slices && db.packed) {
    const int64_t row_bytes = (int64_t)db.dpad * (db.elem == Elem::FP8 ? 1 : 2);
    slices = std::max<int64_t>(slices, div_up(rows * row_bytes, kL2SliceBytes));
  }
  slices = std::min<int64_t>(slices, div_up(rows, 256));
  slices = std::max<int64_t>(slices, 1);
  fa.rows_per_slice = round_up(div_up(rows, slices), 64);
  fa.num_slices = (int)div_up(rows, fa.rows_per_slice);
  fa.cap = cap;
  fa.counts = counts;
  fa.cand = cand;
  if (skip_tiles > 0) {
    RPLSH_REQUIRE(db.trank && db.rank_tiles > 0 && lv.stride == 1 && !fa.dense, "tile skip: bad level");
    fa.tile_rank = db.trank;
    fa.rank_tiles = db.rank_tiles;
    fa.skip_tiles = skip_tiles;
  }
  if (skip_tiles == 0 && row_begin == 0) launch_fill_i32(counts, nq, fa.dense ? (int)lv.m : 0, st);
  if (db.packed) {
    launch_filter_packed(db.elem, db.dpad, variant, fa, st);
  } else {
    launch_filter(db.elem, db.dpad, fa, st);
  }
}

// One level over a cluster-residual copy: every padded row, one workgroup
// per (query block, slice); the lists start empty.
static void run_filter_resid(const DbView& db, int copy, const float* q, const float* tau, int nq,
                             const int* nq_dev, int cap, int* counts, uint64_t* cand, int variant,
                             hipStream_t st) {
  const DbView::ResidView& c = db.rcopy[copy];
  FilterArgs fa{};
  fa.x = c.enc;
  fa.ldx = db.dpad;
  fa.dvalid_bytes = db.dvalid_bytes;
  fa.xnorm = c.nh;
  fa.xscale = c.scale;
  fa.m = c.npad;
  fa.stride = 1;
  fa.rid = c.rid;
  fa.slice_t0 = c.slice_t0;
  fa.slice_cl = c.slice_cl;
  fa.num_slices = (int)c.slices;
  fa.cent = db.rcent;
  fa.qf32 = q;
  fa.ldq = db.d;
  fa.d = db.d;
  fa.tau = tau;
  fa.nq = nq;
  fa.nq_dev = nq_dev;
  fa.num_qtiles = (int)div_up(nq, filter_resid_block_queries(db.elem, db.dpad, variant));
  fa.cap = cap;
  fa.counts = counts;
  fa.cand = cand;
  launch_fill_i32(counts, nq, 0, st);
  launch_filter_resid(db.elem, db.dpad, variant, fa, st);
}

// Carry: a level whose predecessor read whole tiles of the sample copy starts
// from the predecessor's survivors that pass its own threshold and does not
// score those rows again (LDS variants): a sample level starts at the
// predecessor's last whole tile, the full level skips the predecessor's tiles.
// Every level then keeps exactly the rows it would have kept scanning all.
// Returns the number of sample tiles carried into level li (0: no carry).
static int64_t carried_tiles(const DbView& db, const std::vector<Level>& L, size_t li, int variant) {
  if (li == 0 || li >= L.size()) return 0;
  const Level& lp = L[li - 1];
  const Level& lv = L[li];
  const bool lds = variant == kFilterLdsWide || variant == kFilterLdsNarrow || variant == kFilterLdsWide8;
  if (!lds || !db.packed || !db.trank || lp.stride <= 1 || db.srows < lp.m) return 0;
  if (lv.stride == 1) return lp.m / 32;                       // into the full level
  if (db.srows >= lv.m && lp.m / 32 * 32 < lv.m) return lp.m / 32;  // into a sample level
  return 0;
}

// Filter variant of level li of nlevels: the sample levels may run another
// kernel than the full level (same arithmetic and survivors).
static int level_variant(const SearchParams& p, size_t li, size_t nlevels) {
  return li + 1 < nlevels && p.sample_variant >= 0 ? p.sample_variant : p.variant;
}

// Capturable part of a batch: query norms and encoding, filter levels,
// selects, re-rank and the failed-query list, all on st with no host
// synchronisation (nothing is allocated once ws has seen this size). When
// nq_dev is set it holds the live query count (<= nq); rows beyond it are
// skipped by every kernel.
void enqueue_batch(const DbView& db, const float* q, int nq, const int* nq_dev,
                   const SearchParams& p, int32_t* out_i, float* out_d, Workspace& ws,
                   hipStream_t st, PhaseEvents* pe) {
  const int d = db.d;
  const int kp = (int)std::min<int64_t>(p.kp, db.n);
  const bool resid = db.nrcopies > 0;
  const std::vector<Level> L = resid ? std::vector<Level>{} : plan_levels(db.n, kp, p.c_factor, p.rho);
  const std::vector<ResidLevel> RL = resid ? resid_levels(db, kp, p.c_factor, p.rho) : std::vector<ResidLevel>{};
  const int64_t cap64 = batch_cap(db, kp, p);
  RPLSH_REQUIRE(cap64 < (1ll << 31), "candidate capacity overflow");
  const int cap = (int)cap64;
  const size_t nlevels = resid ? RL.size() : L.size();
  auto rec = [&](int i) {
    if (pe) pe->rec(i, st);
  };

  // 1. query norms and encoding
  float* qn = ws.qnorm.as<float>(nq);
  launch_l2norm(q, nq, d, d, qn, 1, st);
  const int eb = db.elem == Elem::F32 ? 4 : (db.elem == Elem::BF16 ? 2 : 1);
  void* qenc = ws.qenc.ensure((size_t)nq * db.dpad * eb);
  float* qscale = nullptr;
  switch (db.elem) {
    case Elem::F32: launch_pad_rows(q, nq, d, d, db.dpad, static_cast<float*>(qenc), st); break;
    case Elem::BF16: launch_encode_bf16(q, nq, d, d, db.dpad, static_cast<uint16_t*>(qenc), st); break;
    case Elem::FP8:
      qscale = ws.qscale.as<float>(nq);
      launch_encode_e4m3(q, nq, d, d, db.dpad, static_cast<uint8_t*>(qenc), qscale, st);
      break;
  }

  // Residual search: the exact FP32 view and padded FP32 queries for its
  // smallest sample levels.
  DbView db32;
  float* qpad32 = nullptr;
  if (resid && RL.front().f32_stride > 0) {
    db32.x = db.x;
    db32.ldx = db.ldx;
    db32.n = db.n;
    db32.d = d;
    db32.xnorm = db.xnorm;
    db32.xn_max = db.xn_max;
    db32.elem = Elem::F32;
    db32.dpad = filter_dpad(Elem::F32, d);
    db32.xenc = db.x;
    db32.ldenc = db.ldx;
    db32.dvalid_bytes = (int)(db.ldx * 4);
    qpad32 = ws.qpad.as<float>((size_t)nq * db32.dpad);
    launch_pad_rows(q, nq, d, d, db32.dpad, qpad32, st);
  }

  // 2. filter levels
  float* tau = ws.tau.as<float>(nq);
  float* tau2 = ws.tau2.as<float>(nq);
  int* counts = ws.counts.as<int>(nq);
  uint64_t* cand = ws.cand.as<uint64_t>((size_t)nq * cap);
  uint64_t* topk = ws.topk.as<uint64_t>((size_t)nq * kp);
  float* bound = ws.bound.as<float>(nq);
  int* status = ws.status.as<int>(nq);
  float* relk = ws.relk.as<float>(nq);
  BatchCounters* ctr = ws.ctr.as<BatchCounters>(1);
  static_assert(sizeof(BatchCounters) == 4 * sizeof(int), "counters cleared as four words");
  launch_fill_i32(reinterpret_cast<int*>(ctr), 4, 0, st);
  launch_fill_f32(tau, nq, INFINITY, st);
  launch_fill_i32(status, nq, 0, st);
  rec(0);
  for (size_t li = 0; li < nlevels; ++li) {
    const bool last = li + 1 == nlevels;
    if (last) rec(1);
    int rank_next;
    if (resid && RL[li].f32_stride > 0) {
      const Level lv{RL[li].f32_stride, div_up(db.n, RL[li].f32_stride), RL[li].rank_next};
      run_filter(db32, lv, false, qpad32, nullptr, tau, nq, nq_dev, cap, counts, cand, p.target_wgs, 0, 0, 0,
                 st);
      rank_next = RL[li].rank_next;
    } else if (resid) {
      run_filter_resid(db, RL[li].copy, q, tau, nq, nq_dev, cap, counts, cand, p.variant, st);
      rank_next = RL[li].rank_next;
    } else {
      const int64_t carried = carried_tiles(db, L, li, level_variant(p, li, nlevels));
      run_filter(db, L[li], li == 0, qenc, qscale, tau, nq, nq_dev, cap, counts, cand, p.target_wgs,
                 level_variant(p, li, nlevels), last ? carried : 0, last ? 0 : 32 * carried, st,
                 last && p.l2_slices);
      rank_next = L[li].rank_next;
    }
    if (last) rec(2);
    SelectArgs sa{};
    sa.cand = cand;
    sa.counts = counts;
    sa.cap = cap;
    sa.nq = nq;
    sa.nq_dev = nq_dev;
    sa.tau_in = tau;
    if (rank_next > 0) {
      sa.K = rank_next;
      sa.topk_mode = 0;
      sa.tau_out = tau2;
      sa.status = nullptr;  // a loose intermediate threshold only costs time
      launch_select(sa, st);
      std::swap(tau, tau2);
      const int64_t carry_next = resid ? 0 : carried_tiles(db, L, li + 1, level_variant(p, li + 1, nlevels));
      if (carry_next > 0) {
        CarryArgs ca{};
        ca.cand = cand;
        ca.counts = counts;
        ca.cap = cap;
        ca.nq = nq;
        ca.nq_dev = nq_dev;
        ca.tau = tau;
        ca.row_limit = 32 * carry_next;
        const bool into_full = li + 2 == nlevels;
        ca.step = into_full ? db.sstep : 0;
        ca.tiles = db.stiles;
        ca.status = status;
        launch_carry(ca, st);
      }
    } else {
      sa.K = kp;
      sa.topk_mode = 1;
      sa.topk_out = topk;
      sa.bound_out = bound;
      sa.status = status;
      launch_select(sa, st);
      rec(3);
    }
  }
  launch_sum_counts(counts, nq, cap, &ctr->survivors, st);

  // 3. re-rank
  RerankArgs ra{};
  ra.x = db.x;
  ra.ldx = db.ldx;
  ra.xnorm = db.xnorm;
  ra.d = d;
  ra.q = q;
  ra.ldq = d;
  ra.qnorm = qn;
  ra.topk = topk;
  ra.kp = kp;
  ra.k = p.k;
  ra.nq = nq;
  ra.nq_dev = nq_dev;
  ra.exact = p.exact;
  ra.bound = p.certify ? bound : nullptr;
  ra.eps_coeff = 4.0f * (float)(d + 4) * kUnitRoundoff;
  ra.band_coeff = 2.0f * (float)(d + 2) * kUnitRoundoff;
  ra.xn_max = db.xn_max;
  ra.fp64_boundary = p.fp64_boundary;
  ra.out_idx = out_i;
  ra.out_dist = out_d;
  ra.status = status;
  ra.relk_out = relk;
  ra.fp64_count = &ctr->fp64;
  rec(4);
  launch_rerank(ra, st);
  rec(5);

  // 4. failed queries
  int* flist = ws.flist.as<int>(nq);
  launch_collect_failed(status, nq, nq_dev, flist, &ctr->failed, st);
}

// The same batch as enqueue_batch, as argument records for the persistent
// kernel (persist.hip) instead of launches: the same levels, thresholds,
// capacities, selects and re-rank. Its filter scores every row of a level
// (no carry between levels, like the direct variant), which keeps exactly the
// rows the carrying variants keep, so the selected candidates and the results
// are the same. Filter units are 32 queries x rows_per_slice rows (a multiple
// of 128: the four waves of a unit take 32-row tiles in turn), about
// target_units per level.
void plan_persistent_batch(const DbView& db, const float* q, int nq, const SearchParams& p,
                           int32_t* out_i, float* out_d, Workspace& ws, int64_t target_units,
                           PersistBatchPlan& out) {
  RPLSH_REQUIRE(db.packed && db.elem != Elem::F32, "persistent kernel: packed BF16 / FP8 copies only");
  RPLSH_REQUIRE(nq > 0, "persistent kernel: empty batch");
  const int d = db.d;
  const int kp = (int)std::min<int64_t>(p.kp, db.n);
  RPLSH_REQUIRE(kp <= kPersistMaxKp, "persistent kernel: k' %d above %d", kp, kPersistMaxKp);
  const std::vector<Level> L = plan_levels(db.n, kp, p.c_factor, p.rho);
  const int64_t cap64 = level_cap(L, kp, p.c_factor);
  RPLSH_REQUIRE(cap64 < (1ll << 31), "candidate capacity overflow");
  const int cap = (int)cap64;
  const int eb = db.elem == Elem::BF16 ? 2 : 1;

  float* qn = ws.qnorm.as<float>(nq);
  void* qenc = ws.qenc.ensure((size_t)nq * db.dpad * eb);
  float* qscale = db.elem == Elem::FP8 ? ws.qscale.as<float>(nq) : nullptr;
  float* tau = ws.tau.as<float>(nq);
  float* tau2 = ws.tau2.as<float>(nq);
  int* counts = ws.counts.as<int>(nq);
  uint64_t* cand = ws.cand.as<uint64_t>((size_t)nq * cap);
  uint64_t* topk = ws.topk.as<uint64_t>((size_t)nq * kp);
  float* bound = ws.bound.as<float>(nq);
  int* status = ws.status.as<int>(nq);
  float* relk = ws.relk.as<float>(nq);
  int* flist = ws.flist.as<int>(nq);
  BatchCounters* ctr = ws.ctr.as<BatchCounters>(1);

  PrepArgs& pr = out.prep;
  pr = PrepArgs{};
  pr.q = q;
  pr.nq = nq;
  pr.d = d;
  pr.dpad = db.dpad;
  pr.fp8 = db.elem == Elem::FP8 ? 1 : 0;
  pr.qnorm = qn;
  pr.qenc = qenc;
  pr.qscale = qscale;
  pr.tau = tau;
  pr.status = status;
  pr.counts = counts;
  pr.ctr = reinterpret_cast<int*>(ctr);

  out.filter.clear();
  out.select.clear();
  out.count_reset.clear();
  for (size_t li = 0; li < L.size(); ++li) {
    const Level& lv = L[li];
    FilterArgs fa{};
    filter_source(fa, db, lv);
    fa.dense = li == 0 && lv.m <= cap ? 1 : 0;
    if (li == 0) pr.count_init = fa.dense ? (int)lv.m : 0;
    fa.q = qenc;
    fa.qscale = qscale;
    fa.tau = tau;
    fa.nq = nq;
    fa.num_qtiles = (int)div_up(nq, 32);
    int64_t slices = std::max<int64_t>(1, div_up(target_units, fa.num_qtiles));
    slices = std::min<int64_t>(slices, div_up(lv.m, 128));
    fa.rows_per_slice = round_up(div_up(lv.m, slices), 128);
    fa.num_slices = (int)div_up(lv.m, fa.rows_per_slice);
    fa.cap = cap;
    fa.counts = counts;
    fa.cand = cand;
    out.filter.push_back(fa);

    SelectArgs sa{};
    sa.cand = cand;
    sa.counts = counts;
    sa.cap = cap;
    sa.nq = nq;
    sa.tau_in = tau;
    if (lv.rank_next > 0) {
      sa.K = lv.rank_next;
      sa.topk_mode = 0;
      sa.tau_out = tau2;
      sa.status = nullptr;
      std::swap(tau, tau2);
      out.count_reset.push_back(0);  // the next level appends from zero
    } else {
      sa.K = kp;
      sa.topk_mode = 1;
      sa.topk_out = topk;
      sa.bound_out = bound;
      sa.status = status;
      out.count_reset.push_back(-1);  // the survivor total reads them
    }
    out.select.push_back(sa);
  }

  RerankTailArgs& rt = out.rerank;
  rt = RerankTailArgs{};
  RerankArgs& ra = rt.ra;
  ra.x = db.x;
  ra.ldx = db.ldx;
  ra.xnorm = db.xnorm;
  ra.d = d;
  ra.q = q;
  ra.ldq = d;
  ra.qnorm = qn;
  ra.topk = topk;
  ra.kp = kp;
  ra.k = p.k;
  ra.nq = nq;
  ra.exact = p.exact;
  ra.bound = p.certify ? bound : nullptr;
  ra.eps_coeff = 4.0f * (float)(d + 4) * kUnitRoundoff;
  ra.band_coeff = 2.0f * (float)(d + 2) * kUnitRoundoff;
  ra.xn_max = db.xn_max;
  ra.fp64_boundary = p.fp64_boundary;
  ra.out_idx = out_i;
  ra.out_dist = out_d;
  ra.status = status;
  ra.relk_out = relk;
  ra.fp64_count = &ctr->fp64;
  rt.counts = counts;
  rt.flist = flist;
  rt.failed = &ctr->failed;
  rt.survivors = &ctr->survivors;
}

// Host side after enqueue_batch: waits for st, reads the counters and re-runs
// the failed queries (eagerly, with wider margins) into out_i / out_d.
void finish_batch(const DbView& db, const float* q, int nq, const SearchParams& p,
                  int32_t* out_i, float* out_d, std::vector<Workspace>& wss, int depth,
                  hipStream_t st, rplsh_search_stats& stats, const PhaseEvents* pe,
                  const BatchCounters* ready) {
  Workspace& ws = wss.at(depth);
  const int d = db.d;
  const int kp = (int)std::min<int64_t>(p.kp, db.n);
  // One read-back of the batch's counters into the workspace's pinned buffer
  // (or a stack copy when the workspace must not allocate).
  BatchCounters stack_ctr{};
  const BatchCounters* hc = ready;
  if (!hc) {
    BatchCounters* t = ws.no_alloc ? &stack_ctr : ws.hctr.as<BatchCounters>(1);
    RPLSH_HIP(hipMemcpyAsync(t, ws.ctr.p, sizeof(BatchCounters), hipMemcpyDeviceToHost, st));
    RPLSH_HIP(hipStreamSynchronize(st));
    hc = t;
  }
  const int nfail = hc->failed, nfp64 = hc->fp64;
  const long long nsurv = hc->survivors;
  stats.boundary_fp64 += nfp64;
  if (depth == 0) {
    if (db.nrcopies > 0) {
      stats.levels = (int32_t)resid_levels(db, kp, p.c_factor, p.rho).size();
      stats.rows_filter_full = db.rcopy[0].npad;
    } else {
      const std::vector<Level> L = plan_levels(db.n, kp, p.c_factor, p.rho);
      stats.levels = (int32_t)L.size();
      stats.rows_filter_full = L.back().m - 32 * carried_tiles(db, L, L.size() - 1, p.variant);
    }
    stats.filter_variant = encode_variants(p.variant, p.sample_variant, p.l2_slices);
    stats.target_wgs = (int32_t)p.target_wgs;
    if (pe && pe->on) {
      stats.ms_filter_samples = pe->ms(0, 1);
      stats.ms_filter_full = pe->ms(1, 2);
      stats.ms_select = pe->ms(2, 3);
      stats.ms_rerank = pe->ms(4, 5);
    }
    stats.survivors = nsurv;
    stats.candidates = (int64_t)nq * kp;
    stats.kprime = kp;
  }
  if (nfail == 0) return;
  if (depth + 1 >= (int)wss.size())
    fail(RPLSH_ERR_STATE, "search: %d queries still unresolved after %d retries", nfail, depth);
  stats.retried_queries += nfail;

  // Re-run the failed subset with four times the survivor margin and, for
  // certification failures, four times the candidates.
  Workspace& sub = wss.at(depth + 1);
  float* qs = sub.qsub.as<float>((size_t)nfail * d);
  int32_t* is = sub.isub.as<int32_t>((size_t)nfail * p.k);
  float* ds = sub.dsub.as<float>((size_t)nfail * p.k);
  launch_gather_rows(q, d, static_cast<const int*>(ws.flist.p), nfail, d, qs, st);
  SearchParams p2 = p;
  p2.c_factor = p.c_factor * 4;
  p2.kp = (int)std::min<int64_t>(std::min<int64_t>(db.n, 4096), (int64_t)p.kp * 4);
  search_batch(db, qs, nfail, p2, is, ds, wss, depth + 1, st, stats);
  launch_scatter_results(is, ds, static_cast<const int*>(ws.flist.p), nfail, p.k, out_i, out_d, st);
  RPLSH_HIP(hipStreamSynchronize(st));
}

// One batch of nq device-resident FP32 queries [nq][d] -> device outputs.
void search_batch(const DbView& db, const float* q, int nq, const SearchParams& p,
                  int32_t* out_i, float* out_d, std::vector<Workspace>& wss, int depth,
                  hipStream_t st, rplsh_search_stats& stats) {
  if (nq <= 0) return;
  PhaseEvents pe;
  if (depth == 0 && !wss.at(0).no_alloc) pe.create();
  enqueue_batch(db, q, nq, nullptr, p, out_i, out_d, wss.at(depth), st, depth == 0 ? &pe : nullptr);
  finish_batch(db, q, nq, p, out_i, out_d, wss, depth, st, stats, depth == 0 ? &pe : nullptr);
}

static int plain_kprime(int k) { return std::max(2 * k, k + 32); }

size_t search_workspace_bytes(const DbView& db, int nq, const SearchParams& p) {
  const int kp = (int)std::min<int64_t>(p.kp, db.n);
  const size_t cap = (size_t)batch_cap(db, kp, p);
  const size_t eb = db.elem == Elem::F32 ? 4 : (db.elem == Elem::BF16 ? 2 : 1);
  size_t b = 0;
  b += (size_t)nq * 4 * 9;                 // norms, scales, tau x2, counts, bound, status, relk, list
  b += (size_t)nq * db.dpad * eb;          // encoded queries
  if (db.nrcopies > 0) b += (size_t)nq * filter_dpad(Elem::F32, db.d) * 4;  // padded FP32 queries
  b += (size_t)nq * cap * 8;               // candidate words
  b += (size_t)nq * kp * 8;                // top-k' words
  return b + 4096;
}

}  // namespace rplsh

using namespace rplsh;

// ------------------------------------------------------------------ IndexBF


hipStream_t bf_stream(rplsh_bf* b) {
  void* s = nullptr;
  if (rplsh_resources_default_stream(b->res, b->cfg.device, &s) != RPLSH_OK)
    fail(RPLSH_ERR_STATE, "default stream: %s", rplsh_last_error());
  return (hipStream_t)s;
}

static Elem elem_of(int precision) {
  switch (precision) {
    case RPLSH_PREC_FP32: return Elem::F32;
    case RPLSH_PREC_BF16: return Elem::BF16;
    case RPLSH_PREC_FP8: return Elem::FP8;
  }
  fail(RPLSH_ERR_INVALID_ARG, "unknown precision %d", precision);
}

static int elem_bytes(Elem e) { return e == Elem::F32 ? 4 : (e == Elem::BF16 ? 2 : 1); }

// (Re-)encode rows [from, n) of the candidate copy for the configured precision.
static void bf_free_sample(rplsh_bf* b) {
  if (b->senc) (void)hipFree(b->senc);
  if (b->trank) (void)hipFree(b->trank);
  b->trank = nullptr;
  b->trank_cap = 0;
  if (b->sscale) (void)hipFree(b->sscale);
  if (b->snorm) (void)hipFree(b->snorm);
  if (b->snh) (void)hipFree(b->snh);
  b->senc = nullptr;
  b->sscale = nullptr;
  b->snorm = nullptr;
  b->snh = nullptr;
  b->srows = b->scap = b->sstep = 0;
  b->sprec = -1;
}

// Sample copy for the sample levels: srows = max(n / 8, 32768) rows rounded
// down to whole tiles (at most the n / 32 whole tiles of the index), sample
// tile s = index tile (s * P) mod (n / 32) with P ~ 0.618 n / 32 coprime to
// it, so every prefix is a spread-out set of distinct tiles; trank maps each
// index tile back to its sample tile (the full level skips the tiles a
// carried sample level scored). Rebuilt whenever the rows change.
static void bf_encode_sample(rplsh_bf* b, Elem e, int dpad, hipStream_t st) {
  constexpr int64_t kMinRows = 32768;
  if (b->n < 2 * kMinRows) {
    bf_free_sample(b);
    return;
  }
  const int64_t tiles = b->n / 32;
  const int64_t sm = std::min<int64_t>(tiles, std::max<int64_t>(b->n / 8, kMinRows) / 32) * 32;
  if (sm > b->scap || b->sprec != b->cfg.precision) {
    bf_free_sample(b);
    const int64_t alloc_tiles = div_up(sm + sm / 4, 32);  // room to grow with the index
    RPLSH_HIP(hipMalloc(&b->senc, (size_t)(alloc_tiles * packed_tile_bytes(e, dpad))));
    if (e == Elem::FP8) RPLSH_HIP(hipMalloc((void**)&b->sscale, (size_t)alloc_tiles * 32 * sizeof(float)));
    RPLSH_HIP(hipMalloc((void**)&b->snorm, (size_t)alloc_tiles * 32 * sizeof(float)));
    RPLSH_HIP(hipMalloc((void**)&b->snh, (size_t)(alloc_tiles * 32 + 1) * sizeof(float)));
    b->scap = alloc_tiles * 32;
    b->sprec = b->cfg.precision;
  }
  const int64_t rank_tiles = div_up(b->n, 32);
  if (rank_tiles > b->trank_cap) {
    if (b->trank) (void)hipFree(b->trank);
    b->trank = nullptr;
    b->trank_cap = 0;
    const int64_t cap = rank_tiles + rank_tiles / 4;
    RPLSH_HIP(hipMalloc((void**)&b->trank, (size_t)cap * sizeof(int32_t)));
    b->trank_cap = cap;
  }
  int64_t p = std::max<int64_t>(1, (int64_t)std::llround(0.6180339887498949 * (double)tiles));
  while (std::gcd(p, tiles) != 1) ++p;
  b->srows = sm;
  b->sstep = p;
  launch_sample_rows(b->xn, b->xscale, tiles, sm, p, b->snorm, b->snh, b->scap, b->sscale, st);
  launch_encode_sample(e, dpad, b->x, tiles, b->dim, b->ldx, sm, p, b->senc, b->sscale, st);
  launch_tile_ranks(tiles, sm / 32, p, b->trank, rank_tiles, st);
}

static void bf_encode(rplsh_bf* b, int64_t from, hipStream_t st) {
  const Elem e = elem_of(b->cfg.precision);
  if (e == Elem::F32) {
    b->enc_precision = b->cfg.precision;
    b->dpad = filter_dpad(e, b->dim);
    bf_free_sample(b);
    resid_free(b);
    return;
  }
  const int dpad = filter_dpad(e, b->dim);
  RPLSH_REQUIRE(dpad > 0, "dimension %d not supported by precision %d", b->dim, b->cfg.precision);
  if (b->enc_precision != b->cfg.precision || !b->xenc) {
    if (b->xenc) RPLSH_HIP(hipFree(b->xenc));
    if (b->xscale) RPLSH_HIP(hipFree(b->xscale));
    if (b->xnh) RPLSH_HIP(hipFree(b->xnh));
    b->xenc = nullptr;
    b->xscale = nullptr;
    b->xnh = nullptr;
    if (b->capacity > 0) {
      const int64_t tiles = div_up(b->capacity, 32);
      RPLSH_HIP(hipMalloc(&b->xenc, (size_t)(tiles * packed_tile_bytes(e, dpad))));
      if (e == Elem::FP8) RPLSH_HIP(hipMalloc((void**)&b->xscale, (size_t)tiles * 32 * sizeof(float)));
      RPLSH_HIP(hipMalloc((void**)&b->xnh, (size_t)(tiles * 32 + 1) * sizeof(float)));
    }
    b->enc_precision = b->cfg.precision;
    b->dpad = dpad;
    from = 0;
  }
  if (b->n - from > 0) {
    launch_encode_packed(e, dpad, b->x, b->n, b->dim, b->ldx, from, b->xenc, b->xscale, st);
    launch_neg_half_norms(b->xn, from, b->n, b->xnh, div_up(b->capacity, 32) * 32, st);
    bf_encode_sample(b, e, dpad, st);
  }
  // The residual copy serves the full level of a multi-level search (more
  // than 2,048 rows; below that one dense level reads the regular copy).
  if (b->cfg.residual_clusters > 0 && b->n > 2048)
    resid_build(b, st);
  else
    resid_free(b);
}

static DbView bf_view(const rplsh_bf* b) {
  DbView v;
  v.x = b->x;
  v.ldx = b->ldx;
  v.n = b->n;
  v.d = b->dim;
  v.xnorm = b->xn;
  v.xn_max = b->xn_max;
  v.elem = elem_of(b->cfg.precision);
  v.dpad = b->dpad;
  if (v.elem == Elem::F32) {
    v.xenc = b->x;
    v.ldenc = b->ldx;
    v.dvalid_bytes = (int)(b->ldx * 4);
  } else {
    v.xenc = b->xenc;
    v.ldenc = b->dpad;
    v.dvalid_bytes = b->dpad * elem_bytes(v.elem);
    v.xscale = b->xscale;
    v.senc = b->senc;
    v.sscale = b->sscale;
    v.snorm = b->snorm;
    v.snh = b->snh;
    v.snh_inf = b->scap;
    v.xnh = b->xnh;
    v.xnh_inf = div_up(b->capacity, 32) * 32;
    v.srows = b->senc ? b->srows : 0;
    v.sstep = b->sstep;
    v.stiles = b->n / 32;
    v.trank = b->senc ? b->trank : nullptr;
    v.rank_tiles = div_up(b->n, 32);
    v.packed = true;
    if (!b->rcopies.empty() && b->rprec == b->cfg.precision && b->rn == b->n) {
      v.nrcopies = (int)b->rcopies.size();
      v.rcent = b->rcent;
      for (int j = 0; j < v.nrcopies; ++j) {
        const auto& c = b->rcopies[(size_t)j];
        auto& r = v.rcopy[j];
        r.enc = c.enc;
        r.scale = c.scale;
        r.nh = c.nh;
        r.rid = c.rid;
        r.slice_t0 = c.slice_t0;
        r.slice_cl = c.slice_cl;
        r.slices = c.slices;
        r.npad = c.npad;
        r.rows = c.rows;
        r.stride = c.stride;
      }
    }
  }
  return v;
}

static void bf_free_data(rplsh_bf* b) {
  if (b->x) (void)hipFree(b->x);
  if (b->xn) (void)hipFree(b->xn);
  if (b->xenc) (void)hipFree(b->xenc);
  if (b->xscale) (void)hipFree(b->xscale);
  bf_free_sample(b);
  resid_free(b);
  if (b->xnh) (void)hipFree(b->xnh);
  b->xnh = nullptr;
  b->x = nullptr;
  b->xn = nullptr;
  b->xenc = nullptr;
  b->xscale = nullptr;
  b->n = b->capacity = 0;
  b->xn_max = 0.0f;
  b->enc_precision = -1;
}

namespace rplsh {

DbView bf_db_view(const rplsh_bf* b) { return bf_view(b); }

// A tuning entry's variant field: a full-level variant 0..5, optionally with
// a sample-level variant 0..5 in bits 8-15 and the L2-sized slices bit 16
// (encode_variants).
static bool valid_variants(int v) {
  const int f = full_variant_of(v), sv = sample_variant_of(v);
  return v >= 0 && (v >> 17) == 0 && f >= kFilterLdsWide && f <= kFilterLdsWide8 && sv >= -1 &&
         sv <= kFilterLdsWide8;
}

SearchParams bf_search_params(const rplsh_bf* b, int k, int64_t nq) {
  SearchParams p;
  p.k = k;
  p.exact = 1;  // IndexBF::search always asks for exact distances (index_bf.cpp:68-69)
  const bool plain = elem_of(b->cfg.precision) == Elem::F32;
  if (plain) {
    p.kp = plain_kprime(k);
    p.certify = true;
    p.fp64_boundary = 0;
  } else {
    p.kp = k * b->cfg.kprime_factor;
    p.certify = false;
    p.fp64_boundary = b->cfg.fp64_boundary;
  }
  // Autotuned entry for this k with the nearest batch size.
  const rplsh_tune_result* best = nullptr;
  double bd = 0;
  for (const auto& t : b->tunes) {
    if (t.k != k) continue;
    const double dist = nq > 0 ? std::fabs(std::log((double)nq / (double)std::max<int64_t>(1, t.batch_nq))) : 0.0;
    if (!best || dist < bd) {
      best = &t;
      bd = dist;
    }
  }
  if (best) {
    if (!plain && best->kprime_factor > 0) p.kp = k * best->kprime_factor;
    if (best->target_wgs > 0) p.target_wgs = best->target_wgs;
    if (best->c_factor > 0) p.c_factor = best->c_factor;
    if (valid_variants(best->variant)) {
      p.variant = full_variant_of(best->variant);
      p.sample_variant = sample_variant_of(best->variant);
      p.l2_slices = l2_slices_of(best->variant);
    }
    if (best->rho > 0) p.rho = best->rho;
  }
  return p;
}

int bf_device_of(const rplsh_bf* b) { return b->cfg.device; }

void bf_begin_change(rplsh_bf* b) {
  if (b->serving.load() > 0)
    fail(RPLSH_ERR_STATE, "index is registered with a running server and cannot be changed");
  b->generation.fetch_add(1);
}

void bf_prepare(rplsh_bf* b) {
  if (b->enc_precision == b->cfg.precision || b->n == 0) return;
  DeviceScope scope(b->cfg.device);
  hipStream_t st = bf_stream(b);
  bf_encode(b, 0, st);
  RPLSH_HIP(hipStreamSynchronize(st));
}

}  // namespace rplsh

extern "C" {

void rplsh_bf_config_default(rplsh_bf_config* cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->device = 0;
  cfg->memory_space = 0;
  cfg->precision = RPLSH_PREC_FP32;
  cfg->kprime_factor = 4;
  cfg->fp64_boundary = 1;
  cfg->persistent_max_queries = 32;
  cfg->residual_clusters = 0;
}

rplsh_status rplsh_bf_create(rplsh_resources* res, int dim, const rplsh_bf_config* cfg,
                             rplsh_bf** out) {
  return guard([&] {
    RPLSH_REQUIRE(out, "out is NULL");
    RPLSH_REQUIRE(res, "resources is NULL");
    RPLSH_REQUIRE(dim > 0, "dimension must be > 0 (got %d)", dim);
    rplsh_bf_config c;
    rplsh_bf_config_default(&c);
    if (cfg) c = *cfg;
    int count = 0;
    rplsh_device_count(&count);
    RPLSH_REQUIRE(c.device >= 0 && c.device < count, "device %d out of range [0, %d)", c.device, count);
    RPLSH_REQUIRE(c.memory_space == 0, "only MemorySpace::Device is supported");
    (void)elem_of(c.precision);
    RPLSH_REQUIRE(filter_dpad(elem_of(c.precision), dim) > 0,
                  "dimension %d is not supported by precision %d (FP32: up to %d; BF16/FP8: up to 256)",
                  dim, c.precision, kMaxKLoopDim);
    if (c.kprime_factor <= 0) c.kprime_factor = 4;
    RPLSH_REQUIRE(c.residual_clusters >= 0, "residual_clusters must be >= 0");
    if (rplsh_resources_init_device(res, c.device) != RPLSH_OK)
      fail(RPLSH_ERR_STATE, "%s", rplsh_last_error());
    auto* b = new rplsh_bf();
    b->res = res;
    b->dim = dim;
    b->cfg = c;
    b->ldx = round_up(dim, 4);
    b->ws.resize(4);
    b->dpad = filter_dpad(elem_of(c.precision), dim);
    *out = b;
  });
}

rplsh_status rplsh_bf_destroy(rplsh_bf* b) {
  return guard([&] {
    if (!b) return;
    if (b->serving.load() > 0)
      fail(RPLSH_ERR_STATE, "index is registered with a running server and cannot be destroyed");
    DeviceScope scope(b->cfg.device);
    (void)hipDeviceSynchronize();
    bf_free_data(b);
    if (b->pst) persist_state_destroy(b->pst);
    delete b;
  });
}

rplsh_status rplsh_bf_train(rplsh_bf* b, int64_t, const float*) {
  return guard([&] { RPLSH_REQUIRE(b, "index is NULL"); });
}

rplsh_status rplsh_bf_reset(rplsh_bf* b) {
  return guard([&] {
    RPLSH_REQUIRE(b, "index is NULL");
    bf_begin_change(b);
    DeviceScope scope(b->cfg.device);
    RPLSH_HIP(hipStreamSynchronize(bf_stream(b)));
    bf_free_data(b);
  });
}

rplsh_status rplsh_bf_add(rplsh_bf* b, int64_t num, const float* data) {
  return guard([&] {
    RPLSH_REQUIRE(b, "index is NULL");
    if (num <= 0) return;
    RPLSH_REQUIRE(data, "data is NULL");
    bf_begin_change(b);
    RPLSH_REQUIRE(b->n + num < (1ll << 31), "index size limited to 2^31 rows");
    DeviceScope scope(b->cfg.device);
    hipStream_t st = bf_stream(b);
    const int64_t need = b->n + num;
    if (need > b->capacity) {
      // Grow: new buffers, copy the resident rows, re-encode everything.
      const int64_t cap = b->capacity == 0 ? need : std::max(need, b->capacity + b->capacity / 2);
      float *nx = nullptr, *nn = nullptr;
      RPLSH_HIP(hipMalloc(&nx, (size_t)cap * b->ldx * sizeof(float)));
      RPLSH_HIP(hipMalloc(&nn, (size_t)cap * sizeof(float)));
      if (b->n > 0) {
        RPLSH_HIP(hipMemcpyAsync(nx, b->x, (size_t)b->n * b->ldx * sizeof(float), hipMemcpyDeviceToDevice, st));
        RPLSH_HIP(hipMemcpyAsync(nn, b->xn, (size_t)b->n * sizeof(float), hipMemcpyDeviceToDevice, st));
      }
      RPLSH_HIP(hipStreamSynchronize(st));
      if (b->x) RPLSH_HIP(hipFree(b->x));
      if (b->xn) RPLSH_HIP(hipFree(b->xn));
      if (b->xenc) RPLSH_HIP(hipFree(b->xenc));
      if (b->xscale) RPLSH_HIP(hipFree(b->xscale));
      if (b->xnh) RPLSH_HIP(hipFree(b->xnh));
      b->xenc = nullptr;
      b->xscale = nullptr;
      b->xnh = nullptr;
      b->enc_precision = -1;
      b->x = nx;
      b->xn = nn;
      b->capacity = cap;
    }
    float* dst = b->x + b->n * b->ldx;
    if (b->ldx == b->dim) {
      RPLSH_HIP(hipMemcpyAsync(dst, data, (size_t)num * b->dim * sizeof(float), hipMemcpyDefault, st));
    } else {
      RPLSH_HIP(hipMemsetAsync(dst, 0, (size_t)num * b->ldx * sizeof(float), st));
      RPLSH_HIP(hipMemcpy2DAsync(dst, b->ldx * sizeof(float), data, b->dim * sizeof(float),
                                 b->dim * sizeof(float), num, hipMemcpyDefault, st));
    }
    const int64_t from = b->n;
    b->n = need;
    launch_l2norm(dst, num, b->dim, b->ldx, b->xn + from, 1, st);
    if (!b->xn_max_dev) RPLSH_HIP(hipMalloc(&b->xn_max_dev, sizeof(float)));
    RPLSH_HIP(hipMemcpyAsync(b->xn_max_dev, &b->xn_max, sizeof(float), hipMemcpyHostToDevice, st));
    launch_max_reduce(b->xn + from, num, b->xn_max_dev, st);
    bf_encode(b, b->enc_precision == b->cfg.precision ? from : 0, st);
    RPLSH_HIP(hipMemcpyAsync(&b->xn_max, b->xn_max_dev, sizeof(float), hipMemcpyDeviceToHost, st));
    RPLSH_HIP(hipStreamSynchronize(st));
  });
}

rplsh_status rplsh_bf_set_precision(rplsh_bf* b, int precision, int kprime_factor) {
  return guard([&] {
    RPLSH_REQUIRE(b, "index is NULL");
    (void)elem_of(precision);
    RPLSH_REQUIRE(filter_dpad(elem_of(precision), b->dim) > 0, "dimension not supported");
    bf_begin_change(b);
    DeviceScope scope(b->cfg.device);
    if (b->cfg.precision != precision) b->tunes.clear();
    b->cfg.precision = precision;
    if (kprime_factor > 0) b->cfg.kprime_factor = kprime_factor;
    hipStream_t st = bf_stream(b);
    bf_encode(b, 0, st);
    RPLSH_HIP(hipStreamSynchronize(st));
  });
}

rplsh_status rplsh_bf_set_persistent(rplsh_bf* b, int max_queries) {
  return guard([&] {
    RPLSH_REQUIRE(b, "index is NULL");
    RPLSH_REQUIRE(max_queries >= 0, "max_queries must be >= 0");
    b->cfg.persistent_max_queries = max_queries;
  });
}

rplsh_status rplsh_bf_search(rplsh_bf* b, int64_t nq, const float* queries, int k,
                             int32_t* indices, float* distances) {
  return guard([&] {
    RPLSH_REQUIRE(b, "index is NULL");
    if (nq <= 0) return;
    RPLSH_REQUIRE(queries && indices && distances, "NULL pointer");
    RPLSH_REQUIRE(k > 0, "k must be > 0 (got %d)", k);
    RPLSH_REQUIRE(nq < (1ll << 31) && (int64_t)nq * k < (1ll << 40), "batch too large");
    DeviceScope scope(b->cfg.device);
    hipStream_t st = bf_stream(b);
    const int dev = b->cfg.device;
    rplsh_search_stats stats{};
    stats.queries = nq;
    hipEvent_t e0, e1;
    RPLSH_HIP(hipEventCreate(&e0));
    RPLSH_HIP(hipEventCreate(&e1));
    RPLSH_HIP(hipEventRecord(e0, st));

    const float* qd = queries;
    if (pointer_device(queries) != dev) {
      float* t = b->qdev.as<float>((size_t)nq * b->dim);
      RPLSH_HIP(hipMemcpyAsync(t, queries, (size_t)nq * b->dim * sizeof(float), hipMemcpyDefault, st));
      qd = t;
    }
    const bool i_dev = pointer_device(indices) == dev;
    const bool d_dev = pointer_device(distances) == dev;
    int32_t* id = i_dev ? indices : b->idev.as<int32_t>((size_t)nq * k);
    float* dd = d_dev ? distances : b->ddev.as<float>((size_t)nq * k);

    if (b->n == 0) {
      // Empty index: FLT_MAX / -1 (search.cu:93-100).
      launch_fill_i32(id, nq * k, -1, st);
      launch_fill_f32(dd, nq * k, FLT_MAX, st);
    } else {
      if (b->enc_precision != b->cfg.precision) bf_encode(b, 0, st);
      const DbView v = bf_view(b);
      const SearchParams p = bf_search_params(b, k, nq);
      RPLSH_REQUIRE(std::min<int64_t>(p.kp, b->n) <= 4096, "k' = %d exceeds 4096", p.kp);
      if (b->cfg.persistent_max_queries > 0 && nq <= b->cfg.persistent_max_queries &&
          persist_supported(v, p)) {
        // Small batch: the persistent kernel (same results; every row of a
        // level scored, no carry between levels).
        if (!b->pst) b->pst = persist_state_create();
        PhaseEvents pe;
        pe.create();
        persist_run(*b->pst, {PersistBatch{&v, qd, (int)nq, &p, id, dd, &b->ws[0]}}, st, &pe);
        RPLSH_HIP(hipStreamSynchronize(st));
        finish_batch(v, qd, (int)nq, p, id, dd, b->ws, 0, st, stats, &pe,
                     static_cast<const BatchCounters*>(b->ws[0].hctr.p));
        stats.rows_filter_full = v.n;
        stats.persistent = 1;
      } else {
        search_batch(v, qd, (int)nq, p, id, dd, b->ws, 0, st, stats);
      }
    }
    if (!i_dev) RPLSH_HIP(hipMemcpyAsync(indices, id, (size_t)nq * k * sizeof(int32_t), hipMemcpyDefault, st));
    if (!d_dev) RPLSH_HIP(hipMemcpyAsync(distances, dd, (size_t)nq * k * sizeof(float), hipMemcpyDefault, st));
    RPLSH_HIP(hipEventRecord(e1, st));
    RPLSH_HIP(hipEventSynchronize(e1));
    RPLSH_HIP(hipEventElapsedTime(&stats.ms_total, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->last = stats;
  });
}

rplsh_status rplsh_bf_num_data(const rplsh_bf* b, int64_t* n) {
  return guard([&] {
    RPLSH_REQUIRE(b && n, "NULL argument");
    *n = b->n;
  });
}

rplsh_status rplsh_bf_dimension(const rplsh_bf* b, int* d) {
  return guard([&] {
    RPLSH_REQUIRE(b && d, "NULL argument");
    *d = b->dim;
  });
}

rplsh_status rplsh_bf_device(const rplsh_bf* b, int* device) {
  return guard([&] {
    RPLSH_REQUIRE(b && device, "NULL argument");
    *device = b->cfg.device;
  });
}

rplsh_status rplsh_bf_data_dev(const rplsh_bf* b, const float** data, const float** norms) {
  return guard([&] {
    RPLSH_REQUIRE(b, "index is NULL");
    if (data) *data = b->x;
    if (norms) *norms = b->xn;
  });
}

rplsh_status rplsh_bf_last_stats(const rplsh_bf* b, rplsh_search_stats* out) {
  return guard([&] {
    RPLSH_REQUIRE(b && out, "NULL argument");
    *out = b->last;
  });
}

// searchL2Distance on caller device buffers, plain FP32 path. Temporaries
// (norms when not given, rows padded to 16 bytes when x is not, and the
// search workspaces) are kept per device between calls and only grow; a
// call holds its device's set for its whole duration.
namespace {
struct KeptL2 {
  std::mutex mu;
  L2Scratch s;
};
KeptL2& kept_l2(int device) {
  static std::mutex mu;
  static std::vector<std::unique_ptr<KeptL2>> per_device;
  std::lock_guard<std::mutex> lk(mu);
  if ((int)per_device.size() <= device) per_device.resize(device + 1);
  if (!per_device[device]) {
    per_device[device] = std::make_unique<KeptL2>();
    per_device[device]->s.wss.resize(4);
  }
  return *per_device[device];
}
}  // namespace

rplsh_status rplsh_search_l2_dev(const float* x, int64_t n, int d, const float* xnorms,
                                 const float* q, int64_t nq, int k, int32_t* indices,
                                 float* distances, int exact, void* stream) {
  return guard([&] {
    KeptL2& kept = kept_l2(current_device());
    std::lock_guard<std::mutex> lk(kept.mu);
    search_l2_device(x, n, d, xnorms, q, nq, k, indices, distances, exact, (hipStream_t)stream, kept.s);
  });
}

}  // extern "C"

namespace rplsh {

void search_l2_device(const float* x, int64_t n, int d, const float* xnorms, const float* q, int64_t nq,
                      int k, int32_t* indices, float* distances, int exact, hipStream_t st,
                      L2Scratch& sc) {
  RPLSH_REQUIRE(d > 0 && filter_dpad(Elem::F32, d) > 0, "dimension %d unsupported (1..%d)", d, kMaxKLoopDim);
  RPLSH_REQUIRE(n >= 0 && nq >= 0, "bad sizes");
  RPLSH_REQUIRE(k > 0, "k must be > 0");
  RPLSH_REQUIRE(nq < (1ll << 31) && n < (1ll << 31), "sizes limited to 2^31 rows");
  if (nq == 0) return;
  RPLSH_REQUIRE(q && indices && distances && (n == 0 || x), "NULL pointer");
  if (n == 0) {
    launch_fill_i32(indices, nq * k, -1, st);
    launch_fill_f32(distances, nq * k, FLT_MAX, st);
    RPLSH_HIP(hipStreamSynchronize(st));
    return;
  }
  RPLSH_REQUIRE(plain_kprime(k) <= 4096 || n <= 4096, "k too large");
  DbView v;
  v.n = n;
  v.d = d;
  v.elem = Elem::F32;
  v.dpad = filter_dpad(Elem::F32, d);
  const bool aligned = (d % 4 == 0) && ((uintptr_t)x % 16 == 0);
  if (aligned) {
    v.x = x;
    v.ldx = d;
  } else {
    const int ld = (int)round_up(d, 4);
    float* t = sc.xpad.as<float>((size_t)n * ld);
    launch_pad_rows(x, n, d, d, ld, t, st);
    v.x = t;
    v.ldx = ld;
  }
  if (xnorms) {
    v.xnorm = xnorms;
  } else {
    float* t = sc.xn.as<float>(n);
    launch_l2norm(v.x, n, d, v.ldx, t, 1, st);
    v.xnorm = t;
  }
  float* mx = sc.xmax.as<float>(1);
  launch_fill_f32(mx, 1, 0.0f, st);
  launch_max_reduce(v.xnorm, n, mx, st);
  RPLSH_HIP(hipMemcpyAsync(&v.xn_max, mx, sizeof(float), hipMemcpyDeviceToHost, st));
  RPLSH_HIP(hipStreamSynchronize(st));
  v.xenc = v.x;
  v.ldenc = v.ldx;
  v.dvalid_bytes = (int)(v.ldx * 4);
  SearchParams p;
  p.k = k;
  p.kp = plain_kprime(k);
  p.exact = exact;
  p.certify = true;
  rplsh_search_stats stats{};
  search_batch(v, q, (int)nq, p, indices, distances, sc.wss, 0, st, stats);
  RPLSH_HIP(hipStreamSynchronize(st));
}

size_t search_l2_device_bytes(int64_t n, int d, bool have_norms, bool aligned, int64_t nq, int k) {
  if (n <= 0 || nq <= 0) return 0;
  DbView v;
  v.n = n;
  v.d = d;
  v.elem = Elem::F32;
  v.dpad = filter_dpad(Elem::F32, d);
  SearchParams p;
  p.k = k;
  p.kp = plain_kprime(k);
  SearchParams p2 = p;
  p2.c_factor = p.c_factor * 4;
  p2.kp = (int)std::min<int64_t>(std::min<int64_t>(n, 4096), (int64_t)p.kp * 4);
  const int64_t r = std::min<int64_t>(nq, 64);
  constexpr size_t kSlop = 32 * 256;  // alignment of the pieces
  size_t b = kSlop + 256;
  if (!aligned) b += (size_t)n * round_up(d, 4) * 4;
  if (!have_norms) b += (size_t)n * 4;
  b += search_workspace_bytes(v, (int)nq, p) + kSlop;
  b += search_workspace_bytes(v, (int)r, p2) + (size_t)r * ((size_t)d * 4 + (size_t)k * 8) + kSlop;
  return b;
}

}  // namespace rplsh

namespace {

// recall@k of got against truth (rows of k indices; -1 entries ignored).
double host_recall(const std::vector<int32_t>& got, const std::vector<int32_t>& truth, int64_t nq,
                   int k) {
  double tot = 0;
  int64_t rows = 0;
  std::vector<int32_t> a, t;
  for (int64_t i = 0; i < nq; ++i) {
    a.assign(got.begin() + i * k, got.begin() + (i + 1) * k);
    t.assign(truth.begin() + i * k, truth.begin() + (i + 1) * k);
    std::sort(a.begin(), a.end());
    std::sort(t.begin(), t.end());
    int valid = 0, hit = 0;
    for (int32_t v : t) valid += v >= 0;
    if (valid == 0) continue;
    size_t x = 0, y = 0;
    while (x < a.size() && y < t.size()) {
      if (a[x] < t[y]) ++x;
      else if (t[y] < a[x]) ++y;
      else {
        hit += a[x] >= 0;
        ++x;
        ++y;
      }
    }
    tot += (double)hit / valid;
    ++rows;
  }
  return rows ? tot / rows : 1.0;
}

}  // namespace

extern "C" {

rplsh_status rplsh_bf_autotune(rplsh_bf* b, const float* sample, int64_t nsample, int k,
                               int64_t batch_nq, float target_recall, rplsh_tune_result* out) {
  return guard([&] {
    RPLSH_REQUIRE(b && sample && out, "NULL argument");
    RPLSH_REQUIRE(nsample > 0 && k > 0 && batch_nq > 0, "bad sizes");
    RPLSH_REQUIRE(b->n > 0, "autotune needs a non-empty index");
    bf_begin_change(b);
    DeviceScope scope(b->cfg.device);
    bf_prepare(b);
    hipStream_t st = bf_stream(b);
    const int d = b->dim;
    const int ns = (int)std::min<int64_t>(nsample, batch_nq);
    DevBuf qs, qb, gi, gd;
    float* dq = qs.as<float>((size_t)ns * d);
    RPLSH_HIP(hipMemcpyAsync(dq, sample, (size_t)ns * d * sizeof(float), hipMemcpyDefault, st));
    std::vector<Workspace> wss(4);
    rplsh_search_stats stats{};
    const DbView v = bf_view(b);
    const bool plain = v.elem == Elem::F32;
    const int kmax_f = std::max(1, 4096 / k);

    rplsh_tune_result r{};
    r.k = k;
    r.batch_nq = batch_nq;
    r.recall = 1.0f;

    // 1. k' factor by recall against the plain path (exact in canonical FP32).
    if (!plain && target_recall <= 0.0f) {
      r.kprime_factor = b->cfg.kprime_factor;
      r.recall = -1.0f;  // not measured
    } else if (!plain) {
      DbView vf = v;
      vf.elem = Elem::F32;
      vf.dpad = filter_dpad(Elem::F32, d);
      vf.xenc = b->x;
      vf.ldenc = b->ldx;
      vf.dvalid_bytes = (int)(b->ldx * 4);
      vf.xscale = nullptr;
      vf.packed = false;
      SearchParams pf;
      pf.k = k;
      pf.kp = plain_kprime(k);
      pf.certify = true;
      int32_t* ti = gi.as<int32_t>((size_t)ns * k);
      float* td = gd.as<float>((size_t)ns * k);
      search_batch(vf, dq, ns, pf, ti, td, wss, 0, st, stats);
      std::vector<int32_t> truth((size_t)ns * k), got((size_t)ns * k);
      RPLSH_HIP(hipMemcpy(truth.data(), ti, truth.size() * 4, hipMemcpyDeviceToHost));
      DevBuf ri, rd;
      int32_t* oi = ri.as<int32_t>((size_t)ns * k);
      float* od = rd.as<float>((size_t)ns * k);
      const int factors[] = {1, 2, 3, 4, 6, 8, 12, 16};
      int chosen = 0;
      float chosen_recall = 0.0f;
      for (int f : factors) {
        if (f > kmax_f) break;
        SearchParams p = bf_search_params(b, k, 0);
        p.kp = f * k;
        search_batch(v, dq, ns, p, oi, od, wss, 0, st, stats);
        RPLSH_HIP(hipMemcpy(got.data(), oi, got.size() * 4, hipMemcpyDeviceToHost));
        const double rec = host_recall(got, truth, ns, k);
        chosen = f;
        chosen_recall = (float)rec;
        if (rec >= target_recall) break;
      }
      r.kprime_factor = chosen;
      r.recall = chosen_recall;
    }

    // 2. Grid decomposition and survivor target by measured batch time.
    float* dqb = qb.as<float>((size_t)batch_nq * d);
    for (int64_t off = 0; off < batch_nq; off += ns) {
      const int64_t cnt = std::min<int64_t>(ns, batch_nq - off);
      RPLSH_HIP(hipMemcpyAsync(dqb + off * d, dq, (size_t)cnt * d * sizeof(float),
                               hipMemcpyDeviceToDevice, st));
    }
    int32_t* oi = gi.as<int32_t>((size_t)batch_nq * k);
    float* od = gd.as<float>((size_t)batch_nq * k);
    hipEvent_t e0, e1;
    RPLSH_HIP(hipEventCreate(&e0));
    RPLSH_HIP(hipEventCreate(&e1));
    auto time_it = [&](const SearchParams& p) {
      stats = rplsh_search_stats{};
      search_batch(v, dqb, (int)batch_nq, p, oi, od, wss, 0, st, stats);  // warm-up
      float best = 1e30f;
      for (int rep = 0; rep < 2; ++rep) {
        RPLSH_HIP(hipEventRecord(e0, st));
        search_batch(v, dqb, (int)batch_nq, p, oi, od, wss, 0, st, stats);
        RPLSH_HIP(hipEventRecord(e1, st));
        RPLSH_HIP(hipEventSynchronize(e1));
        float ms = 0;
        RPLSH_HIP(hipEventElapsedTime(&ms, e0, e1));
        best = std::min(best, ms);
      }
      return best;
    };
    SearchParams base = bf_search_params(b, k, 0);
    if (!plain) base.kp = k * r.kprime_factor;
    SearchParams def = base;
    def.target_wgs = 4096;
    def.c_factor = 3;
    def.rho = 0;
    def.variant = kFilterLdsWide;
    if (!plain) def.kp = k * b->cfg.kprime_factor;
    r.ms_default = time_it(def);
    r.ms_tuned = 1e30f;
    // (survivor multiple c, rank rho of the last threshold): the last sample
    // level has stride c*k'/rho, so a larger rho buys a tighter c at the same
    // risk P(Gamma(rho) < rho/c) of a query keeping fewer than k' rows.
    // Time of re-running a few failed queries (wider margins, one more pass).
    auto retry_ms = [&](const SearchParams& p) -> double {
      SearchParams p2 = p;
      p2.c_factor = p.c_factor * 4;
      p2.kp = (int)std::min<int64_t>(std::min<int64_t>(v.n, 4096), (int64_t)p.kp * 4);
      const int nr = (int)std::min<int64_t>(ns, 8);
      search_batch(v, dq, nr, p2, oi, od, wss, 0, st, stats);  // warm-up
      RPLSH_HIP(hipEventRecord(e0, st));
      search_batch(v, dq, nr, p2, oi, od, wss, 0, st, stats);
      RPLSH_HIP(hipEventRecord(e1, st));
      RPLSH_HIP(hipEventSynchronize(e1));
      float ms = 0;
      RPLSH_HIP(hipEventElapsedTime(&ms, e0, e1));
      return ms;
    };
    const int wgs[] = {4096, 8192, 16384};
    // With a carried last sample level (LDS variants) a denser last sample
    // costs little (the full level skips its tiles), so tight c pays off.
    const int cr[][2] = {{3, 0}, {4, 16}, {2, 32}, {3, 32}, {2, 48}, {2, 64}, {8, 8}};
    // Packed copies: LDS wide / narrow, ring wide (the direct and ring-narrow
    // variants are left to explicit tuning entries).
    const int vars_packed[] = {kFilterLdsWide, kFilterLdsNarrow, kFilterRingWide};
    const int vars_plain[] = {kFilterLdsWide};
    const int* vars = v.packed ? vars_packed : vars_plain;
    const int nvar = v.packed ? 3 : 1;
    struct Cand {
      SearchParams p;
      float ms;       // sweep time (best of two runs)
      float penalty;  // expected retry cost
    };
    std::vector<Cand> cands;
    for (int vi = 0; vi < nvar; ++vi)
      for (int tw : wgs)
        for (const auto& c : cr) {
          SearchParams p = base;
          p.target_wgs = tw;
          p.c_factor = c[0];
          p.rho = c[1];
          p.variant = vars[vi];
          const float ms = time_it(p);
          // The timing batch repeats the sample, so a query that keeps too
          // few rows fails in all its copies or in none: add the expected
          // cost of re-running the failed queries of a batch of distinct
          // queries (one more pass over the database), unless the timed runs
          // already paid for one.
          float penalty = 0.0f;
          if (stats.retried_queries == 0 && plan_levels(v.n, p.kp, p.c_factor, p.rho).size() > 1) {
            const int rho = final_rank(p.kp, p.c_factor, p.rho);
            const double pf = short_probability(rho, (double)survivor_target(p.kp, p.c_factor) / p.kp);
            const double p_any = 1.0 - std::pow(1.0 - pf, (double)batch_nq);
            if (p_any > 1e-3) penalty = (float)(p_any * retry_ms(p));
          }
          cands.push_back({p, ms, penalty});
        }
    // Configurations within a few percent of each other are not told apart by
    // one best-of-two timing (the clock drifts with power and temperature over
    // the sweep): re-time the fastest few in interleaved rounds and keep the
    // one with the lowest total.
    std::sort(cands.begin(), cands.end(),
              [](const Cand& x, const Cand& y) { return x.ms + x.penalty < y.ms + y.penalty; });
    const size_t nfinal = std::min<size_t>(cands.size(), 4);
    std::vector<float> total(nfinal, 0.0f);
    const int kRounds = 5;
    for (int round = 0; round < kRounds; ++round)
      for (size_t i = 0; i < nfinal; ++i) total[i] += time_it(cands[i].p);
    size_t pick = 0;
    for (size_t i = 1; i < nfinal; ++i)
      if (total[i] + kRounds * cands[i].penalty < total[pick] + kRounds * cands[pick].penalty) pick = i;
    r.ms_tuned = total[pick] / kRounds + cands[pick].penalty;
    r.target_wgs = cands[pick].p.target_wgs;
    r.c_factor = cands[pick].p.c_factor;
    r.rho = cands[pick].p.rho;
    r.variant = cands[pick].p.variant;
    // 3. The sample levels with the other LDS kernel (same survivors): their
    // loose thresholds make them emission-heavy, which the narrow kernel's
    // third wave per SIMD hides better for FP8 (10M x 96, 10k queries: sample
    // levels 1.41 -> 1.22 ms with the wide kernel on the full level).
    if (v.packed && v.nrcopies == 0 && (r.variant == kFilterLdsWide || r.variant == kFilterLdsNarrow)) {
      SearchParams p = base;
      p.target_wgs = r.target_wgs;
      p.c_factor = r.c_factor;
      p.rho = r.rho;
      p.variant = r.variant;
      SearchParams pm = p;
      pm.sample_variant = r.variant == kFilterLdsWide ? kFilterLdsNarrow : kFilterLdsWide;
      float ms_same = 0.0f, ms_mixed = 0.0f;
      for (int round = 0; round < kRounds; ++round) {
        ms_same += time_it(p);
        ms_mixed += time_it(pm);
      }
      if (ms_mixed < ms_same) {
        r.ms_tuned -= (ms_same - ms_mixed) / kRounds;
        r.variant = encode_variants(pm.variant, pm.sample_variant);
      }
    }
    // 4. The full level in L2-sized slices: the same arithmetic over shorter
    // slices, whose re-reads by the other query blocks stay in L2 (10M x 96
    // BF16: 4.33 -> 2.86 GB over HBM per launch). Kept unless slower by more
    // than the timing noise (0.5 %): at equal speed fewer HBM bytes leave
    // bandwidth and power to whatever else runs on the chip.
    if (v.packed && v.nrcopies == 0) {
      SearchParams p = base;
      p.target_wgs = r.target_wgs;
      p.c_factor = r.c_factor;
      p.rho = r.rho;
      p.variant = full_variant_of(r.variant);
      p.sample_variant = sample_variant_of(r.variant);
      SearchParams pl = p;
      pl.l2_slices = true;
      float ms_cur = 0.0f, ms_l2 = 0.0f;
      for (int round = 0; round < kRounds; ++round) {
        ms_cur += time_it(p);
        ms_l2 += time_it(pl);
      }
      if (ms_l2 <= 1.005f * ms_cur) {
        r.ms_tuned += (ms_l2 - ms_cur) / kRounds;
        r.variant = encode_variants(pl.variant, pl.sample_variant, true);
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->tunes.erase(std::remove_if(b->tunes.begin(), b->tunes.end(),
                                  [&](const rplsh_tune_result& t) {
                                    return t.k == k && t.batch_nq == batch_nq;
                                  }),
                   b->tunes.end());
    b->tunes.push_back(r);
    *out = r;
  });
}

rplsh_status rplsh_bf_set_tuning(rplsh_bf* b, const rplsh_tune_result* t) {
  return guard([&] {
    RPLSH_REQUIRE(b && t, "NULL argument");
    RPLSH_REQUIRE(t->k > 0 && t->batch_nq > 0, "bad tuning entry");
    RPLSH_REQUIRE(t->kprime_factor >= 0 && (int64_t)t->kprime_factor * t->k <= 4096, "k' too large");
    RPLSH_REQUIRE(valid_variants(t->variant), "unknown filter variant %d", t->variant);
    RPLSH_REQUIRE(t->rho >= 0 && t->rho <= 256, "bad rho %d", t->rho);
    bf_begin_change(b);
    b->tunes.erase(std::remove_if(b->tunes.begin(), b->tunes.end(),
                                  [&](const rplsh_tune_result& x) {
                                    return x.k == t->k && x.batch_nq == t->batch_nq;
                                  }),
                   b->tunes.end());
    b->tunes.push_back(*t);
  });
}

rplsh_status rplsh_bf_clear_tuning(rplsh_bf* b) {
  return guard([&] {
    RPLSH_REQUIRE(b, "index is NULL");
    bf_begin_change(b);
    b->tunes.clear();
  });
}

}  // extern "C"

// ---- index files -------------------------------------------------------------
// Layout (little endian): "RPLSHBF1", u32 version (1), i32 dim, i64 n,
// rplsh_bf_config, i32 tuning entries, entries (rplsh_tune_result), then the
// n FP32 rows unpadded [n][dim]. Norms and candidate copies are rebuilt on
// load (they are functions of the rows).
namespace {
constexpr char kIndexMagic[8] = {'R', 'P', 'L', 'S', 'H', 'B', 'F', '1'};
constexpr uint32_t kIndexVersion = 3;  // 3: configuration with residual_clusters
}  // namespace

rplsh_status rplsh_bf_save(rplsh_bf* b, const char* path) {
  return guard([&] {
    RPLSH_REQUIRE(b && path, "NULL argument");
    DeviceScope scope(b->cfg.device);
    hipStream_t st = bf_stream(b);
    RPLSH_HIP(hipStreamSynchronize(st));
    FILE* fp = std::fopen(path, "wb");
    if (!fp) fail(RPLSH_ERR_INVALID_ARG, "%s: cannot open for writing", path);
    bool ok = true;
    auto put = [&](const void* p, size_t bytes) { ok = ok && std::fwrite(p, 1, bytes, fp) == bytes; };
    const int32_t dim = b->dim, ntune = (int32_t)b->tunes.size();
    const int64_t n = b->n;
    put(kIndexMagic, 8);
    put(&kIndexVersion, 4);
    put(&dim, 4);
    put(&n, 8);
    put(&b->cfg, sizeof(b->cfg));
    put(&ntune, 4);
    for (const auto& t : b->tunes) put(&t, sizeof(t));
    // Rows, through a host buffer of at most ~64 MB.
    const int64_t per = std::max<int64_t>(1, (64ll << 20) / ((int64_t)dim * 4));
    std::vector<float> buf((size_t)std::min<int64_t>(per, std::max<int64_t>(n, 1)) * dim);
    try {
      for (int64_t r = 0; r < n && ok; r += per) {
        const int64_t c = std::min(per, n - r);
        RPLSH_HIP(hipMemcpy2D(buf.data(), (size_t)dim * 4, b->x + r * b->ldx, (size_t)b->ldx * 4,
                              (size_t)dim * 4, (size_t)c, hipMemcpyDeviceToHost));
        put(buf.data(), (size_t)c * dim * 4);
      }
    } catch (...) {
      std::fclose(fp);
      throw;
    }
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok) fail(RPLSH_ERR_INVALID_ARG, "%s: write failed", path);
  });
}

rplsh_status rplsh_bf_load(rplsh_resources* res, const char* path, int device, rplsh_bf** out) {
  return guard([&] {
    RPLSH_REQUIRE(res && path && out, "NULL argument");
    FILE* fp = std::fopen(path, "rb");
    if (!fp) fail(RPLSH_ERR_INVALID_ARG, "%s: cannot open", path);
    std::vector<char> head(8);
    uint32_t version = 0;
    int32_t dim = 0, ntune = 0;
    int64_t n = 0;
    rplsh_bf_config cfg{};
    bool ok = std::fread(head.data(), 1, 8, fp) == 8 && std::memcmp(head.data(), kIndexMagic, 8) == 0 &&
              std::fread(&version, 4, 1, fp) == 1 && version == kIndexVersion &&
              std::fread(&dim, 4, 1, fp) == 1 && std::fread(&n, 8, 1, fp) == 1 &&
              std::fread(&cfg, sizeof(cfg), 1, fp) == 1 && std::fread(&ntune, 4, 1, fp) == 1 && dim > 0 &&
              n >= 0 && ntune >= 0 && ntune < (1 << 16);
    std::vector<rplsh_tune_result> tunes((size_t)std::max(ntune, 0));
    if (ok && ntune > 0) ok = std::fread(tunes.data(), sizeof(rplsh_tune_result), (size_t)ntune, fp) == (size_t)ntune;
    std::vector<float> rows;
    if (ok && n > 0) {
      rows.resize((size_t)n * dim);
      ok = std::fread(rows.data(), 4, rows.size(), fp) == rows.size();
    }
    std::fclose(fp);
    if (!ok) fail(RPLSH_ERR_INVALID_ARG, "%s: not an index file of this library (version %u) or truncated", path,
                  kIndexVersion);
    if (device >= 0) cfg.device = device;
    rplsh_bf* b = nullptr;
    if (rplsh_bf_create(res, dim, &cfg, &b) != RPLSH_OK) fail(RPLSH_ERR_STATE, "%s", rplsh_last_error());
    if (n > 0 && rplsh_bf_add(b, n, rows.data()) != RPLSH_OK) {
      const std::string msg = rplsh_last_error();
      rplsh_bf_destroy(b);
      fail(RPLSH_ERR_STATE, "%s", msg.c_str());
    }
    b->tunes = tunes;
    *out = b;
  });
}
