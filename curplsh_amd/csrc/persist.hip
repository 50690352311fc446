// Persistent serving kernel: whole searches -- query preparation, every
// filter level with its select, and the exact re-rank -- for several batches
// at once, against indexes of different sizes (one kernel per element kind
// and padded dimension).
//
// The host writes a program: per batch, its phases in order (prep, filter and
// select of each level, re-rank), each a run of work units with the argument
// record of that phase. Step s of a group is the s-th phase of every batch
// that has one; the steps run as consecutive launches of the kernel on one
// stream, so a step sees everything the previous one wrote and no workgroup
// ever waits for another. Within a launch the resident workgroups take the
// step's units from a device counter until none is left (batches of uneven
// sizes share the workgroups), with a grid of at most what the device holds at
// once.
//
// Units per phase: prep and select take four queries (a wave each); a filter
// unit takes 32 queries x rows_per_slice rows, its four waves taking the
// slice's 32-row tiles in turn, with the query operands in registers and the
// row operands loaded straight from the packed copy (the tile layout is the
// MFMA operand order, so every load is 16 or 32 contiguous bytes per lane);
// a re-rank unit takes one query. The per-tile arithmetic, the selects and the
// re-rank are the same device code and instruction sequence as the launched
// pipeline's (packed_mma.hpp, select_body.hpp, encode.hpp), so results are
// bit-identical to rplsh_bf_search.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstdint>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "encode.hpp"
#include "engine.hpp"
#include "kernels.hpp"
#include "packed_mma.hpp"
#include "select_body.hpp"

namespace rplsh {

namespace {

constexpr int kPBlock = 256;

// Bytes of dynamic LDS a launch carries: the larger of the selects' (four
// waves x (256-bin histogram + kWaveSelK words)) and the re-rank's region at
// the longest list, with the FP64 band and the staged rows of d = 256.
constexpr size_t kSelectLds = (size_t)kWaveSelQ * (256 * 4 + kWaveSelK * 8);
inline size_t persist_lds_bytes() {
  const size_t rr = rerank_lds_bytes(kPersistMaxKp, true, 256, true);
  return rr > kSelectLds ? rr : kSelectLds;
}

// ---- phase bodies --------------------------------------------------------

__device__ __forceinline__ void prep_unit(const PrepArgs& a, int local) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int q = local * 4 + w;
  if (local == 0 && threadIdx.x < 4) a.ctr[threadIdx.x] = 0;
  if (q >= a.nq) return;
  const float* row = a.q + (int64_t)q * a.d;
  if (lane == 0) {
    // Canonical norm: one fmaf chain over j (as l2norm_kernel).
    float s = 0.0f;
    for (int j = 0; j < a.d; ++j) s = fmaf(row[j], row[j], s);
    a.qnorm[q] = s;
    a.tau[q] = INFINITY;
    a.status[q] = 0;
    a.counts[q] = a.count_init;
  }
  if (!a.fp8) {
    uint16_t* out = static_cast<uint16_t*>(a.qenc) + (int64_t)q * a.dpad;
    for (int j = lane; j < a.dpad; j += 64) out[j] = j < a.d ? bf16_rne(row[j]) : (uint16_t)0;
  } else {
    float mx = 0.0f;
    for (int j = lane; j < a.d; j += 64) mx = fmaxf(mx, fabsf(row[j]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    const float s = e4m3_row_scale(mx);
    uint8_t* out = static_cast<uint8_t*>(a.qenc) + (int64_t)q * a.dpad;
    for (int j = lane; j < a.dpad; j += 64) out[j] = j < a.d ? e4m3_rne_sat(row[j] / s) : 0;
    if (lane == 0) a.qscale[q] = s;
  }
}

// Pointers read from the program are generic to the compiler; the filter's
// loads go through global (address space 1) pointers, so they are counted on
// vmcnt alone and can stay in flight together.
template <class T>
__device__ __forceinline__ T gload(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}

// WIDE: several tiles per wave and iteration, about 16 KB of row operands in
// flight (the filter-only kernel); otherwise one tile (the general kernel,
// whose other bodies need the registers).
template <Elem E, int DPAD, bool WIDE>
__device__ __forceinline__ void filter_unit(const FilterArgs& args, int local) {
  // A private copy: the candidate stores below cannot alias it, so its fields
  // stay in registers instead of being re-read after every store.
  const FilterArgs a = args;
  using G = Geometry<E, DPAD>;
  constexpr int LB = G::LB;
  constexpr int NC = G::NC;
  constexpr int TPW = !WIDE || G::TILE >= 12288 ? 1 : (G::TILE >= 6144 ? 2 : 4);
  const int qtile = local % a.num_qtiles;
  const int slice = local / a.num_qtiles;
  const int64_t r_begin = (int64_t)slice * a.rows_per_slice;
  const int64_t r_end = min(r_begin + a.rows_per_slice, a.m);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int h = lane >> 5, col = lane & 31;
  const int qr = qtile * 32 + col;
  const int nq = a.nq;

  Frag<LB> qf[NC];
  float thr = INFINITY;
  int qsc = 127;
  if (qr < nq) {
    const char* qp = reinterpret_cast<const char*>(a.q) + (int64_t)qr * G::ROWB;
#pragma unroll
    for (int c = 0; c < NC; ++c) qf[c] = load_frag_global<LB>(qp + 2 * LB * c + LB * h);
    thr = fmaxf(-0.5f * gload(a.tau + qr), -FLT_MAX);
    if (E == Elem::FP8) qsc = e8m0_of(gload(a.qscale + qr));
  } else {
#pragma unroll
    for (int c = 0; c < NC; ++c) qf[c] = zero_frag<LB>();
  }
  const char* xb = reinterpret_cast<const char*>(a.x);
  const int64_t stride = a.stride;
  int* cnt = a.counts + qr;
  uint64_t* dst = a.cand + (int64_t)qr * a.cap;

#pragma unroll 1
  for (int64_t t0 = r_begin + 32 * TPW * wave; t0 < r_end; t0 += 128 * TPW) {
    // Operand A of tiles t0 + 32u: this lane's row (the same 32 rows for both
    // halves of the wave, each half its part of every K-chunk). Lanes past the
    // end load the last row instead: their output rows carry C = -inf and are
    // never emitted, and unconditional loads stay in flight together.
    Frag<LB> af[TPW][NC];
    int xsc[TPW];
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
      int64_t j = t0 + 32 * u + col;
      if (j >= r_end) j = r_end - 1;
      const int64_t p = j * stride;
      const char* ap = xb + (p >> 5) * G::TILE + ((p & 31) + 32 * h) * LB;
#pragma unroll
      for (int c = 0; c < NC; ++c) af[u][c] = load_frag_global<LB>(ap + c * G::CHUNK);
      xsc[u] = E == Elem::FP8 ? e8m0_of(gload(a.xscale + p)) : 127;
    }
    // C = -||x||^2 / 2 of rows t0 + 32u + 8g + 4h + i (-inf past the end),
    // loaded while the row operands are in flight.
    f32x16 ci[TPW];
    if (stride == 1 && t0 + 32 * TPW <= r_end) {
      // Whole tiles at stride 1: aligned 16-byte loads (t0 is a multiple of 32).
#pragma unroll
      for (int u = 0; u < TPW; ++u)
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          const f32x4 v = gload(reinterpret_cast<const f32x4*>(a.xnorm + t0 + 32 * u + 8 * g + 4 * h));
#pragma unroll
          for (int i = 0; i < 4; ++i) ci[u][4 * g + i] = -0.5f * v[i];
        }
    } else {
#pragma unroll
      for (int u = 0; u < TPW; ++u)
#pragma unroll
        for (int g = 0; g < 4; ++g)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int64_t r = t0 + 32 * u + 8 * g + 4 * h + i;
            const float v = gload(a.xnorm + (r < r_end ? r : r_end - 1) * stride);
            ci[u][4 * g + i] = r < r_end ? -0.5f * v : -INFINITY;
          }
    }
    // Every load of the tiles is issued before the first product.
    if constexpr (WIDE) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
      const int64_t tu = t0 + 32 * u;
      f32x16 acc;
#pragma unroll
      for (int c = 0; c < NC; ++c) acc = mfma_chunk<E>(af[u][c], qf[c], c == 0 ? ci[u] : acc, xsc[u], qsc);
      if (a.dense) {
        // Every row survives (tau = +inf): row r's word goes to slot r.
        if (qr < nq) {
#pragma unroll
          for (int i = 0; i < 16; ++i) {
            const int64_t r = tu + 8 * (i >> 2) + 4 * h + (i & 3);
            if (r < r_end) dst[r] = pack_cand(-2.0f * acc[i], (int32_t)(r * stride));
          }
        }
        continue;
      }
      float m = acc[0];
#pragma unroll
      for (int i = 1; i < 16; ++i) m = fmaxf(m, acc[i]);
      if (__any(m >= thr)) {
        if (m >= thr) {
#pragma unroll
          for (int i = 0; i < 16; ++i) {
            const float v = acc[i];
            if (v >= thr) {
              const int64_t r = tu + 8 * (i >> 2) + 4 * h + (i & 3);
              const int pos = atomicAdd(cnt, 1);
              if (pos < a.cap) dst[pos] = pack_cand(-2.0f * v, (int32_t)(r * stride));
            }
          }
        }
      }
    }
  }
}

__device__ __forceinline__ void select_unit(const SelectArgs& a, int local, int count_reset, char* lds) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int q = local * 4 + w;
  if (q >= a.nq) return;
  int* hist = reinterpret_cast<int*>(lds) + w * 256;
  uint64_t* out = reinterpret_cast<uint64_t*>(lds + kWaveSelQ * 256 * 4) + w * kWaveSelK;
  select_wave_query(a, q, hist, out);
  // The wave has read this query's count and list; the next level appends
  // to it from count_reset.
  if (count_reset >= 0 && lane == 0) const_cast<int*>(a.counts)[q] = count_reset;
}

template <int NP>
__device__ __forceinline__ void rerank_unit(const RerankTailArgs& t, int q, uint64_t* lds, float* sq, float* sxn, int* sh) {
  rerank_query<NP>(t.ra, q, lds, sq, sxn, sh);
  __syncthreads();
  if (threadIdx.x == 0) {
    // collect_failed and sum_counts of the launched pipeline.
    if (__hip_atomic_load(t.ra.status + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
      const int pos = atomicAdd(t.failed, 1);
      t.flist[pos] = q;
    }
    atomicAdd(reinterpret_cast<unsigned long long*>(t.survivors), (unsigned long long)t.counts[q]);
  }
}

// ---- the kernel ----------------------------------------------------------

// Any phase kind's unit (the general step kernel).
template <Elem E, int DPAD>
__device__ __forceinline__ void run_unit(const PersistPhase& P, const char* arg, int local, uint64_t* plds, float* sq_fixed,
                         float* sxn, int* sh) {
  {
    switch (P.kind) {
      case kPkPrep:
        prep_unit(*reinterpret_cast<const PrepArgs*>(arg), local);
        break;
      case kPkFilter:
        filter_unit<E, DPAD, false>(*reinterpret_cast<const FilterArgs*>(arg), local);
        break;
      case kPkSelect:
        select_unit(*reinterpret_cast<const SelectArgs*>(arg), local, P.count_reset,
                    reinterpret_cast<char*>(plds));
        break;
      default: {
        const RerankTailArgs& t = *reinterpret_cast<const RerankTailArgs*>(arg);
        if (rerank_staged(t.ra.x, t.ra.ldx, t.ra.d)) {
          if (t.ra.d <= 128) rerank_unit<1>(t, local, plds, sq_fixed, sxn, sh);
          else rerank_unit<2>(t, local, plds, sq_fixed, sxn, sh);
        } else {
          rerank_unit<0>(t, local, plds, sq_fixed, sxn, sh);
        }
        break;
      }
    }
  }
}

// One step: units [unit_begin, unit_end) of phases [phase_begin, phase_end),
// taken from *claim (zeroed before the launch). FILTER: a step of filter
// phases only, compiled without the other bodies (fewer registers, so more
// waves per SIMD hide the row loads); otherwise every kind.
template <Elem E, int DPAD, bool FILTER>
__global__ __launch_bounds__(kPBlock, 2) void persistent_step_kernel(const char* __restrict__ prog,
                                                                     int phase_begin, int phase_end,
                                                                     int unit_begin, int unit_end,
                                                                     int* __restrict__ claim) {
  extern __shared__ __attribute__((aligned(16))) uint64_t plds[];
  __shared__ float sq_fixed[256];
  __shared__ float sxn[kSelWaves * 64];
  __shared__ int sh[4];
  __shared__ int s_unit;
  const PersistPhase* ph = reinterpret_cast<const PersistPhase*>(prog + sizeof(PersistHeader));
  for (;;) {
    if (threadIdx.x == 0) s_unit = unit_begin + atomicAdd(claim, 1);
    __syncthreads();
    const int u = __builtin_amdgcn_readfirstlane(s_unit);
    if (u >= unit_end) break;
    // The phase holding unit u (phases are sorted by their first unit).
    int lo = phase_begin, hi = phase_end - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (ph[mid].begin <= u) lo = mid;
      else hi = mid - 1;
    }
    lo = __builtin_amdgcn_readfirstlane(lo);
    const PersistPhase P = ph[lo];
    const int local = u - P.begin;
    const char* arg = prog + P.arg;
    if constexpr (FILTER) {
      filter_unit<E, DPAD, true>(*reinterpret_cast<const FilterArgs*>(arg), local);
    } else {
      run_unit<E, DPAD>(P, arg, local, plds, sq_fixed, sxn, sh);
    }
    __syncthreads();  // s_unit is rewritten by the next claim
  }
}

typedef void (*PersistKernel)(const char*, int, int, int, int, int*);

template <Elem E, int DPAD>
PersistKernel kernel_for(bool filter) {
  return filter ? persistent_step_kernel<E, DPAD, true> : persistent_step_kernel<E, DPAD, false>;
}

PersistKernel persist_kernel(Elem e, int dpad, bool filter) {
  if (e == Elem::BF16) {
    switch (dpad) {
      case 32: return kernel_for<Elem::BF16, 32>(filter);
      case 64: return kernel_for<Elem::BF16, 64>(filter);
      case 96: return kernel_for<Elem::BF16, 96>(filter);
      case 128: return kernel_for<Elem::BF16, 128>(filter);
      case 160: return kernel_for<Elem::BF16, 160>(filter);
      case 192: return kernel_for<Elem::BF16, 192>(filter);
      case 256: return kernel_for<Elem::BF16, 256>(filter);
      default: return nullptr;
    }
  }
  if (e == Elem::FP8) {
    switch (dpad) {
      case 64: return kernel_for<Elem::FP8, 64>(filter);
      case 128: return kernel_for<Elem::FP8, 128>(filter);
      case 256: return kernel_for<Elem::FP8, 256>(filter);
      default: return nullptr;
    }
  }
  return nullptr;
}

}  // namespace

struct PersistState {
  DevBuf prog, sync;   // program; one unit counter per step
  HostBuf hprog;       // program staging
  int resident = 0;    // workgroups the device holds at once: general kernel
  int resident_f = 0;  // filter-only kernel
};

PersistState* persist_state_create() { return new PersistState(); }
void persist_state_destroy(PersistState* s) { delete s; }

bool persist_supported(const DbView& db, const SearchParams& p) {
  // The cluster-residual copy's full level runs on the launched pipeline.
  return db.packed && db.nrcopies == 0 && db.n > 0 && db.elem != Elem::F32 && db.d <= 256 &&
         persist_kernel(db.elem, db.dpad, false) != nullptr && std::min<int64_t>(p.kp, db.n) <= kPersistMaxKp;
}

void persist_run(PersistState& S, const std::vector<PersistBatch>& batches, hipStream_t st,
                 PhaseEvents* pe) {
  RPLSH_REQUIRE(!batches.empty(), "persistent kernel: no batches");
  const Elem e = batches[0].db->elem;
  const int dpad = batches[0].db->dpad;
  PersistKernel kern = persist_kernel(e, dpad, false);
  PersistKernel kern_f = persist_kernel(e, dpad, true);
  RPLSH_REQUIRE(kern && kern_f, "persistent kernel: element kind / dimension not instantiated");
  const size_t lds = persist_lds_bytes();
  if (S.resident == 0) {
    int dev = 0, cus = 0, per_cu = 0, per_cu_f = 0;
    RPLSH_HIP(hipGetDevice(&dev));
    RPLSH_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    RPLSH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kern),
                                                           kPBlock, lds));
    RPLSH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_f, reinterpret_cast<const void*>(kern_f),
                                                           kPBlock, 0));
    S.resident = std::max(1, cus * std::max(1, per_cu));
    S.resident_f = std::max(1, cus * std::max(1, per_cu_f));
  }
  const int64_t target_units = S.resident_f;

  // Plans, then the program: phases ordered by position (all batches' preps,
  // then their first filters, ...); position s is step s.
  std::vector<PersistBatchPlan> plans(batches.size());
  size_t max_phases = 0;
  for (size_t b = 0; b < batches.size(); ++b) {
    const PersistBatch& pb = batches[b];
    RPLSH_REQUIRE(pb.db->elem == e && pb.db->dpad == dpad, "persistent kernel: mixed element kinds / dimensions");
    plan_persistent_batch(*pb.db, pb.q, pb.nq, *pb.p, pb.out_i, pb.out_d, *pb.ws, target_units, plans[b]);
    max_phases = std::max(max_phases, 2 + 2 * plans[b].filter.size());
  }
  struct Rec {
    int kind, units, count_reset;
    const void* arg;
    size_t bytes;
  };
  std::vector<std::vector<Rec>> per(batches.size());
  for (size_t b = 0; b < batches.size(); ++b) {
    const PersistBatchPlan& pl = plans[b];
    const int nq = batches[b].nq;
    const int qunits = (int)div_up(nq, 4);
    per[b].push_back({kPkPrep, qunits, -1, &pl.prep, sizeof(PrepArgs)});
    for (size_t l = 0; l < pl.filter.size(); ++l) {
      const FilterArgs& fa = pl.filter[l];
      per[b].push_back({kPkFilter, fa.num_qtiles * fa.num_slices, -1, &fa, sizeof(FilterArgs)});
      per[b].push_back({kPkSelect, qunits, pl.count_reset[l], &pl.select[l], sizeof(SelectArgs)});
    }
    per[b].push_back({kPkRerank, nq, -1, &pl.rerank, sizeof(RerankTailArgs)});
  }
  // Batches are aligned at their last phase (the re-rank): with fewer levels
  // a batch starts later, so every step holds one kind of phase or the
  // prep / select mix (filters at odd steps, re-ranks at the last).
  std::vector<PersistPhase> phases;
  std::vector<std::pair<const void*, size_t>> args;
  std::vector<int> step_phase{0}, step_unit{0};  // boundaries of the steps
  std::vector<char> step_filter;                 // step of filter phases only
  int units = 0;
  for (size_t pos = 0; pos < max_phases; ++pos) {
    bool only_filter = true;
    for (size_t b = 0; b < batches.size(); ++b) {
      const size_t shift = max_phases - per[b].size();
      if (pos < shift) continue;
      const Rec& r = per[b][pos - shift];
      only_filter = only_filter && r.kind == kPkFilter;
      PersistPhase P{};
      P.kind = r.kind;
      P.begin = units;
      P.units = r.units;
      P.count_reset = r.count_reset;
      P.arg = (int)args.size();  // index for now, byte offset below
      args.push_back({r.arg, r.bytes});
      units += r.units;
      phases.push_back(P);
    }
    step_phase.push_back((int)phases.size());
    step_unit.push_back(units);
    step_filter.push_back(only_filter ? 1 : 0);
  }
  size_t off = round_up(sizeof(PersistHeader) + phases.size() * sizeof(PersistPhase), 16);
  std::vector<size_t> arg_off(args.size());
  for (size_t i = 0; i < args.size(); ++i) {
    arg_off[i] = off;
    off = round_up(off + args[i].second, 16);
  }
  RPLSH_REQUIRE(off < (size_t)INT32_MAX, "persistent program too large");
  char* hp = S.hprog.as<char>(off);
  PersistHeader hdr{};
  hdr.nphases = (int)phases.size();
  hdr.total_units = units;
  std::memcpy(hp, &hdr, sizeof(hdr));
  for (size_t i = 0; i < phases.size(); ++i) {
    phases[i].arg = (int)arg_off[phases[i].arg];
    std::memcpy(hp + sizeof(PersistHeader) + i * sizeof(PersistPhase), &phases[i], sizeof(PersistPhase));
  }
  for (size_t i = 0; i < args.size(); ++i) std::memcpy(hp + arg_off[i], args[i].first, args[i].second);

  const int nsteps = (int)max_phases;
  char* dprog = S.prog.as<char>(off);
  int* claims = S.sync.as<int>(nsteps);
  RPLSH_HIP(hipMemcpyAsync(dprog, hp, off, hipMemcpyHostToDevice, st));
  launch_fill_i32(claims, nsteps, 0, st);
  // One batch: steps 1, 3, .., 2L-1 filter, 2, .., 2L select, 2L+1 re-rank.
  const int last_filter = nsteps - 3;
  auto rec = [&](int i) {
    if (pe && batches.size() == 1) pe->rec(i, st);
  };
  for (int sidx = 0; sidx < nsteps; ++sidx) {
    const int ub = step_unit[sidx], ue = step_unit[sidx + 1];
    if (sidx == 1) rec(0);
    if (sidx == last_filter) rec(1);
    if (sidx == nsteps - 1) rec(4);
    if (ue <= ub) continue;
    const char* cprog = dprog;
    int ph0 = step_phase[sidx], ph1 = step_phase[sidx + 1];
    int* claim = claims + sidx;
    void* kargs[] = {&cprog, &ph0, &ph1, const_cast<int*>(&ub), const_cast<int*>(&ue), &claim};
    if (step_filter[sidx]) {
      const int grid = std::min(ue - ub, S.resident_f);
      RPLSH_HIP(hipLaunchKernel(reinterpret_cast<const void*>(kern_f), dim3(grid), dim3(kPBlock), kargs, 0, st));
    } else {
      const int grid = std::min(ue - ub, S.resident);
      RPLSH_HIP(hipLaunchKernel(reinterpret_cast<const void*>(kern), dim3(grid), dim3(kPBlock), kargs, lds, st));
    }
    if (sidx == last_filter) rec(2);
    if (sidx == last_filter + 1) rec(3);
    if (sidx == nsteps - 1) rec(5);
  }
  // Every batch's counters come back with the caller's next synchronisation.
  for (const PersistBatch& pb : batches) {
    BatchCounters* hc = pb.ws->hctr.as<BatchCounters>(1);
    RPLSH_HIP(hipMemcpyAsync(hc, pb.ws->ctr.p, sizeof(BatchCounters), hipMemcpyDeviceToHost, st));
  }
}

}  // namespace rplsh
