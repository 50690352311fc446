// Internal engine interface shared by the search entry points (search.hip)
// and the serving loop (server.hip). Nothing here crosses the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"
#include "rplsh.h"

namespace rplsh {

// A fixed block of device memory handed out front to back (256-byte
// aligned pieces); never allocates. Used where a caller cannot allocate.
struct Arena {
  char* base = nullptr;
  size_t cap = 0;
  size_t used = 0;
  void* take(size_t b) {
    const size_t at = (used + 255) & ~(size_t)255;
    if (at > cap || b > cap - at)
      fail(RPLSH_ERR_OUT_OF_MEMORY, "scratch exhausted: %zu more bytes needed beyond the %zu-byte scratch",
           at + b - cap, cap);
    used = at + b;
    return base + at;
  }
};

// Growable device buffer (hipMalloc only when a call needs more than before),
// or, when `arena` is set, pieces of that arena (no allocation, no free).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  Arena* arena = nullptr;
  DevBuf() = default;
  explicit DevBuf(Arena* a) : arena(a) {}
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes), arena(o.arena) {
    o.p = nullptr;
    o.bytes = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      if (p && !arena) (void)hipFree(p);
      p = o.p;
      bytes = o.bytes;
      arena = o.arena;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DevBuf() {
    if (p && !arena) (void)hipFree(p);
  }
  void* ensure(size_t b) {
    if (b <= bytes) return p;
    if (arena) {
      p = arena->take(b);
      bytes = b;
      return p;
    }
    if (p) RPLSH_HIP(hipFree(p));
    p = nullptr;
    bytes = 0;
    size_t nb = std::max(b, (size_t)256);
    RPLSH_HIP(hipMalloc(&p, nb));
    bytes = nb;
    return p;
  }
  template <typename T>
  T* as(size_t count) {
    return static_cast<T*>(ensure(count * sizeof(T)));
  }
};

// Growable pinned host buffer (hipHostMalloc): staging for request I/O and
// counter read-backs, so every asynchronous copy has a page-locked end.
struct HostBuf {
  void* p = nullptr;
  size_t bytes = 0;
  HostBuf() = default;
  HostBuf(const HostBuf&) = delete;
  HostBuf& operator=(const HostBuf&) = delete;
  HostBuf(HostBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
    o.p = nullptr;
    o.bytes = 0;
  }
  HostBuf& operator=(HostBuf&& o) noexcept {
    if (this != &o) {
      if (p) (void)hipHostFree(p);
      p = o.p;
      bytes = o.bytes;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~HostBuf() {
    if (p) (void)hipHostFree(p);
  }
  void* ensure(size_t b) {
    if (b <= bytes) return p;
    if (p) RPLSH_HIP(hipHostFree(p));
    p = nullptr;
    bytes = 0;
    size_t nb = std::max(b, (size_t)4096);
    RPLSH_HIP(hipHostMalloc(&p, nb, hipHostMallocDefault));
    bytes = nb;
    return p;
  }
  template <typename T>
  T* as(size_t count) {
    return static_cast<T*>(ensure(count * sizeof(T)));
  }
};

// Per-batch counters, cleared at the start of a batch and read back with one
// copy after it.
struct BatchCounters {
  long long survivors;  // final-level survivors kept (sum over queries)
  int failed;           // queries to re-run
  int fp64;             // boundary candidates re-ranked in FP64
};

// Events around the phases of a batch (depth 0 only), read after the sync.
struct PhaseEvents {
  hipEvent_t ev[8] = {};
  bool on = false;
  void create() {
    for (auto& e : ev) RPLSH_HIP(hipEventCreate(&e));
    on = true;
  }
  ~PhaseEvents() {
    if (on)
      for (auto e : ev) (void)hipEventDestroy(e);
  }
  void rec(int i, hipStream_t st) {
    if (on) RPLSH_HIP(hipEventRecord(ev[i], st));
  }
  float ms(int a, int b) const {
    float t = 0.0f;
    if (on) RPLSH_HIP(hipEventElapsedTime(&t, ev[a], ev[b]));
    return t;
  }
};

struct Workspace {
  DevBuf qnorm, qenc, qscale, qpad, tau, tau2, counts, cand, topk, bound, status, relk;
  DevBuf flist, ctr;  // failed-query list; BatchCounters
  DevBuf qsub, isub, dsub;  // a retried subset of queries
  HostBuf hctr;             // pinned read-back target of ctr
  bool no_alloc = false;  // buffers from an arena; ctr read back into host stack memory; no events
  Workspace() = default;
  // Every device buffer taken from `a` (nothing allocated, counters read
  // back into pageable stack memory).
  explicit Workspace(Arena* a)
      : qnorm(a), qenc(a), qscale(a), qpad(a), tau(a), tau2(a), counts(a), cand(a), topk(a),
        bound(a), status(a), relk(a), flist(a), ctr(a), qsub(a), isub(a), dsub(a),
        no_alloc(true) {}
};

// The database as the search kernels see it.
struct DbView {
  const float* x = nullptr;  // FP32 rows [n][ldx]
  int64_t ldx = 0;
  int64_t n = 0;
  int d = 0;
  const float* xnorm = nullptr;
  float xn_max = 0.0f;
  Elem elem = Elem::F32;
  int dpad = 0;
  const void* xenc = nullptr;  // candidate rows [n][ldenc] of elem
  int64_t ldenc = 0;
  int dvalid_bytes = 0;
  const float* xscale = nullptr;  // FP8 per-row scales
  // Packed sample copy for the sample levels: whole tiles, row r is row
  // sample_source_row(r, sstep, stiles) (device.hpp), r < srows (0: none);
  // same layout, with its own norm and scale arrays. trank[u] (u <
  // rank_tiles, the packed copy's tiles): the sample tile holding tile u, or
  // INT32_MAX.
  const void* senc = nullptr;
  const float* sscale = nullptr;
  const float* snorm = nullptr;
  int64_t srows = 0;
  int64_t sstep = 0, stiles = 0;
  const int32_t* trank = nullptr;
  int64_t rank_tiles = 0;
  // -||x||^2 / 2 per row, with one -inf entry (index *_inf) for rows past
  // the end: the C operand the ring kernels copy into LDS as is.
  const float* xnh = nullptr;
  int64_t xnh_inf = 0;
  const float* snh = nullptr;
  int64_t snh_inf = 0;
  bool packed = false;            // xenc is in 32-row MFMA tiles (filter_packed.hip)
  // Cluster-residual copies (resid.hip), when built: rcopy[0] holds every
  // row, rcopy[j] (j > 0) a systematic sample of every rstride[j]-th row in
  // centroid order; each in tiles of one centroid. Every filter level then
  // reads one of them (resid_levels in search.hip).
  static constexpr int kMaxResidCopies = 8;
  int nrcopies = 0;
  const float* rcent = nullptr;
  struct ResidView {
    const void* enc = nullptr;
    const float* scale = nullptr;
    const float* nh = nullptr;
    const int32_t* rid = nullptr;
    const int32_t* slice_t0 = nullptr;
    const int32_t* slice_cl = nullptr;
    int64_t slices = 0;
    int64_t npad = 0;    // padded rows
    int64_t rows = 0;    // real rows
    int64_t stride = 1;  // sampling stride
  } rcopy[kMaxResidCopies];
};

struct SearchParams {
  int k = 0;
  int kp = 0;          // candidates re-ranked per query
  int exact = 1;
  int fp64_boundary = 0;
  bool certify = false;
  int c_factor = 3;    // last-level target survivors = c_factor * kp
  int target_wgs = 4096;  // filter workgroups aimed for (query tiles x database slices)
  int variant = 0;        // packed filter variant (FilterVariant) of the full level
  int sample_variant = -1;  // variant of the sample levels (-1: the same)
  bool l2_slices = false;   // full level: slices of at most kL2SliceBytes of rows
  int rho = 0;            // rank of the last threshold in its sample (0: automatic)
};

// Tuning entries (and the graph cache keys) carry both variants in one int:
// the full level's in bits 0-7, 1 + the sample levels' in bits 8-15 (0: the
// same kernel for every level), and in bit 16 whether the full level runs in
// L2-sized slices.
constexpr int kVariantL2Slices = 1 << 16;
inline int encode_variants(int full, int sample, bool l2_slices = false) {
  return full | (sample >= 0 ? (sample + 1) << 8 : 0) | (l2_slices ? kVariantL2Slices : 0);
}
inline int full_variant_of(int v) { return v & 0xff; }
inline int sample_variant_of(int v) { return ((v >> 8) & 0xff) - 1; }
inline bool l2_slices_of(int v) { return (v & kVariantL2Slices) != 0; }

// Rows per slice of a full level in L2-sized slices: a workgroup streams its
// slice while the other query blocks' workgroups of that slice, dispatched
// next to it on the same XCD, read it again; slices of at most this many
// bytes keep those re-reads in the XCD's 4 MB L2 beside the query operands
// (10M x 96 BF16, 20 query blocks: 2.13x the algorithmic bytes over HBM at
// ~12,000-row slices, 1.41x at 3,072, at the same speed).
constexpr int64_t kL2SliceBytes = 576 * 1024;


// The capturable part of a batch (no host synchronisation; allocates only
// when ws grows). nq_dev, if set, is a device int with the live query count.
void enqueue_batch(const DbView& db, const float* q, int nq, const int* nq_dev,
                   const SearchParams& p, int32_t* out_i, float* out_d, Workspace& ws,
                   hipStream_t st, PhaseEvents* pe);
// Waits for st, reads the counters (unless `ready` already holds them),
// re-runs failed queries eagerly.
void finish_batch(const DbView& db, const float* q, int nq, const SearchParams& p,
                  int32_t* out_i, float* out_d, std::vector<Workspace>& wss, int depth,
                  hipStream_t st, rplsh_search_stats& stats, const PhaseEvents* pe,
                  const BatchCounters* ready = nullptr);

// One batch of nq device-resident FP32 queries [nq][d] -> device outputs
// [nq][k], on stream st, with the workspaces wss (one per retry depth).
// Synchronises st.
void search_batch(const DbView& db, const float* q, int nq, const SearchParams& p,
                  int32_t* out_i, float* out_d, std::vector<Workspace>& wss, int depth,
                  hipStream_t st, rplsh_search_stats& stats);

// A batch as argument records for the persistent kernel (search.hip).
struct PersistBatchPlan {
  PrepArgs prep;
  std::vector<FilterArgs> filter;  // per level
  std::vector<SelectArgs> select;  // per level
  std::vector<int> count_reset;    // per level: counts after its select (-1: unchanged)
  RerankTailArgs rerank;
};
void plan_persistent_batch(const DbView& db, const float* q, int nq, const SearchParams& p,
                           int32_t* out_i, float* out_d, Workspace& ws, int64_t target_units,
                           PersistBatchPlan& out);

// Persistent serving kernel (persist.hip): whether it takes a batch, and
// one run over several batches of the same element kind and padded
// dimension (one launch per step of the batches' phases). Each batch's
// depth-0 results, status, failed list and counters land where enqueue_batch
// puts them (finish_batch completes the batch). Everything is enqueued on st;
// once st is synchronised each batch's counters are in ws->hctr.
bool persist_supported(const DbView& db, const SearchParams& p);
struct PersistState;  // per-stream program and synchronisation buffers
PersistState* persist_state_create();
void persist_state_destroy(PersistState* s);
struct PersistBatch {
  const DbView* db;
  const float* q;  // device FP32 queries
  int nq;
  const SearchParams* p;
  int32_t* out_i;
  float* out_d;
  Workspace* ws;
};
// pe (single batch only): phase events as search_batch records them (sample
// levels, full level, its select, re-rank).
void persist_run(PersistState& state, const std::vector<PersistBatch>& batches, hipStream_t st,
                 PhaseEvents* pe = nullptr);

// Device bytes search_batch allocates for one batch (depth-0 workspace).
size_t search_workspace_bytes(const DbView& db, int nq, const SearchParams& p);

// searchL2Distance on device buffers (plain FP32 path, results on `st`,
// synchronised). wss: workspaces per retry depth (4); `arena`, if set, backs
// the temporary norms / padded rows and every workspace buffer (no
// allocation anywhere); otherwise they come from the kept buffers passed in.
struct L2Scratch {
  DevBuf xpad, xn, xmax;
  std::vector<Workspace> wss;
};
void search_l2_device(const float* x, int64_t n, int d, const float* xnorms, const float* q, int64_t nq,
                      int k, int32_t* indices, float* distances, int exact, hipStream_t st,
                      L2Scratch& scratch);
// Device bytes search_l2_device needs for depth 0 plus a retry of
// min(nq, 64) queries (what the no-allocation entry point checks up front).
size_t search_l2_device_bytes(int64_t n, int d, bool have_norms, bool aligned, int64_t nq, int k);

// The view and the parameters IndexBF::search uses for k on an index; the
// candidate copy must already match the configured precision.
DbView bf_db_view(const rplsh_bf* b);
// The tuned entry for (k, nq) is used when the index has one (rplsh_bf_autotune).
SearchParams bf_search_params(const rplsh_bf* b, int k, int64_t nq = 0);
int bf_device_of(const rplsh_bf* b);
// Brings the candidate copy up to date with the configured precision.
void bf_prepare(rplsh_bf* b);

}  // namespace rplsh

// IndexBF state (C-ABI handle), shared by search.hip, server.hip and tune.hip.
struct rplsh_bf {
  rplsh_resources* res = nullptr;
  int dim = 0;
  rplsh_bf_config cfg{};
  int64_t n = 0;
  int64_t capacity = 0;
  int64_t ldx = 0;
  float* x = nullptr;
  float* xn = nullptr;
  float* xn_max_dev = nullptr;
  float xn_max = 0.0f;
  void* xenc = nullptr;
  float* xscale = nullptr;
  // Packed sample copy (sample levels): whole tiles, row r < srows is row
  // sample_source_row(r, sstep, n / 32); trank: its inverse per packed tile.
  void* senc = nullptr;
  int32_t* trank = nullptr;
  int64_t trank_cap = 0;
  float* sscale = nullptr;
  float* snorm = nullptr;
  float* snh = nullptr;   // -snorm / 2, then one -inf
  float* xnh = nullptr;   // packed copies: -||x||^2 / 2, then one -inf
  int64_t srows = 0, scap = 0, sstep = 0;
  int sprec = -1;  // precision the sample buffers were allocated for
  int enc_precision = -1;  // precision xenc currently holds
  int dpad = 0;
  // Cluster-residual copies (cfg.residual_clusters > 0, BF16 / FP8): rk
  // centroids rcent [rk][dim], built for the first rn rows; copies of
  // sampling stride 1, 8, 64, ... (rows: original row or -1 for padding,
  // tcl: centroid per tile, slices of one centroid), their residual rows for
  // precision rprec.
  struct ResidCopy {
    int64_t stride = 1, rows = 0, npad = 0, slices = 0;
    int32_t* rid = nullptr;
    int32_t* tcl = nullptr;
    int32_t* slice_t0 = nullptr;
    int32_t* slice_cl = nullptr;
    void* enc = nullptr;
    float* scale = nullptr;
    float* nh = nullptr;
  };
  std::vector<ResidCopy> rcopies;
  float* rcent = nullptr;
  int rk = 0;
  int64_t rn = 0;
  int rprec = -1;
  std::vector<rplsh::Workspace> ws;
  rplsh::DevBuf qdev, idev, ddev;
  rplsh_search_stats last{};
  std::vector<rplsh_tune_result> tunes;  // autotuned parameters per (k, batch size)
  rplsh::PersistState* pst = nullptr;     // persistent-kernel program buffers (small searches)
  // Servers holding the index (rplsh_server_add_index until the server is
  // destroyed): while non-zero every call that changes the rows, their
  // encoding or the tuning fails with RPLSH_ERR_STATE.
  std::atomic<int> serving{0};
  // Bumped by every such change; cached serving graphs record the value they
  // were captured at and are dropped when it moved.
  std::atomic<uint64_t> generation{0};
};

namespace rplsh {
// Cluster-residual copy of b's rows (resid.hip): (re)builds the centroids and
// layout when the rows changed since the last build, and the residual rows
// when the precision changed; synchronises st. resid_free releases it.
void resid_build(rplsh_bf* b, hipStream_t st);
void resid_free(rplsh_bf* b);

// Fails with RPLSH_ERR_STATE while a server holds the index; otherwise bumps
// its generation (the caller is about to change it).
void bf_begin_change(rplsh_bf* b);

// HBM budget (memory.hip): accounting-only charges for buffers their users
// own (the serving loop's batch workspaces and cached graphs), beside the
// block reservations of rplsh_memory_reserve.
rplsh_memory* memory_create(int device, size_t budget);
void memory_destroy(rplsh_memory* m);
bool memory_try_charge(rplsh_memory* m, size_t bytes);  // false if it would pass the budget
void memory_force_charge(rplsh_memory* m, size_t bytes);
void memory_uncharge(rplsh_memory* m, size_t bytes);
size_t memory_budget(const rplsh_memory* m);
int memory_device(const rplsh_memory* m);
size_t memory_reserved(rplsh_memory* m);
size_t memory_peak(rplsh_memory* m);
void memory_reset_peak(rplsh_memory* m);
}  // namespace rplsh
