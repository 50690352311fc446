// Serving form: continuous batching over resident indexes.
//
// Each of the `slots` workers owns a HIP stream (default priority) and a
// search workspace. An idle worker takes the oldest queued request, adds the
// queued requests for the same (index, k) behind it while the batch stays
// within max_batch_queries, and admits the batch when its workspace fits in
// the HBM budget left by the batches in flight; otherwise it waits until a
// finishing batch releases memory. Batches on different slots run
// concurrently on the device. Because every query is filtered, selected and
// re-ranked independently of the others in its batch, a request's result is
// the same alone and in any mix.
//
// Batches of up to graph_max_queries queries replay a hipGraph captured once
// per (index, k, power-of-two bucket) on the slot, from the shape's second
// batch on (a shape seen once runs eagerly): the pipeline up to the
// failed-query list is launch-free after capture, the live query count is a
// device scalar written before each replay, and the rare failed queries are
// resolved eagerly afterwards. Captured graphs hold their own workspaces;
// least recently used ones are evicted beyond graph_cache_bytes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cfloat>
#include <string>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.hpp"
#include "engine.hpp"
#include "rplsh.h"

using namespace rplsh;

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Request {
  uint64_t id = 0;
  int index = 0;
  rplsh_bf* b = nullptr;  // resolved under the lock at submit
  int64_t nq = 0;
  const float* q = nullptr;
  int k = 0;
  int32_t* I = nullptr;
  float* D = nullptr;
  Clock::time_point t_submit, t_admit, t_done;
  bool done = false;
  rplsh_request_info info{};
  std::string err;
};

// Everything a captured pipeline depends on besides its buffers: the index
// and its generation (bumped by any change to it), k, the batch bucket and
// the resolved search parameters. A batch replays a graph only if all match.
struct GraphKey {
  int index = -1;
  uint64_t gen = 0;
  int k = 0, bucket = 0;
  int kp = 0, c_factor = 0, target_wgs = 0, variant = 0, rho = 0, fp64_boundary = 0;
  bool certify = false;
  bool operator==(const GraphKey& o) const {
    return index == o.index && gen == o.gen && k == o.k && bucket == o.bucket && kp == o.kp &&
           c_factor == o.c_factor && target_wgs == o.target_wgs && variant == o.variant &&
           rho == o.rho && fp64_boundary == o.fp64_boundary && certify == o.certify;
  }
  uint64_t hash() const {
    uint64_t h = 1469598103934665603ull;
    for (uint64_t v : {(uint64_t)index, gen, (uint64_t)k, (uint64_t)bucket, (uint64_t)kp,
                       (uint64_t)c_factor, (uint64_t)target_wgs, (uint64_t)variant, (uint64_t)rho,
                       (uint64_t)fp64_boundary, (uint64_t)certify})
      h = (h ^ v) * 1099511628211ull;
    return h;
  }
};

GraphKey make_key(int index, const rplsh_bf* b, int k, int bucket, const SearchParams& p) {
  GraphKey g;
  g.index = index;
  g.gen = b->generation.load();
  g.k = k;
  g.bucket = bucket;
  g.kp = p.kp;
  g.c_factor = p.c_factor;
  g.target_wgs = p.target_wgs;
  g.variant = encode_variants(p.variant, p.sample_variant, p.l2_slices);
  g.rho = p.rho;
  g.fp64_boundary = p.fp64_boundary;
  g.certify = p.certify;
  return g;
}

// A captured pipeline: fixed device buffers for up to key.bucket queries,
// the live count in nqd, its own workspaces.
struct GraphEntry {
  GraphKey key;
  hipGraphExec_t exec = nullptr;
  std::vector<Workspace> wss;
  DevBuf q, I, D, nqd;
  size_t bytes = 0;
  uint64_t last_use = 0;
  ~GraphEntry() {
    if (exec) (void)hipGraphExecDestroy(exec);
  }
};

struct Slot {
  hipStream_t st = nullptr;
  std::vector<Workspace> wss;
  DevBuf qbuf, ibuf, dbuf;
  HostBuf hq, hi, hd;  // pinned staging of the batch's queries and results
  hipEvent_t ev0 = nullptr, ev1 = nullptr;  // device time of a batch
  std::vector<std::unique_ptr<GraphEntry>> graphs;  // touched by this slot's worker only
  size_t graph_bytes = 0;
  PersistState* pst = nullptr;                      // persistent kernel program buffers
  std::vector<std::vector<Workspace>> pws;          // workspaces per batch of a persistent launch
  std::thread th;
};

}  // namespace

struct rplsh_server {
  rplsh_server_config cfg{};
  std::vector<rplsh_bf*> indexes;
  std::mutex mu;
  std::condition_variable cv;       // workers: queue / budget changes
  std::condition_variable done_cv;  // waiters: completions
  std::deque<Request*> queue;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
  uint64_t next_id = 1;
  // The HBM budget batches are admitted against (own, or given in the
  // config): admitted batches charge their workspaces to it while they run,
  // cached graphs theirs while cached.
  rplsh_memory* mem = nullptr;
  bool own_mem = false;
  size_t graph_bytes = 0;  // workspaces held by cached graphs (all slots)
  int inflight = 0;
  bool stop = false;
  std::vector<std::unique_ptr<Slot>> slots;
  // statistics
  std::vector<double> lat;
  int64_t submitted = 0, completed = 0, rejected = 0, queries = 0, batches = 0;
  int peak_inflight = 0;
  bool have_first = false;
  Clock::time_point first_submit, last_done;
  int64_t graph_launches = 0, graph_captures = 0;
  int64_t persistent_launches = 0, persistent_batches = 0;
  uint64_t use_clock = 0;
  // Batches seen per graph key: a graph is captured on a key's second batch,
  // so one-off shapes never pay for a capture.
  std::unordered_map<uint64_t, int64_t> shape_seen;
};

namespace {

// Retried queries of one batch get at most this many queries' worth of
// workspace in the admission count (a retry re-runs only the failed queries,
// with 4x the survivor margin and k'; more failures than this still run, on
// memory the budget did not see).
constexpr int64_t kRetryReserveQueries = 64;

// Batches one persistent launch takes at most.
constexpr int kMaxPersistBatches = 16;

size_t retry_bytes(const DbView& v, int64_t nq, const SearchParams& p) {
  SearchParams p2 = p;
  p2.c_factor = p.c_factor * 4;
  p2.kp = (int)std::min<int64_t>(std::min<int64_t>(v.n, 4096), (int64_t)p.kp * 4);
  const int64_t r = std::min<int64_t>(nq, kRetryReserveQueries);
  return search_workspace_bytes(v, (int)r, p2) + (size_t)r * (v.d * 4 + (size_t)p.k * 8);
}

// Device bytes an eager batch of nq queries needs: its depth-0 workspace with
// the parameters it will run with, its query/result buffers and the retry
// reserve.
size_t batch_bytes(rplsh_bf* b, int64_t nq, int k) {
  const DbView v = bf_db_view(b);
  const SearchParams p = bf_search_params(b, k, nq);
  return search_workspace_bytes(v, (int)nq, p) + (size_t)nq * (v.d * 4 + (size_t)k * 8) +
         retry_bytes(v, nq, p);
}

void finish(rplsh_server* s, Request* r, int status) {
  r->t_done = Clock::now();
  r->done = true;
  r->info.status = status;
  r->info.latency_ms = ms_between(r->t_submit, r->t_done);
  r->info.queue_ms = ms_between(r->t_submit, r->t_admit);
  if (status == RPLSH_OK) {
    s->completed += 1;
    s->queries += r->nq;
    s->lat.push_back(r->info.latency_ms);
  } else if (status == RPLSH_ERR_BUDGET) {
    s->rejected += 1;
  }
  s->last_done = r->t_done;
}

size_t workspace_bytes(const std::vector<Workspace>& wss) {
  size_t b = 0;
  for (const auto& w : wss) {
    for (const DevBuf* d : {&w.qnorm, &w.qenc, &w.qscale, &w.qpad, &w.tau, &w.tau2, &w.counts,
                            &w.cand, &w.topk, &w.bound, &w.status, &w.relk, &w.flist, &w.ctr,
                            &w.qsub, &w.isub, &w.dsub})
      b += d->bytes;
  }
  return b;
}

int bucket_of(int64_t total) {
  int b = 16;
  while (b < total) b <<= 1;
  return b;
}

// How a batch will run, decided at admission by the worker that admits it.
struct Plan {
  rplsh_bf* b = nullptr;
  int index = 0, k = 0;
  int64_t total = 0;
  SearchParams p;
  GraphEntry* g = nullptr;  // a cached graph of this slot to replay
  bool capture = false;     // capture a graph for this batch's key
  GraphKey key;
  size_t need = 0;          // bytes reserved while the batch runs
  size_t graph_part = 0;    // of need: the graph's workspace, moved to the cache once captured
  std::unique_ptr<GraphEntry> temp_graph;  // captured but not cached (no room in the budget)
  // Persistent kernel: the head batch plus `extra` batches of other (index,
  // k) pairs of the same element kind and padded dimension, in one launch.
  bool persistent = false;
  struct Extra {
    rplsh_bf* b = nullptr;
    int index = 0, k = 0;
    int64_t total = 0;
    SearchParams p;
    std::vector<Request*> reqs;
  };
  std::vector<Extra> extra;
};

GraphEntry* find_graph(Slot& sl, const GraphKey& key) {
  for (auto& g : sl.graphs)
    if (g->key == key) return g.get();
  return nullptr;
}

// Bytes a graph for this key would hold (its bucket-sized workspace and
// buffers; the retry depths are allocated on use).
size_t graph_bytes_for(const DbView& v, const GraphKey& key, const SearchParams& p) {
  return search_workspace_bytes(v, key.bucket, p) + (size_t)key.bucket * (v.d * 4 + (size_t)key.k * 8) + 256;
}

// Drops this slot's graphs whose index changed since capture, then least
// recently used ones until `want` more bytes fit in the slot's cache share.
// Called with the server lock held, by the slot's own worker or while no
// batch is in flight (a graph is only used by a running batch of its slot).
void evict_graphs(rplsh_server* s, Slot& sl, size_t want) {
  const size_t share = s->cfg.graph_cache_bytes / (size_t)std::max(1, s->cfg.slots);
  for (auto it = sl.graphs.begin(); it != sl.graphs.end();) {
    const GraphKey& k = (*it)->key;
    if (k.gen != s->indexes[k.index]->generation.load()) {
      sl.graph_bytes -= (*it)->bytes;
      s->graph_bytes -= (*it)->bytes;
      memory_uncharge(s->mem, (*it)->bytes);
      it = sl.graphs.erase(it);
    } else {
      ++it;
    }
  }
  while (!sl.graphs.empty() && sl.graph_bytes + want > share) {
    auto it = std::min_element(sl.graphs.begin(), sl.graphs.end(),
                               [](const auto& a, const auto& b) { return a->last_use < b->last_use; });
    sl.graph_bytes -= (*it)->bytes;
    s->graph_bytes -= (*it)->bytes;
    memory_uncharge(s->mem, (*it)->bytes);
    sl.graphs.erase(it);
  }
}

// Captures the pipeline for p.key on this slot (without the server lock).
std::unique_ptr<GraphEntry> capture_graph(Slot& sl, const DbView& v, const Plan& pl) {
  auto g = std::make_unique<GraphEntry>();
  g->key = pl.key;
  g->wss.resize(4);
  const int bucket = pl.key.bucket, k = pl.k;
  float* q = g->q.as<float>((size_t)bucket * v.d);
  int32_t* I = g->I.as<int32_t>((size_t)bucket * k);
  float* D = g->D.as<float>((size_t)bucket * k);
  int* nqd = g->nqd.as<int>(1);
  // Size the workspace with a run over zero live queries, then capture.
  launch_fill_i32(nqd, 1, 0, sl.st);
  enqueue_batch(v, q, bucket, nqd, pl.p, I, D, g->wss[0], sl.st, nullptr);
  RPLSH_HIP(hipStreamSynchronize(sl.st));
  hipGraph_t graph = nullptr;
  RPLSH_HIP(hipStreamBeginCapture(sl.st, hipStreamCaptureModeThreadLocal));
  try {
    enqueue_batch(v, q, bucket, nqd, pl.p, I, D, g->wss[0], sl.st, nullptr);
  } catch (...) {
    (void)hipStreamEndCapture(sl.st, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  RPLSH_HIP(hipStreamEndCapture(sl.st, &graph));
  const hipError_t ie = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  RPLSH_HIP(ie);
  g->bytes = workspace_bytes(g->wss) + g->q.bytes + g->I.bytes + g->D.bytes + g->nqd.bytes;
  return g;
}

// Runs one admitted batch on a slot (called without the lock). Request I/O
// goes through the slot's pinned staging buffers: one host-to-device copy of
// the batch's queries, one device-to-host copy of each result array.
int run_batch(rplsh_server* s, Slot& sl, Plan& pl, const std::vector<Request*>& batch,
              std::string& err, float& device_ms) {
  try {
    const DbView v = bf_db_view(pl.b);
    const int k = pl.k;
    const int64_t total = pl.total;
    if (pl.capture) {
      std::unique_ptr<GraphEntry> g = capture_graph(sl, v, pl);
      std::lock_guard<std::mutex> lk(s->mu);
      evict_graphs(s, sl, g->bytes);
      // The estimate charged at admission becomes the graph's real size;
      // this slot's other cached graphs (none in use) make room if needed.
      memory_uncharge(s->mem, pl.graph_part);
      pl.need -= pl.graph_part;
      pl.graph_part = 0;
      while (!memory_try_charge(s->mem, g->bytes)) {
        if (sl.graphs.empty()) {
          // No room to keep it: charged while this batch runs, then dropped.
          memory_force_charge(s->mem, g->bytes);
          pl.need += g->bytes;
          pl.temp_graph = std::move(g);
          break;
        }
        evict_graphs(s, sl, SIZE_MAX / 2);
      }
      s->graph_captures += 1;
      if (pl.temp_graph) {
        pl.g = pl.temp_graph.get();
      } else {
        sl.graph_bytes += g->bytes;
        s->graph_bytes += g->bytes;
        g->last_use = ++s->use_clock;
        sl.graphs.push_back(std::move(g));
        pl.g = sl.graphs.back().get();
      }
    }
    float* q;
    int32_t* I;
    float* D;
    if (pl.g) {
      q = static_cast<float*>(pl.g->q.p);
      I = static_cast<int32_t*>(pl.g->I.p);
      D = static_cast<float*>(pl.g->D.p);
    } else {
      q = sl.qbuf.as<float>((size_t)total * v.d);
      I = sl.ibuf.as<int32_t>((size_t)total * k);
      D = sl.dbuf.as<float>((size_t)total * k);
    }
    float* hq = sl.hq.as<float>((size_t)total * v.d);
    int32_t* hi = sl.hi.as<int32_t>((size_t)total * k);
    float* hd = sl.hd.as<float>((size_t)total * k);
    int64_t off = 0;
    for (auto* r : batch) {
      std::memcpy(hq + off * v.d, r->q, (size_t)r->nq * v.d * sizeof(float));
      off += r->nq;
    }
    RPLSH_HIP(hipEventRecord(sl.ev0, sl.st));
    RPLSH_HIP(hipMemcpyAsync(q, hq, (size_t)total * v.d * sizeof(float), hipMemcpyHostToDevice, sl.st));
    rplsh_search_stats stats{};
    if (v.n == 0) {
      // Empty index: FLT_MAX / -1 (search.cu:93-100).
      launch_fill_i32(I, total * k, -1, sl.st);
      launch_fill_f32(D, total * k, FLT_MAX, sl.st);
    } else if (pl.g) {
      launch_fill_i32(static_cast<int*>(pl.g->nqd.p), 1, (int)total, sl.st);
      RPLSH_HIP(hipGraphLaunch(pl.g->exec, sl.st));
      finish_batch(v, q, (int)total, pl.p, I, D, pl.g->wss, 0, sl.st, stats, nullptr);
    } else {
      search_batch(v, q, (int)total, pl.p, I, D, sl.wss, 0, sl.st, stats);
    }
    RPLSH_HIP(hipMemcpyAsync(hi, I, (size_t)total * k * sizeof(int32_t), hipMemcpyDeviceToHost, sl.st));
    RPLSH_HIP(hipMemcpyAsync(hd, D, (size_t)total * k * sizeof(float), hipMemcpyDeviceToHost, sl.st));
    RPLSH_HIP(hipEventRecord(sl.ev1, sl.st));
    RPLSH_HIP(hipStreamSynchronize(sl.st));
    RPLSH_HIP(hipEventElapsedTime(&device_ms, sl.ev0, sl.ev1));
    off = 0;
    for (auto* r : batch) {
      std::memcpy(r->I, hi + off * k, (size_t)r->nq * k * sizeof(int32_t));
      std::memcpy(r->D, hd + off * k, (size_t)r->nq * k * sizeof(float));
      off += r->nq;
    }
    return RPLSH_OK;
  } catch (const Error& e) {
    err = e.what();
    return e.status;
  } catch (const std::bad_alloc&) {
    err = "host allocation failed";
    return RPLSH_ERR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    err = e.what();
    return RPLSH_ERR_STATE;
  }
}

// Runs a persistent-kernel group on a slot (called without the lock): every
// batch's queries in one host-to-device copy, the kernel's steps over all
// batches, the rare failed queries re-run eagerly per batch, results back in
// one device-to-host copy per array.
int run_persistent(rplsh_server* s, Slot& sl, Plan& pl, const std::vector<Request*>& head,
                   std::string& err, float& device_ms) {
  try {
    struct B {
      rplsh_bf* b;
      int k;
      int64_t nq;
      SearchParams p;
      const std::vector<Request*>* reqs;
      DbView v;
      size_t qoff, ooff;
    };
    std::vector<B> bs;
    bs.push_back({pl.b, pl.k, pl.total, pl.p, &head, bf_db_view(pl.b), 0, 0});
    for (const auto& e : pl.extra) bs.push_back({e.b, e.k, e.total, e.p, &e.reqs, bf_db_view(e.b), 0, 0});
    size_t qn = 0, on = 0;
    for (auto& x : bs) {
      x.qoff = qn;
      x.ooff = on;
      qn += (size_t)x.nq * x.v.d;
      on += (size_t)x.nq * x.k;
    }
    float* q = sl.qbuf.as<float>(qn);
    int32_t* I = sl.ibuf.as<int32_t>(on);
    float* D = sl.dbuf.as<float>(on);
    float* hq = sl.hq.as<float>(qn);
    int32_t* hi = sl.hi.as<int32_t>(on);
    float* hd = sl.hd.as<float>(on);
    for (auto& x : bs) {
      size_t off = x.qoff;
      for (auto* r : *x.reqs) {
        std::memcpy(hq + off, r->q, (size_t)r->nq * x.v.d * sizeof(float));
        off += (size_t)r->nq * x.v.d;
      }
    }
    if (sl.pws.size() < bs.size()) sl.pws.resize(bs.size());
    std::vector<PersistBatch> list;
    for (size_t i = 0; i < bs.size(); ++i) {
      if (sl.pws[i].size() < 4) sl.pws[i].resize(4);
      const B& x = bs[i];
      list.push_back({&x.v, q + x.qoff, (int)x.nq, &x.p, I + x.ooff, D + x.ooff, &sl.pws[i][0]});
    }
    RPLSH_HIP(hipEventRecord(sl.ev0, sl.st));
    RPLSH_HIP(hipMemcpyAsync(q, hq, qn * sizeof(float), hipMemcpyHostToDevice, sl.st));
    persist_run(*sl.pst, list, sl.st);
    // Results and counters in one synchronisation; the rare batch with failed
    // queries re-runs them (finish_batch) and copies its results again.
    RPLSH_HIP(hipMemcpyAsync(hi, I, on * sizeof(int32_t), hipMemcpyDeviceToHost, sl.st));
    RPLSH_HIP(hipMemcpyAsync(hd, D, on * sizeof(float), hipMemcpyDeviceToHost, sl.st));
    RPLSH_HIP(hipEventRecord(sl.ev1, sl.st));
    RPLSH_HIP(hipStreamSynchronize(sl.st));
    rplsh_search_stats stats{};
    bool redo = false;
    for (size_t i = 0; i < bs.size(); ++i) {
      const B& x = bs[i];
      const BatchCounters* hc = static_cast<const BatchCounters*>(sl.pws[i][0].hctr.p);
      if (hc->failed == 0) continue;
      finish_batch(x.v, q + x.qoff, (int)x.nq, x.p, I + x.ooff, D + x.ooff, sl.pws[i], 0, sl.st, stats,
                   nullptr, hc);
      const size_t cnt = (size_t)x.nq * x.k;
      RPLSH_HIP(hipMemcpyAsync(hi + x.ooff, I + x.ooff, cnt * sizeof(int32_t), hipMemcpyDeviceToHost, sl.st));
      RPLSH_HIP(hipMemcpyAsync(hd + x.ooff, D + x.ooff, cnt * sizeof(float), hipMemcpyDeviceToHost, sl.st));
      redo = true;
    }
    if (redo) {
      RPLSH_HIP(hipEventRecord(sl.ev1, sl.st));
      RPLSH_HIP(hipStreamSynchronize(sl.st));
    }
    RPLSH_HIP(hipEventElapsedTime(&device_ms, sl.ev0, sl.ev1));
    for (auto& x : bs) {
      size_t off = x.ooff;
      for (auto* r : *x.reqs) {
        std::memcpy(r->I, hi + off, (size_t)r->nq * x.k * sizeof(int32_t));
        std::memcpy(r->D, hd + off, (size_t)r->nq * x.k * sizeof(float));
        off += (size_t)r->nq * x.k;
      }
    }
    (void)s;
    return RPLSH_OK;
  } catch (const Error& e) {
    err = e.what();
    return e.status;
  } catch (const std::bad_alloc&) {
    err = "host allocation failed";
    return RPLSH_ERR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    err = e.what();
    return RPLSH_ERR_STATE;
  }
}

// Completes a request that can never be admitted (called with the lock held).
void reject(rplsh_server* s, Request* r, int status, const std::string& why) {
  r->t_admit = Clock::now();
  r->err = why;
  finish(s, r, status);
  s->done_cv.notify_all();
}

void worker(rplsh_server* s, Slot* sl) {
  (void)hipSetDevice(s->cfg.device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    std::vector<Request*> batch;
    Plan pl;
    for (;;) {
      if (s->queue.empty()) {
        if (s->stop) return;
        s->cv.wait(lk);
        continue;
      }
      Request* head = s->queue.front();
      rplsh_bf* b = head->b;
      size_t hb = 0;
      std::string why;
      try {
        hb = batch_bytes(b, head->nq, head->k);
      } catch (const std::exception& e) {
        hb = SIZE_MAX;
        why = e.what();
      }
      const size_t budget = memory_budget(s->mem);
      if (hb == SIZE_MAX || hb > budget) {
        s->queue.pop_front();
        reject(s, head, hb == SIZE_MAX ? RPLSH_ERR_STATE : RPLSH_ERR_BUDGET,
               hb == SIZE_MAX ? why : "batch workspace exceeds the HBM budget");
        continue;
      }
      size_t avail = budget - std::min(budget, memory_reserved(s->mem));
      if (hb > avail) {
        if (s->inflight == 0 && s->graph_bytes > 0) {
          // Cached graphs stand in the way and nothing runs, so none is in
          // use: release them all.
          for (auto& o : s->slots) evict_graphs(s, *o, SIZE_MAX / 2);
          continue;
        }
        // A finishing batch, or a reservation released elsewhere on a shared
        // budget, makes room: look again shortly.
        s->cv.wait_for(lk, std::chrono::milliseconds(2));
        continue;
      }
      // Admit the head and the same-(index, k) requests queued behind it.
      s->queue.pop_front();
      batch.push_back(head);
      int64_t total = head->nq;
      size_t need = hb;
      for (auto it = s->queue.begin(); it != s->queue.end();) {
        Request* r = *it;
        if (r->index == head->index && r->k == head->k &&
            total + r->nq <= s->cfg.max_batch_queries) {
          size_t nb = SIZE_MAX;
          try {
            nb = batch_bytes(b, total + r->nq, head->k);
          } catch (...) {
          }
          if (nb != SIZE_MAX && nb <= avail) {
            batch.push_back(r);
            total += r->nq;
            need = nb;
            it = s->queue.erase(it);
            continue;
          }
        }
        ++it;
      }
      pl.b = b;
      pl.index = head->index;
      pl.k = head->k;
      pl.total = total;
      pl.need = need;
      try {
        pl.p = bf_search_params(b, pl.k, total);
        const DbView v = bf_db_view(b);
        if (s->cfg.persistent_max_queries > 0 && total <= s->cfg.persistent_max_queries &&
            persist_supported(v, pl.p)) {
          // Persistent kernel: add queued batches of other (index, k) pairs of
          // the same element kind and padded dimension, each within the
          // per-batch query limit, while the group's workspaces fit.
          pl.persistent = true;
          size_t group_need = need;
          int64_t group_q = total;
          const int64_t max_group_q = 4 * (int64_t)s->cfg.persistent_max_queries;
          for (auto it = s->queue.begin(); it != s->queue.end();) {
            Request* r = *it;
            bool taken = false;
            if (group_q + r->nq <= max_group_q && r->nq <= s->cfg.persistent_max_queries) {
              const DbView rv = bf_db_view(r->b);
              if (rv.elem == v.elem && rv.dpad == v.dpad && rv.n > 0) {
                Plan::Extra* e = nullptr;
                for (auto& x : pl.extra)
                  if (x.index == r->index && x.k == r->k) e = &x;
                const bool fresh = e == nullptr;
                if (fresh && pl.extra.size() + 1 >= (size_t)kMaxPersistBatches) {
                  ++it;
                  continue;
                }
                const int64_t nt = (fresh ? 0 : e->total) + r->nq;
                try {
                  const SearchParams rp = bf_search_params(r->b, r->k, nt);
                  if (nt <= s->cfg.persistent_max_queries && persist_supported(rv, rp)) {
                    const size_t before = fresh ? 0 : batch_bytes(r->b, e->total, r->k);
                    const size_t after = batch_bytes(r->b, nt, r->k);
                    if (group_need - before + after <= avail) {
                      if (fresh) {
                        pl.extra.emplace_back();
                        e = &pl.extra.back();
                        e->b = r->b;
                        e->index = r->index;
                        e->k = r->k;
                      }
                      e->total = nt;
                      e->p = rp;
                      e->reqs.push_back(r);
                      group_need = group_need - before + after;
                      group_q += r->nq;
                      taken = true;
                    }
                  }
                } catch (...) {
                }
              }
            }
            if (taken) it = s->queue.erase(it);
            else ++it;
          }
          pl.need = group_need;
        } else if (s->cfg.use_graphs && v.n > 0 && total <= s->cfg.graph_max_queries) {
          pl.key = make_key(pl.index, b, pl.k, bucket_of(total), pl.p);
          pl.g = find_graph(*sl, pl.key);
          if (pl.g) {
            pl.need = retry_bytes(v, total, pl.p);
            pl.g->last_use = ++s->use_clock;
          } else if (++s->shape_seen[pl.key.hash()] >= 2) {
            // Capture if the graph's workspace fits beside what runs now.
            const size_t gw = graph_bytes_for(v, pl.key, pl.p);
            const size_t rb = retry_bytes(v, total, pl.p);
            if (gw + rb <= avail && gw <= s->cfg.graph_cache_bytes / (size_t)std::max(1, s->cfg.slots)) {
              pl.capture = true;
              pl.need = gw + rb;
              pl.graph_part = gw;
            }
          }
        }
      } catch (...) {
        pl.g = nullptr;
        pl.capture = false;
        pl.need = need;
        pl.graph_part = 0;
        for (auto eit = pl.extra.rbegin(); eit != pl.extra.rend(); ++eit)
          for (auto rit = eit->reqs.rbegin(); rit != eit->reqs.rend(); ++rit) s->queue.push_front(*rit);
        pl.extra.clear();
        pl.persistent = false;
      }
      if (!memory_try_charge(s->mem, pl.need)) {
        // Taken meanwhile by a reservation on a shared budget: requeue.
        for (auto eit = pl.extra.rbegin(); eit != pl.extra.rend(); ++eit)
          for (auto rit = eit->reqs.rbegin(); rit != eit->reqs.rend(); ++rit) s->queue.push_front(*rit);
        for (auto rit = batch.rbegin(); rit != batch.rend(); ++rit) s->queue.push_front(*rit);
        batch.clear();
        pl = Plan();
        s->cv.wait_for(lk, std::chrono::milliseconds(2));
        continue;
      }
      break;
    }
    const auto now = Clock::now();
    const int group = pl.persistent ? 1 + (int)pl.extra.size() : 0;
    for (auto* r : batch) {
      r->t_admit = now;
      r->info.batch_requests = (int)batch.size();
      r->info.batch_queries = pl.total;
      r->info.kernel_batches = group;
    }
    for (auto& e : pl.extra)
      for (auto* r : e.reqs) {
        r->t_admit = now;
        r->info.batch_requests = (int)e.reqs.size();
        r->info.batch_queries = e.total;
        r->info.kernel_batches = group;
      }
    s->inflight += 1;
    s->batches += 1 + (int64_t)pl.extra.size();
    s->peak_inflight = std::max(s->peak_inflight, s->inflight);
    const bool replay = pl.g != nullptr || pl.capture;
    if (replay) s->graph_launches += 1;
    if (pl.persistent) {
      s->persistent_launches += 1;
      s->persistent_batches += group;
    }
    lk.unlock();
    std::string err;
    float device_ms = 0.0f;
    const int st = pl.persistent ? run_persistent(s, *sl, pl, batch, err, device_ms)
                                 : run_batch(s, *sl, pl, batch, err, device_ms);
    lk.lock();
    memory_uncharge(s->mem, pl.need);
    s->inflight -= 1;
    for (auto* r : batch) {
      r->err = err;
      r->info.device_ms = device_ms;
      finish(s, r, st);
    }
    for (auto& e : pl.extra)
      for (auto* r : e.reqs) {
        r->err = err;
        r->info.device_ms = device_ms;
        finish(s, r, st);
      }
    s->done_cv.notify_all();
    s->cv.notify_all();
  }
}

void destroy_slots(rplsh_server* s) {
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  for (auto& sl : s->slots)
    if (sl->th.joinable()) sl->th.join();
  DeviceScope scope(s->cfg.device);
  for (auto& sl : s->slots) {
    if (sl->st) (void)hipStreamSynchronize(sl->st);
    sl->graphs.clear();
    sl->wss.clear();
    sl->pws.clear();
    if (sl->pst) persist_state_destroy(sl->pst);
    sl->pst = nullptr;
    sl->qbuf = DevBuf();
    sl->ibuf = DevBuf();
    sl->dbuf = DevBuf();
    sl->hq = HostBuf();
    sl->hi = HostBuf();
    sl->hd = HostBuf();
    if (sl->ev0) (void)hipEventDestroy(sl->ev0);
    if (sl->ev1) (void)hipEventDestroy(sl->ev1);
    if (sl->st) (void)hipStreamDestroy(sl->st);
  }
  for (rplsh_bf* b : s->indexes) b->serving.fetch_sub(1);
  s->indexes.clear();
}

}  // namespace

extern "C" {

void rplsh_server_config_default(rplsh_server_config* cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->device = 0;
  cfg->slots = 4;
  cfg->hbm_budget_bytes = 0;
  cfg->max_batch_queries = 4096;
  cfg->use_graphs = 1;
  cfg->graph_max_queries = 1024;
  cfg->graph_cache_bytes = (size_t)8 << 30;
  cfg->persistent_max_queries = 256;
}

rplsh_status rplsh_server_create(const rplsh_server_config* cfg, rplsh_server** out) {
  return guard([&] {
    RPLSH_REQUIRE(out, "out is NULL");
    rplsh_server_config c;
    rplsh_server_config_default(&c);
    if (cfg) c = *cfg;
    if (c.slots <= 0) c.slots = 4;
    if (c.max_batch_queries <= 0) c.max_batch_queries = 4096;
    if (c.graph_max_queries <= 0) c.graph_max_queries = 1024;
    if (c.graph_cache_bytes == 0) c.graph_cache_bytes = (size_t)8 << 30;
    int count = 0;
    rplsh_device_count(&count);
    RPLSH_REQUIRE(c.device >= 0 && c.device < count, "device %d out of range [0, %d)", c.device, count);
    DeviceScope scope(c.device);
    if (c.memory) {
      RPLSH_REQUIRE(memory_device(c.memory) == c.device, "budget lives on device %d, server on %d",
                    memory_device(c.memory), c.device);
      c.hbm_budget_bytes = memory_budget(c.memory);
    } else if (c.hbm_budget_bytes == 0) {
      size_t fr = 0, tot = 0;
      RPLSH_HIP(hipMemGetInfo(&fr, &tot));
      c.hbm_budget_bytes = fr / 4;
    }
    auto* s = new rplsh_server();
    s->cfg = c;
    if (c.memory) {
      s->mem = c.memory;
    } else {
      s->mem = memory_create(c.device, c.hbm_budget_bytes);
      s->own_mem = true;
    }
    try {
      for (int i = 0; i < c.slots; ++i) {
        auto sl = std::make_unique<Slot>();
        RPLSH_HIP(hipStreamCreateWithFlags(&sl->st, hipStreamNonBlocking));
        RPLSH_HIP(hipEventCreate(&sl->ev0));
        RPLSH_HIP(hipEventCreate(&sl->ev1));
        sl->wss.resize(4);
        sl->pst = persist_state_create();
        s->slots.push_back(std::move(sl));
      }
      for (auto& sl : s->slots) sl->th = std::thread(worker, s, sl.get());
    } catch (...) {
      destroy_slots(s);  // stops and joins the workers already started
      if (s->own_mem) memory_destroy(s->mem);
      delete s;
      throw;
    }
    *out = s;
  });
}

rplsh_status rplsh_server_destroy(rplsh_server* s) {
  return guard([&] {
    if (!s) return;
    destroy_slots(s);
    if (s->own_mem) memory_destroy(s->mem);
    delete s;
  });
}

rplsh_status rplsh_server_add_index(rplsh_server* s, rplsh_bf* index, int* index_id) {
  return guard([&] {
    RPLSH_REQUIRE(s && index && index_id, "NULL argument");
    RPLSH_REQUIRE(bf_device_of(index) == s->cfg.device, "index lives on another device");
    bf_prepare(index);
    std::lock_guard<std::mutex> lk(s->mu);
    RPLSH_REQUIRE(!s->stop, "server is stopping");
    s->indexes.push_back(index);
    index->serving.fetch_add(1);
    *index_id = (int)s->indexes.size() - 1;
  });
}

rplsh_status rplsh_server_submit(rplsh_server* s, int index_id, int64_t nq, const float* queries,
                                 int k, int32_t* indices, float* distances, uint64_t* request) {
  return guard([&] {
    RPLSH_REQUIRE(s && request, "NULL argument");
    RPLSH_REQUIRE(nq > 0 && queries && indices && distances, "empty request or NULL buffer");
    RPLSH_REQUIRE(k > 0, "k must be > 0");
    std::lock_guard<std::mutex> lk(s->mu);
    RPLSH_REQUIRE(index_id >= 0 && index_id < (int)s->indexes.size(), "unknown index %d", index_id);
    RPLSH_REQUIRE(!s->stop, "server is stopping");
    auto r = std::make_unique<Request>();
    r->id = s->next_id++;
    r->index = index_id;
    r->b = s->indexes[index_id];
    r->nq = nq;
    r->q = queries;
    r->k = k;
    r->I = indices;
    r->D = distances;
    r->t_submit = Clock::now();
    if (!s->have_first) {
      s->have_first = true;
      s->first_submit = r->t_submit;
    }
    s->submitted += 1;
    *request = r->id;
    s->queue.push_back(r.get());
    s->reqs.emplace(r->id, std::move(r));
    s->cv.notify_one();
  });
}

rplsh_status rplsh_server_poll(rplsh_server* s, uint64_t request, int* done) {
  return guard([&] {
    RPLSH_REQUIRE(s && done, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(request);
    RPLSH_REQUIRE(it != s->reqs.end(), "unknown request %llu", (unsigned long long)request);
    *done = it->second->done ? 1 : 0;
  });
}

rplsh_status rplsh_server_wait(rplsh_server* s, uint64_t request, double timeout_s,
                               rplsh_request_info* info) {
  rplsh_status st = guard([&] {
    RPLSH_REQUIRE(s, "server is NULL");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(request);
    RPLSH_REQUIRE(it != s->reqs.end(), "unknown request %llu", (unsigned long long)request);
    Request* r = it->second.get();
    auto pred = [&] { return r->done; };
    if (timeout_s < 0) {
      s->done_cv.wait(lk, pred);
    } else if (!s->done_cv.wait_for(lk, std::chrono::duration<double>(timeout_s), pred)) {
      fail(RPLSH_ERR_STATE, "request %llu not done after %.3f s", (unsigned long long)request, timeout_s);
    }
    if (info) *info = r->info;
    const int rs = r->info.status;
    const std::string err = r->err;
    s->reqs.erase(it);
    if (rs != RPLSH_OK)
      fail((rplsh_status)rs, "request %llu failed: %s", (unsigned long long)request, err.c_str());
  });
  return st;
}

rplsh_status rplsh_server_drain(rplsh_server* s) {
  return guard([&] {
    RPLSH_REQUIRE(s, "server is NULL");
    std::unique_lock<std::mutex> lk(s->mu);
    s->done_cv.wait(lk, [&] { return s->queue.empty() && s->inflight == 0; });
  });
}

rplsh_status rplsh_server_get_stats(rplsh_server* s, rplsh_server_stats* out) {
  return guard([&] {
    RPLSH_REQUIRE(s && out, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(out, 0, sizeof(*out));
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->rejected = s->rejected;
    out->queries = s->queries;
    out->batches = s->batches;
    out->peak_reserved_bytes = memory_peak(s->mem);
    out->peak_inflight = s->peak_inflight;
    out->graph_launches = s->graph_launches;
    out->graph_captures = s->graph_captures;
    out->graph_cache_bytes = s->graph_bytes;
    out->persistent_launches = s->persistent_launches;
    out->persistent_batches = s->persistent_batches;
    if (!s->lat.empty()) {
      std::vector<double> v = s->lat;
      std::sort(v.begin(), v.end());
      auto pct = [&](double p) {
        size_t i = (size_t)std::ceil(p * (double)v.size()) - 1;
        return v[std::min(i, v.size() - 1)];
      };
      out->p50_ms = pct(0.50);
      out->p99_ms = pct(0.99);
      double sum = 0;
      for (double x : v) sum += x;
      out->mean_ms = sum / (double)v.size();
      out->max_ms = v.back();
    }
    if (s->have_first && s->completed > 0) {
      out->wall_s = ms_between(s->first_submit, s->last_done) / 1000.0;
      out->qps = out->wall_s > 0 ? (double)s->queries / out->wall_s : 0.0;
    }
  });
}

rplsh_status rplsh_server_reset_stats(rplsh_server* s) {
  return guard([&] {
    RPLSH_REQUIRE(s, "server is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    s->lat.clear();
    s->submitted = s->completed = s->rejected = s->queries = s->batches = 0;
    s->peak_inflight = 0;
    s->have_first = false;
    s->graph_launches = s->graph_captures = 0;
    s->persistent_launches = s->persistent_batches = 0;
    memory_reset_peak(s->mem);
  });
}

rplsh_status rplsh_server_memory(rplsh_server* s, rplsh_memory** out) {
  return guard([&] {
    RPLSH_REQUIRE(s && out, "NULL argument");
    *out = s->mem;
  });
}

rplsh_status rplsh_bf_batch_bytes(rplsh_bf* index, int64_t nq, int k, size_t* bytes) {
  return guard([&] {
    RPLSH_REQUIRE(index && bytes, "NULL argument");
    RPLSH_REQUIRE(nq >= 0 && k > 0, "bad arguments");
    bf_prepare(index);
    *bytes = batch_bytes(index, nq, k);
  });
}

}  // extern "C"
