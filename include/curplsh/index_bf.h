// IndexBF (reference: include/index_bf.h:8-46): exact brute-force L2 kNN.
//
// Drop-in for the reference class: same constructor, configs and overrides.
// Pointers passed to add / search may be host or device memory (as
// index_bf.cpp:56-66 intends). search always returns exact squared L2
// distances (index_bf.cpp:68-69), ascending, ties broken by the lower index.
// Extensions beyond the reference, all defaulting to its behaviour:
//   * IndexBFConfig::precision / kprimeFactor / fp64Boundary select the
//     precision forms (BF16 or FP8 candidates + FP32/FP64 re-rank);
//   * add may be called repeatedly (the reference allows one add, FIXME at
//     index_bf.cpp:27);
//   * k > number of vectors pads with -1 / FLT_MAX instead of asserting.
#pragma once

#include "index.h"

namespace curplsh {

enum class Precision {
  /// Plain path: FP32 MFMA candidates + canonical FP32 re-rank (bit-exact).
  FP32 = RPLSH_PREC_FP32,
  /// BF16 candidates, k' = kprimeFactor * k, FP32 re-rank, FP64 boundary.
  BF16 = RPLSH_PREC_BF16,
  /// FP8 e4m3 (per-row power-of-two scale) candidates, same re-rank.
  FP8 = RPLSH_PREC_FP8,
};

struct IndexBFConfig : public IndexConfig {
  IndexBFConfig()
      : storeTransposed(false),
        computeExactDistances(false),
        precision(Precision::FP32),
        kprimeFactor(4),
        fp64Boundary(true),
        persistentMaxQueries(32),
        residualClusters(0) {}

  /// Accepted for compatibility; the resident layout is always row-major
  /// FP32 plus the candidate copy of the configured precision.
  bool storeTransposed;

  /// Kept for compatibility: search returns exact distances regardless, as
  /// the reference does (index_bf.cpp:68-69).
  bool computeExactDistances;

  Precision precision;
  int kprimeFactor;
  bool fp64Boundary;
  /// BF16/FP8 searches of up to this many queries run in the persistent
  /// kernel (same results); 0: never.
  int persistentMaxQueries;
  /// BF16/FP8 on clustered rows: > 0 adds a candidate copy of the rows minus
  /// the nearest of this many centroids, read by the full filter level
  /// (keeps the low-precision error at the scale of a cluster's spread).
  int residualClusters;
};

class IndexBF : public Index {
 public:
  IndexBF(DeviceResources* resources, int dim, IndexBFConfig config = IndexBFConfig());

  ~IndexBF() override;

  void train(int num, const float* data) override;

  void add(int num, const float* data) override;

  void search(int num, const float* queries, int k, int* indices, float* distances) override;

  void reset() override;

  /// Switches the candidate precision (re-encodes the resident vectors).
  void setPrecision(Precision p, int kprimeFactor = 0);

  /// Writes the rows and installed tuning entries to a file (no reference
  /// counterpart; the reference has no persistence).
  void save(const char* path);

  /// Replaces the rows and tuning entries by those of a file written by
  /// save() (same dimension); this object's device and precision apply.
  void load(const char* path);

  /// Counters of the last search.
  rplsh_search_stats lastStats() const;

  /// The C-ABI handle.
  rplsh_bf* handle() { return handle_; }

 protected:
  const IndexBFConfig config_;
  rplsh_bf* handle_ = nullptr;
};

}  // namespace curplsh
