// IndexRPLSH: approximate L2 kNN with a random-projection LSH front end and
// the exact re-rank of IndexBF's plain path behind it.
//
// The reference is named for this step but does not implement it: its only
// piece is the Gaussian projection initialiser (matrixRandomInitialize,
// src/internal/matrix_ops.cu:71-76). The class follows the reference's Index
// interface (include/index.h:18-82) like IndexBF does: same constructor
// shape, train / add / search / reset, host or device pointers.
//   * projections: tables x bits standard-normal rows from the seed;
//   * a row's code in a table: one bit per projection, set iff the canonical
//     FP32 dot with the row is >= the projection's offset (its median over
//     the first add, or 0 with center = false);
//   * a query probes its own bucket in every table and, per extra probe, the
//     bucket with its least-certain bit flipped; the first maxCandidates
//     bucket entries (in probe, table order) are de-duplicated and re-ranked
//     exactly; results are exact squared L2 distances of those candidates,
//     ascending, ties by lower index, padded with -1 / FLT_MAX.
#pragma once

#include <cstdint>

#include "index.h"

namespace curplsh {

struct IndexRPLSHConfig : public IndexConfig {
  IndexRPLSHConfig()
      : numTables(16),
        numBits(16),
        numProbes(0),
        maxCandidates(1024),
        seed(1234),
        center(true),
        prerankFactor(0) {}
  /// Hash tables.
  int numTables;
  /// Code bits per table (1..24).
  int numBits;
  /// Extra buckets probed per table (0..numBits).
  int numProbes;
  /// Bucket entries taken per query before de-duplication (<= 8192; above
  /// 4096 only with prerankFactor > 0 and prerankFactor * k <= 4096).
  int maxCandidates;
  /// Seed of the projection matrix.
  unsigned long long seed;
  /// Hyperplanes through the projections' medians over the first add (L2
  /// proximity) rather than through the origin (angle).
  bool center;
  /// 0: exact re-rank of every distinct candidate; f > 0: a BF16 pre-rank keeps
  /// the f * k nearest of them for the exact re-rank (BF16 distance as defined at
  /// rplsh_lsh_config::prerank; 0..64).
  int prerankFactor;
};

class IndexRPLSH : public Index {
 public:
  IndexRPLSH(DeviceResources* resources, int dim, IndexRPLSHConfig config = IndexRPLSHConfig());
  ~IndexRPLSH() override;

  /// No training: the projections depend on the seed only.
  void train(int num, const float* data) override;
  /// Appends vectors and rebuilds the hash tables.
  void add(int num, const float* data) override;
  void search(int num, const float* queries, int k, int* indices, float* distances) override;
  void reset() override;

  /// Counters of the last search.
  rplsh_lsh_stats lastStats() const;
  /// The C-ABI handle.
  rplsh_lsh* handle() { return handle_; }

 protected:
  const IndexRPLSHConfig config_;
  rplsh_lsh* handle_ = nullptr;
};

}  // namespace curplsh
