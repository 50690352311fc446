// TEXMEX datasets (reference: include/dataset.h:14-204) over the C ABI.
// Same names and shapes; deliberate differences: a loaded DatasetIrisa owns and
// frees its arrays (the reference never sets isMemOwner), evaluate() does not
// reorder the caller's arrays (dataset.h:137 sorts them in place), and
// saveXvecsData (a TODO there) exists.
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "memory_space.h"

namespace curplsh {

namespace helper {

/// Loads a .fvecs (T = float) or .ivecs (T = int) file into `space`
/// (loadXvecsData, dataset.h:14-44): num records of dim elements.
template <typename T, typename IndexT>
void loadXvecsData(const char* filename, T*& data, IndexT& num, IndexT& dim, MemorySpace space,
                   int device = 0) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 1, "float / int32 / uint8 elements");
  int64_t n = 0;
  int d = 0;
  checkStatus(rplsh_xvecs_info(filename, (int)sizeof(T), &n, &d), "loadXvecsData");
  num = (IndexT)n;
  dim = (IndexT)d;
  allocMemory(data, (size_t)n * d, space, device);
  checkStatus(rplsh_xvecs_read(filename, (int)sizeof(T), 0, n, data, space == MemorySpace::Device ? 1 : 0,
                               device),
              "loadXvecsData");
}

/// Writes num host records of dim elements (the reference's TODO saveXvecsData).
template <typename T, typename IndexT>
void saveXvecsData(const char* filename, const T* data, IndexT num, IndexT dim, bool append = false) {
  checkStatus(rplsh_xvecs_write(filename, (int)sizeof(T), data, (int64_t)num, (int)dim, append ? 1 : 0),
              "saveXvecsData");
}

}  // namespace helper

template <typename T, typename IndexT>
class Dataset {
 public:
  typedef T DataType;
  typedef IndexT IndexType;

  Dataset() : Dataset(0, 0, 0, 0, nullptr, nullptr, nullptr, "", MemorySpace::Device) {}

  /// Non-owning view of existing arrays (dataset.h:57-68).
  Dataset(IndexT dimension, IndexT gtK, IndexT numBase, IndexT numQuery, T* base, T* query, IndexT* gt,
          const char* name, MemorySpace space)
      : dimension_(dimension), gtK_(gtK), numBase_(numBase), numQuery_(numQuery), base_(base),
        query_(query), gt_(gt), name_(name), space_(space), isMemOwner(false) {}

  Dataset(const Dataset& other) { *this = other; }
  Dataset(Dataset&& other) noexcept { *this = std::move(other); }

  /// Copies are views (dataset.h:74-89).
  Dataset& operator=(const Dataset& other) {
    if (this == &other) return *this;
    release();
    dimension_ = other.dimension_;
    gtK_ = other.gtK_;
    numBase_ = other.numBase_;
    numQuery_ = other.numQuery_;
    base_ = other.base_;
    query_ = other.query_;
    gt_ = other.gt_;
    name_ = other.name_;
    space_ = other.space_;
    isMemOwner = false;
    return *this;
  }

  Dataset& operator=(Dataset&& other) noexcept {
    if (this == &other) return *this;
    release();
    dimension_ = other.dimension_;
    gtK_ = other.gtK_;
    numBase_ = other.numBase_;
    numQuery_ = other.numQuery_;
    base_ = other.base_;
    query_ = other.query_;
    gt_ = other.gt_;
    name_ = std::move(other.name_);
    space_ = other.space_;
    isMemOwner = other.isMemOwner;
    other.isMemOwner = false;
    other.dimension_ = other.numBase_ = other.numQuery_ = 0;
    other.base_ = nullptr;
    other.query_ = nullptr;
    other.gt_ = nullptr;
    return *this;
  }

  virtual ~Dataset() { release(); }

  IndexT getDimension() const { return dimension_; }
  IndexT getGroundTruthK() const { return gtK_; }
  IndexT getNumBase() const { return numBase_; }
  IndexT getNumQuery() const { return numQuery_; }
  const T* getBase() const { return base_; }
  const T* getQuery() const { return query_; }
  const IndexT* getGroundTruth() const { return gt_; }
  std::string getName() const { return name_; }
  MemorySpace getMemorySpace() const { return space_; }

  /// Recall of `indices` ([numQuery][gtK], in this dataset's memory space)
  /// against the ground truth (dataset.h:135-152); diff[i] = gtK - hits (host,
  /// optional). The arrays are not modified.
  virtual float evaluate(const IndexT* indices, IndexT* diff = nullptr) const {
    static_assert(sizeof(IndexT) == 4, "int32 indices");
    const size_t cnt = (size_t)numQuery_ * gtK_;
    std::vector<int32_t> r(cnt), g(cnt), miss((size_t)numQuery_);
    copyMemory(reinterpret_cast<IndexT*>(r.data()), indices, cnt);
    copyMemory(reinterpret_cast<IndexT*>(g.data()), gt_, cnt);
    streamSynchronize(nullptr);
    double recall = 0.0;
    checkStatus(rplsh_recall_at(r.data(), numQuery_, (int)gtK_, g.data(), (int)gtK_, (int)gtK_, miss.data(),
                                &recall),
                "evaluate");
    if (diff)
      for (IndexT i = 0; i < numQuery_; ++i) diff[i] = (IndexT)miss[i];
    return (float)recall;
  }

 protected:
  void release() {
    if (isMemOwner) {
      if (base_) freeMemory(base_);
      if (query_) freeMemory(query_);
      if (gt_) freeMemory(gt_);
    }
    isMemOwner = false;
  }

  IndexT dimension_ = 0;
  IndexT gtK_ = 0;
  IndexT numBase_ = 0;
  IndexT numQuery_ = 0;
  T* base_ = nullptr;
  T* query_ = nullptr;
  IndexT* gt_ = nullptr;
  std::string name_;
  MemorySpace space_ = MemorySpace::Device;
  bool isMemOwner = false;
};

/// `<datadir>/<name>_base.fvecs`, `_query.fvecs`, `_groundtruth.ivecs`
/// (DatasetIrisa, dataset.h:162-190). Owns what it loads.
template <typename IndexT>
class DatasetIrisa : public Dataset<float, IndexT> {
 public:
  DatasetIrisa(const char* name, const char* datadir, MemorySpace space = MemorySpace::Device) {
    this->name_ = name;
    this->space_ = space;
    std::string dir(datadir);
    if (dir.empty() || dir.back() != '/') dir += '/';
    IndexT dimQ = 0, nqGt = 0;
    helper::loadXvecsData<float, IndexT>((dir + name + "_base.fvecs").c_str(), this->base_, this->numBase_,
                                         this->dimension_, space);
    this->isMemOwner = true;
    helper::loadXvecsData<float, IndexT>((dir + name + "_query.fvecs").c_str(), this->query_,
                                         this->numQuery_, dimQ, space);
    helper::loadXvecsData<IndexT, IndexT>((dir + name + "_groundtruth.ivecs").c_str(), this->gt_, nqGt,
                                          this->gtK_, space);
    host_assert(dimQ == this->dimension_ && nqGt == this->numQuery_);
  }
};

class DatasetSIFT : public DatasetIrisa<int> {
 public:
  explicit DatasetSIFT(const char* basedir, MemorySpace space = MemorySpace::Device)
      : DatasetIrisa<int>("sift", basedir, space) {}
};

class DatasetGIST : public DatasetIrisa<int> {
 public:
  explicit DatasetGIST(const char* basedir, MemorySpace space = MemorySpace::Device)
      : DatasetIrisa<int>("gist", basedir, space) {}
};

}  // namespace curplsh
