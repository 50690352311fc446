// C++ host layer of the drop-in: curplsh::DeviceResources, Index, IndexBF and
// the internal free functions, compiled by g++ and calling only the C ABI of
// libcurplsh_amd.so (include/rplsh.h). No HIP header is included here.
#include <cstring>

#include "curplsh/device_resources.h"
#include "curplsh/index.h"
#include "curplsh/index_bf.h"
#include "curplsh/index_rplsh.h"
#include "curplsh/internal/assertions.h"
#include "curplsh/internal/broadcast.h"
#include "curplsh/internal/norm.h"
#include "curplsh/internal/select.h"
#include "curplsh/internal/search.h"

namespace curplsh {

// ------------------------------------------------------- DeviceResources

DeviceResources::DeviceResources() {
  checkStatus(rplsh_resources_create(&handle_), "DeviceResources");
}

DeviceResources::DeviceResources(int device) : DeviceResources() { initializeForDevice(device); }

DeviceResources::DeviceResources(std::initializer_list<int> devices) : DeviceResources() {
  for (int d : devices) initializeForDevice(d);
}

DeviceResources::~DeviceResources() {
  if (handle_) rplsh_resources_destroy(handle_);
}

void DeviceResources::initializeForDevice(int device) {
  host_assert(device < getNumDevices());
  checkStatus(rplsh_resources_init_device(handle_, device), "initializeForDevice");
}

void DeviceResources::setDefaultStream(int device, Stream stream) {
  checkStatus(rplsh_resources_set_default_stream(handle_, device, stream), "setDefaultStream");
}

Stream DeviceResources::getDefaultStream(int device) {
  void* s = nullptr;
  checkStatus(rplsh_resources_default_stream(handle_, device, &s), "getDefaultStream");
  return s;
}

std::vector<Stream> DeviceResources::getAlternateStreams(int device) {
  void* s[8];
  int n = 8;
  checkStatus(rplsh_resources_alternate_streams(handle_, device, s, &n), "getAlternateStreams");
  return std::vector<Stream>(s, s + n);
}

Stream DeviceResources::getAsyncCopyStream(int device) {
  void* s = nullptr;
  checkStatus(rplsh_resources_async_copy_stream(handle_, device, &s), "getAsyncCopyStream");
  return s;
}

void DeviceResources::syncDefaultStream(int device) {
  checkStatus(rplsh_resources_sync_default_stream(handle_, device), "syncDefaultStream");
}

// ----------------------------------------------------------------- Index

// Index::Index (index.cpp:8-27): device and dimension preconditions abort.
Index::Index(DeviceResources* resources, int dimension, IndexConfig config)
    : dimension_(dimension),
      resources_(resources),
      memorySpace_(config.memorySpace),
      device_(config.device) {
  host_assert(device_ < getNumDevices());
  host_assert(dimension_ > 0);
  host_assert(memorySpace_ == MemorySpace::Device);
  host_assert(resources_);
  resources_->initializeForDevice(device_);
}

Index::~Index() = default;

// --------------------------------------------------------------- IndexBF

IndexBF::IndexBF(DeviceResources* resources, int dim, IndexBFConfig config)
    : Index(resources, dim, config), config_(config) {
  rplsh_bf_config c;
  rplsh_bf_config_default(&c);
  c.device = config.device;
  c.memory_space = (int)config.memorySpace;
  c.store_transposed = config.storeTransposed;
  c.compute_exact_distances = config.computeExactDistances;
  c.precision = (int)config.precision;
  c.kprime_factor = config.kprimeFactor;
  c.fp64_boundary = config.fp64Boundary;
  c.persistent_max_queries = config.persistentMaxQueries;
  c.residual_clusters = config.residualClusters;
  checkStatus(rplsh_bf_create(resources->handle(), dim, &c, &handle_), "IndexBF");
  isTrained_ = true;
}

IndexBF::~IndexBF() {
  if (handle_) rplsh_bf_destroy(handle_);
}

void IndexBF::save(const char* path) { checkStatus(rplsh_bf_save(handle_, path), "IndexBF::save"); }

void IndexBF::load(const char* path) {
  rplsh_bf* fresh = nullptr;
  checkStatus(rplsh_bf_load(resources_->handle(), path, device_, &fresh), "IndexBF::load");
  int d = 0;
  int64_t n = 0;
  rplsh_bf_dimension(fresh, &d);
  if (d != dimension_) {
    rplsh_bf_destroy(fresh);
    host_assert(d == dimension_);
  }
  checkStatus(rplsh_bf_set_precision(fresh, (int)config_.precision, config_.kprimeFactor), "IndexBF::load");
  rplsh_bf_destroy(handle_);
  handle_ = fresh;
  checkStatus(rplsh_bf_num_data(handle_, &n), "IndexBF::load");
  numData_ = (int)n;
}

// Nothing to learn (the reference's intent, index_bf.cpp:13-15).
void IndexBF::train(int, const float*) {}

void IndexBF::add(int num, const float* data) {
  if (num <= 0) return;
  checkStatus(rplsh_bf_add(handle_, num, data), "IndexBF::add");
  int64_t n = 0;
  checkStatus(rplsh_bf_num_data(handle_, &n), "IndexBF::add");
  numData_ = (int)n;
}

void IndexBF::search(int num, const float* queries, int k, int* indices, float* distances) {
  if (num <= 0) return;
  checkStatus(rplsh_bf_search(handle_, num, queries, k, indices, distances), "IndexBF::search");
}

void IndexBF::reset() {
  checkStatus(rplsh_bf_reset(handle_), "IndexBF::reset");
  numData_ = 0;
}

void IndexBF::setPrecision(Precision p, int kprimeFactor) {
  checkStatus(rplsh_bf_set_precision(handle_, (int)p, kprimeFactor), "IndexBF::setPrecision");
}

rplsh_search_stats IndexBF::lastStats() const {
  rplsh_search_stats s;
  std::memset(&s, 0, sizeof(s));
  checkStatus(rplsh_bf_last_stats(handle_, &s), "IndexBF::lastStats");
  return s;
}

// ------------------------------------------------------------ IndexRPLSH

IndexRPLSH::IndexRPLSH(DeviceResources* resources, int dim, IndexRPLSHConfig config)
    : Index(resources, dim, config), config_(config) {
  rplsh_lsh_config c;
  rplsh_lsh_config_default(&c);
  c.device = config.device;
  c.tables = config.numTables;
  c.bits = config.numBits;
  c.probes = config.numProbes;
  c.max_candidates = config.maxCandidates;
  c.seed = config.seed;
  c.center = config.center ? 1 : 0;
  c.prerank = config.prerankFactor;
  host_assert(config.memorySpace == MemorySpace::Device);
  checkStatus(rplsh_lsh_create(resources->handle(), dim, &c, &handle_), "IndexRPLSH");
  isTrained_ = true;
}

IndexRPLSH::~IndexRPLSH() {
  if (handle_) rplsh_lsh_destroy(handle_);
}

void IndexRPLSH::train(int, const float*) {}

void IndexRPLSH::add(int num, const float* data) {
  if (num <= 0) return;
  checkStatus(rplsh_lsh_add(handle_, num, data), "IndexRPLSH::add");
  int64_t n = 0;
  checkStatus(rplsh_lsh_num_data(handle_, &n), "IndexRPLSH::add");
  numData_ = (int)n;
}

void IndexRPLSH::search(int num, const float* queries, int k, int* indices, float* distances) {
  if (num <= 0) return;
  checkStatus(rplsh_lsh_search(handle_, num, queries, k, indices, distances), "IndexRPLSH::search");
}

void IndexRPLSH::reset() {
  checkStatus(rplsh_lsh_reset(handle_), "IndexRPLSH::reset");
  numData_ = 0;
}

rplsh_lsh_stats IndexRPLSH::lastStats() const {
  rplsh_lsh_stats s;
  std::memset(&s, 0, sizeof(s));
  checkStatus(rplsh_lsh_last_stats(handle_, &s), "IndexRPLSH::lastStats");
  return s;
}

// ------------------------------------------------------ internal functions

void computeL2Norm(const Tensor<float, 2>& vectors, Tensor<float, 1>& norms, bool squared,
                   Stream stream) {
  host_assert(norms.getSize(0) == vectors.getSize(0));
  checkStatus(rplsh_l2norm_dev(vectors.data(), vectors.getSize(0), (int)vectors.getSize(1),
                               norms.data(), squared ? 1 : 0, stream),
              "computeL2Norm");
}

// Shape preconditions of select.cu (row counts and k agree).
void blockSelect(Tensor<float, 2>& in, Tensor<float, 2>& outK, Tensor<int, 2>& outV, int k,
                 bool selectMax, Stream stream) {
  host_assert(outK.getSize(0) == in.getSize(0) && outV.getSize(0) == in.getSize(0));
  host_assert(outK.getSize(1) == k && outV.getSize(1) == k);
  checkStatus(rplsh_block_select_dev(in.data(), nullptr, in.getSize(0), in.getSize(1), k, selectMax ? 1 : 0,
                                     outK.data(), outV.data(), stream),
              "blockSelect");
}

void blockSelect(Tensor<float, 2>& inK, Tensor<int, 2>& inV, Tensor<float, 2>& outK,
                 Tensor<int, 2>& outV, int k, bool selectMax, Stream stream) {
  host_assert(inV.getSize(0) == inK.getSize(0) && inV.getSize(1) == inK.getSize(1));
  host_assert(outK.getSize(0) == inK.getSize(0) && outV.getSize(0) == inK.getSize(0));
  host_assert(outK.getSize(1) == k && outV.getSize(1) == k);
  checkStatus(rplsh_block_select_dev(inK.data(), inV.data(), inK.getSize(0), inK.getSize(1), k,
                                     selectMax ? 1 : 0, outK.data(), outV.data(), stream),
              "blockSelect");
}

void l2Select(Tensor<float, 2>& productDistances, Tensor<float, 1>& baseNorms, Tensor<float, 2>& distances,
              Tensor<int, 2>& indices, int k, Stream stream) {
  host_assert(baseNorms.getSize(0) == productDistances.getSize(1));
  host_assert(distances.getSize(0) == productDistances.getSize(0) && indices.getSize(0) == distances.getSize(0));
  host_assert(distances.getSize(1) == k && indices.getSize(1) == k);
  checkStatus(rplsh_l2_select_dev(productDistances.data(), productDistances.getSize(0), productDistances.getSize(1),
                                  productDistances.getStride(0), baseNorms.data(), k, distances.data(),
                                  indices.data(), stream),
              "l2Select");
}

void broadcastAlongRowsSum(Tensor<float, 1>& input, Tensor<float, 2>& output, Stream stream) {
  host_assert(input.getSize(0) == output.getSize(0));  // broadcast.cu:43
  checkStatus(rplsh_broadcast_rows_sum_dev(input.data(), output.data(), output.getSize(0), output.getSize(1),
                                           stream),
              "broadcastAlongRowsSum");
}

// Shape preconditions of search.cu:86-89.
void searchL2Distance(DeviceResources* resources, const Tensor<float, 2>& bases,
                      Tensor<float, 1>* basesNorm, const Tensor<float, 2>& queries, int k,
                      Tensor<int, 2>& indices, Tensor<float, 2>& distances,
                      bool needRealDistances) {
  host_assert(distances.getSize(0) == queries.getSize(0));
  host_assert(indices.getSize(0) == queries.getSize(0));
  host_assert(distances.getSize(1) == k);
  host_assert(indices.getSize(1) == k);
  host_assert(bases.getSize(0) == 0 || bases.getSize(1) == queries.getSize(1));
  Stream s = resources->getDefaultStreamCurrentDevice();
  checkStatus(rplsh_search_l2_dev(bases.data(), bases.getSize(0), (int)queries.getSize(1),
                                  basesNorm ? basesNorm->data() : nullptr, queries.data(),
                                  queries.getSize(0), k, indices.data(), distances.data(),
                                  needRealDistances ? 1 : 0, s),
              "searchL2Distance");
}

}  // namespace curplsh
