// Index (reference: include/index.h:18-82). Same virtual interface: train,
// add, search, reset and the getters; dimension and device fixed at
// construction; the DeviceResources are not owned and must outlive the index.
#pragma once

#include "device_resources.h"
#include "memory_space.h"

namespace curplsh {

struct IndexConfig {
  inline IndexConfig() : device(0), memorySpace(MemorySpace::Device) {}

  /// GPU device on which the index is resident
  int device;

  /// What memory space to use for primary storage.
  MemorySpace memorySpace;
};

class Index {
 public:
  Index(DeviceResources* resources, int dimension, IndexConfig config);

  virtual ~Index();

  /// Training on representative vectors (size num * dim).
  virtual void train(int num, const float* data) = 0;

  /// Adds num vectors (size num * dim); labels are assigned sequentially.
  virtual void add(int num, const float* data) = 0;

  /// k nearest neighbours of each query, ascending by distance; rows with
  /// fewer than k results are padded with -1 (index.h:42-44).
  virtual void search(int num, const float* queries, int k, int* indices,
                      float* distances) = 0;

  /// Removes all data.
  virtual void reset() = 0;

  inline int getDimension() const { return dimension_; }
  inline int getNumData() const { return numData_; }
  inline DeviceResources* getResources() { return resources_; }
  inline int getDevice() const { return device_; }

 protected:
  const int dimension_;
  int numData_ = 0;
  bool verbose_ = false;
  bool isTrained_ = false;
  DeviceResources* resources_;
  const MemorySpace memorySpace_;
  const int device_;
};

}  // namespace curplsh
