// Row-major tensor views and owning device tensors (reference:
// include/tensor.h:16-424, include/device_tensor.h:8-186), reduced to what the
// kNN path uses: sizes with 64-bit extents, contiguous rows, narrow() along
// dimension 0, owning vs aliasing device storage.
#pragma once

#include <cstdint>
#include <initializer_list>
#include <utility>

#include "memory_space.h"

namespace curplsh {

template <typename T, int Dim>
class Tensor {
 public:
  Tensor() {
    for (int i = 0; i < Dim; ++i) sizes_[i] = 0;
  }
  Tensor(T* data, std::initializer_list<int64_t> sizes) : data_(data) {
    int i = 0;
    for (auto s : sizes) sizes_[i++] = s;
    for (; i < Dim; ++i) sizes_[i] = 1;
  }
  Tensor(T* data, const int64_t sizes[Dim]) : data_(data) {
    for (int i = 0; i < Dim; ++i) sizes_[i] = sizes[i];
  }

  T* data() { return data_; }
  const T* data() const { return data_; }
  int64_t getSize(int i) const { return sizes_[i]; }
  int64_t getStride(int i) const {
    int64_t s = 1;
    for (int j = i + 1; j < Dim; ++j) s *= sizes_[j];
    return s;
  }
  int64_t numElements() const {
    int64_t n = 1;
    for (int i = 0; i < Dim; ++i) n *= sizes_[i];
    return n;
  }
  size_t getDataMemSize() const { return (size_t)numElements() * sizeof(T); }

  /// Rows [start, start + size) along dimension 0 (tensor.h:381-403).
  Tensor<T, Dim> narrowOutermost(int64_t start, int64_t size) const {
    int64_t s[Dim];
    for (int i = 0; i < Dim; ++i) s[i] = sizes_[i];
    s[0] = size;
    return Tensor<T, Dim>(data_ + start * getStride(0), s);
  }

 protected:
  T* data_ = nullptr;
  int64_t sizes_[Dim];
};

/// Owning (allocated here) or aliasing (wrapping a caller pointer) storage.
template <typename T, int Dim>
class DeviceTensor : public Tensor<T, Dim> {
 public:
  DeviceTensor() = default;
  /// Allocates sizes on `device` in `space`.
  DeviceTensor(std::initializer_list<int64_t> sizes, MemorySpace space = MemorySpace::Device,
               int device = 0)
      : Tensor<T, Dim>(nullptr, sizes), own_(true) {
    if (this->numElements() > 0) allocMemory(this->data_, (size_t)this->numElements(), space, device);
  }
  /// Aliases caller memory; never frees it.
  DeviceTensor(T* data, std::initializer_list<int64_t> sizes) : Tensor<T, Dim>(data, sizes) {}
  ~DeviceTensor() { release(); }
  DeviceTensor(DeviceTensor&& o) noexcept { *this = std::move(o); }
  DeviceTensor& operator=(DeviceTensor&& o) noexcept {
    if (this != &o) {
      release();
      this->data_ = o.data_;
      for (int i = 0; i < Dim; ++i) this->sizes_[i] = o.sizes_[i];
      own_ = o.own_;
      o.data_ = nullptr;
      o.own_ = false;
    }
    return *this;
  }
  DeviceTensor(const DeviceTensor&) = delete;
  DeviceTensor& operator=(const DeviceTensor&) = delete;

  /// Host (or device) -> this copy, stream-ordered.
  void copyFrom(const T* src, Stream stream = nullptr) {
    copyMemory(this->data_, src, (size_t)this->numElements(), stream);
  }
  /// this -> host (or device) copy, stream-ordered.
  void copyTo(T* dst, Stream stream = nullptr) const {
    copyMemory(dst, this->data_, (size_t)this->numElements(), stream);
  }
  bool owner() const { return own_; }

 private:
  void release() {
    if (own_ && this->data_) freeMemory(this->data_);
    this->data_ = nullptr;
    own_ = false;
  }
  bool own_ = false;
};

}  // namespace curplsh
