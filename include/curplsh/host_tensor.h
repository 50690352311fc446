// HostTensor (reference: include/host_tensor.h:7-114): a Tensor whose
// elements live in host memory, owned (new[] / delete[]) or aliased.
#pragma once

#include <memory>

#include "tensor.h"

namespace curplsh {

template <typename T, int Dim, typename IndexT = int,
          template <typename U> class PtrTraits = DefaultPtrTraits>
class HostTensor : public Tensor<T, Dim, IndexT, PtrTraits> {
 public:
  typedef Tensor<T, Dim, IndexT, PtrTraits> Base;
  typedef IndexT IndexType;
  typedef typename PtrTraits<T>::PtrType DataPtrType;

  HostTensor() : Base() {}
  ~HostTensor() { release(); }
  HostTensor(HostTensor&& o) noexcept : Base() { *this = std::move(o); }
  HostTensor& operator=(HostTensor&& o) noexcept {
    if (this != &o) {
      release();
      Base::operator=(std::move(o));
      own_ = o.own_;
      o.own_ = false;
    }
    return *this;
  }
  HostTensor(const HostTensor&) = delete;
  HostTensor& operator=(const HostTensor&) = delete;

  /// Allocates the given sizes on the host.
  explicit HostTensor(const IndexType sizes[Dim]) : Base(nullptr, sizes) { allocate(); }
  HostTensor(std::initializer_list<IndexType> sizes) : Base(nullptr, sizes) { allocate(); }
  /// Aliases host memory.
  HostTensor(DataPtrType data, const IndexType sizes[Dim]) : Base(data, sizes) {}
  HostTensor(DataPtrType data, std::initializer_list<IndexType> sizes) : Base(data, sizes) {}
  HostTensor(DataPtrType data, const IndexType sizes[Dim], const IndexType strides[Dim])
      : Base(data, sizes, strides) {}
  /// Copy of a (host or device) tensor into new host storage.
  explicit HostTensor(const Base& other) : Base(nullptr, other.sizes()) {
    allocate();
    this->copyFrom(other);
  }

  bool owner() const { return own_; }

 private:
  void allocate() {
    this->data_ = new T[this->getNumElements() ? this->getNumElements() : 1];
    own_ = true;
  }
  void release() {
    if (own_) delete[] this->data_;
    if (own_) this->data_ = nullptr;
    own_ = false;
  }
  bool own_ = false;
};

}  // namespace curplsh
