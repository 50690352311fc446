// Strided tensor views and owning / aliasing device tensors (reference:
// include/tensor.h:16-449, include/device_tensor.h:8-186), host-side, over
// the C ABI. The interface is the reference's: Tensor<T, Dim, IndexT,
// PtrTraits> with sizes and strides, operator[] slices down to a scalar
// reference, narrow / view / cast / isCastable, copies between tensors or to
// a pointer; DeviceTensor adds storage that it owns (allocated in a
// MemorySpace on the current device) or aliases (caller memory, never freed).
// Element addressing multiplies in 64 bits, so tensors past 2^31 elements
// index correctly whatever IndexT is, and copies go from source to
// destination (the reference passes them to cudaMemcpy swapped,
// tensor.h:133-138,162-167). Element access through operator[] dereferences
// the address on the host: it is for host or unified (managed) memory, as in
// the reference's drivers.
#pragma once

#include <cstddef>
#include <cstdint>
#include <initializer_list>
#include <utility>

#include "memory_space.h"

namespace curplsh {

/// Pointer type of a tensor's data (traits.h:11-19).
template <typename T>
struct DefaultPtrTraits {
  typedef T* PtrType;
};
template <typename T>
struct RestrictPtrTraits {
  typedef T* __restrict__ PtrType;
};

template <typename TensorType, int SubDim, template <typename U> class PtrTraits>
class TensorSlice;

template <typename T, int Dim, typename IndexT = int,
          template <typename U> class PtrTraits = DefaultPtrTraits>
class Tensor {
  static_assert(Dim > 0, "must have at least 1 dimension");

 public:
  typedef T DataType;
  typedef IndexT IndexType;
  typedef typename PtrTraits<T>::PtrType DataPtrType;
  typedef Tensor<T, Dim, IndexT, PtrTraits> TensorType;

  Tensor() : data_(nullptr) {
    for (int i = 0; i < Dim; ++i) {
      sizes_[i] = 0;
      strides_[i] = 1;
    }
  }
  Tensor(const Tensor&) = default;
  Tensor& operator=(const Tensor&) = default;
  /// A moved-from tensor is empty (no data, zero sizes).
  Tensor(Tensor&& o) noexcept { *this = std::move(o); }
  Tensor& operator=(Tensor&& o) noexcept {
    data_ = o.data_;
    o.data_ = nullptr;
    for (int i = 0; i < Dim; ++i) {
      sizes_[i] = o.sizes_[i];
      strides_[i] = o.strides_[i];
      o.sizes_[i] = 0;
      o.strides_[i] = 0;
    }
    return *this;
  }

  /// Contiguous row-major layout of the given sizes.
  Tensor(DataPtrType data, const IndexType sizes[Dim]) : data_(data) {
    for (int i = 0; i < Dim; ++i) sizes_[i] = sizes[i];
    setContiguousStrides();
  }
  Tensor(DataPtrType data, std::initializer_list<IndexType> sizes) : data_(data) {
    host_assert((int)sizes.size() == Dim);
    int i = 0;
    for (IndexType s : sizes) sizes_[i++] = s;
    setContiguousStrides();
  }
  /// Explicit sizes and strides (elements).
  Tensor(DataPtrType data, const IndexType sizes[Dim], const IndexType strides[Dim]) : data_(data) {
    for (int i = 0; i < Dim; ++i) {
      sizes_[i] = sizes[i];
      strides_[i] = strides[i];
    }
  }

  /// src -> dst, both contiguous with the same element count; synchronous.
  static void copy(const TensorType& src, TensorType& dst) {
    host_assert(src.isContiguous() && dst.isContiguous());
    host_assert(src.getNumElements() == dst.getNumElements());
    if (dst.getNumElements() == 0) return;
    host_assert(src.data() && dst.data());
    copyMemory(dst.data(), src.data(), src.getNumElements());
    streamSynchronize(nullptr);
  }
  void copyFrom(const TensorType& src) { copy(src, *this); }
  void copyTo(TensorType& dst) const { copy(*this, dst); }
  /// All elements to `dst` (host or device memory); synchronous.
  void copyTo(T* dst) const {
    host_assert(isContiguous());
    if (dst == data_ || getNumElements() == 0) return;
    copyMemory(dst, static_cast<const T*>(data_), getNumElements());
    streamSynchronize(nullptr);
  }
  /// Stream-ordered versions (the caller synchronises `stream`).
  void copyTo(T* dst, Stream stream) const {
    host_assert(isContiguous());
    if (getNumElements()) copyMemory(dst, static_cast<const T*>(data_), getNumElements(), stream);
  }
  void copyFrom(const T* src, Stream stream = nullptr) {
    host_assert(isContiguous());
    if (getNumElements()) copyMemory(static_cast<T*>(data_), src, getNumElements(), stream);
  }

  template <typename OtherT, int OtherDim>
  bool isSame(const Tensor<OtherT, OtherDim, IndexT, PtrTraits>& o) const {
    if (Dim != OtherDim) return false;
    for (int i = 0; i < Dim; ++i)
      if (getSize(i) != o.getSize(i) || getStride(i) != o.getStride(i)) return false;
    return true;
  }
  template <typename OtherT, int OtherDim>
  bool isSameSizes(const Tensor<OtherT, OtherDim, IndexT, PtrTraits>& o) const {
    if (Dim != OtherDim) return false;
    for (int i = 0; i < Dim; ++i)
      if (getSize(i) != o.getSize(i)) return false;
    return true;
  }

  const IndexType* sizes() const { return sizes_; }
  const IndexType* strides() const { return strides_; }
  /// Size / stride of dimension i; negative i counts from the innermost.
  IndexType getSize(int i) const { return i >= 0 ? sizes_[i] : sizes_[Dim + i]; }
  IndexType getStride(int i) const { return i >= 0 ? strides_[i] : strides_[Dim + i]; }
  size_t getNumElements() const {
    size_t n = 1;
    for (int i = 0; i < Dim; ++i) n *= (size_t)sizes_[i];
    return n;
  }
  /// Bytes of the elements (assumes a contiguous tensor).
  size_t getDataMemSize() const { return getNumElements() * sizeof(T); }

  /// Slice along the outermost dimension; scalars at depth Dim.
  TensorSlice<TensorType, Dim - 1, PtrTraits> operator[](IndexType i) {
    return TensorSlice<TensorType, Dim, PtrTraits>(*this)[i];
  }
  const TensorSlice<TensorType, Dim - 1, PtrTraits> operator[](IndexType i) const {
    return TensorSlice<TensorType, Dim, PtrTraits>(const_cast<TensorType&>(*this))[i];
  }

  DataPtrType data() { return data_; }
  const DataPtrType data() const { return data_; }
  DataPtrType end() { return data_ + getNumElements(); }
  const DataPtrType end() const { return data_ + getNumElements(); }
  template <typename U>
  typename PtrTraits<U>::PtrType dataAs() {
    return reinterpret_cast<typename PtrTraits<U>::PtrType>(data_);
  }
  template <typename U>
  typename PtrTraits<const U>::PtrType dataAs() const {
    return reinterpret_cast<typename PtrTraits<const U>::PtrType>(data_);
  }

  /// A view as U elements (sizeof(U) a multiple of sizeof(T)) is possible:
  /// aligned data, innermost dimension contiguous and a whole number of U,
  /// outer strides whole numbers of U.
  template <typename U>
  bool isCastable() const {
    static_assert(sizeof(U) >= sizeof(T), "cast to a type at least as large");
    if (sizeof(U) == sizeof(T)) return true;
    if (sizeof(U) % sizeof(T) != 0) return false;
    constexpr IndexType kMul = (IndexType)(sizeof(U) / sizeof(T));
    if (reinterpret_cast<uintptr_t>(data_) % alignof(U) != 0) return false;
    if (strides_[Dim - 1] != 1 || sizes_[Dim - 1] % kMul != 0) return false;
    for (int i = 0; i < Dim - 1; ++i)
      if (strides_[i] % kMul != 0) return false;
    return true;
  }
  template <typename U>
  Tensor<U, Dim, IndexT, PtrTraits> cast() {
    host_assert(isCastable<U>());
    constexpr IndexType kMul = (IndexType)(sizeof(U) / sizeof(T));
    IndexType s[Dim], st[Dim];
    for (int i = 0; i < Dim; ++i) {
      s[i] = sizes_[i];
      st[i] = strides_[i] / kMul;
    }
    s[Dim - 1] = sizes_[Dim - 1] / kMul;
    st[Dim - 1] = 1;
    return Tensor<U, Dim, IndexT, PtrTraits>(reinterpret_cast<typename PtrTraits<U>::PtrType>(data_), s, st);
  }
  template <typename U>
  const Tensor<U, Dim, IndexT, PtrTraits> cast() const {
    return const_cast<TensorType*>(this)->template cast<U>();
  }

  /// True if the elements fill one dense row-major block.
  bool isContiguous() const {
    int64_t expect = 1;
    for (int i = Dim - 1; i >= 0; --i) {
      if (sizes_[i] != 1 && (int64_t)strides_[i] != expect) return false;
      expect *= (int64_t)sizes_[i];
    }
    return true;
  }
  /// True if dimension i holds no padding relative to dimension i + 1.
  bool isContiguous(int i) const {
    return i == Dim - 1 ? strides_[i] == 1
                        : (int64_t)strides_[i] == (int64_t)strides_[i + 1] * (int64_t)sizes_[i + 1];
  }

  /// Elements [start, start + size) of dimension `dim`, same strides
  /// (tensor.h:381-403).
  TensorType narrow(int dim, IndexType start, IndexType size) const {
    host_assert(dim >= 0 && dim < Dim);
    host_assert(start >= 0 && size >= 0 && start + size <= sizes_[dim]);
    IndexType s[Dim];
    for (int i = 0; i < Dim; ++i) s[i] = i == dim ? size : sizes_[i];
    return TensorType(data_ + (int64_t)start * (int64_t)strides_[dim], s, strides_);
  }
  /// The outermost dimension narrowed (narrow(0, start, size)).
  TensorType narrowOutermost(IndexType start, IndexType size) const { return narrow(0, start, size); }

  /// The same elements with other sizes (contiguous tensors only).
  template <int NewDim>
  Tensor<T, NewDim, IndexT, PtrTraits> view(std::initializer_list<IndexType> sizes) const {
    host_assert(isContiguous());
    host_assert((int)sizes.size() == NewDim);
    size_t n = 1;
    for (IndexType s : sizes) n *= (size_t)s;
    host_assert(n == getNumElements());
    return Tensor<T, NewDim, IndexT, PtrTraits>(data_, sizes);
  }

 protected:
  void setContiguousStrides() {
    strides_[Dim - 1] = 1;
    for (int i = Dim - 2; i >= 0; --i) strides_[i] = strides_[i + 1] * sizes_[i + 1];
  }

  DataPtrType data_;
  IndexType sizes_[Dim];
  IndexType strides_[Dim];
};

/// A sub-tensor of the last SubDim dimensions at a fixed position of the
/// others (a pointer plus the tensor's strides); SubDim == 0 is one element.
template <typename TensorType, int SubDim, template <typename U> class PtrTraits>
class TensorSlice {
 public:
  typedef typename TensorType::DataType DataType;
  typedef typename TensorType::IndexType IndexType;
  typedef typename TensorType::DataPtrType DataPtrType;

  explicit TensorSlice(TensorType& t) : tensor_(t), data_(t.data()) {}
  TensorSlice(TensorType& t, DataPtrType data) : tensor_(t), data_(data) {}

  TensorSlice<TensorType, SubDim - 1, PtrTraits> operator[](IndexType i) {
    return TensorSlice<TensorType, SubDim - 1, PtrTraits>(
        tensor_, data_ + (int64_t)i * (int64_t)tensor_.getStride(-SubDim));
  }
  const TensorSlice<TensorType, SubDim - 1, PtrTraits> operator[](IndexType i) const {
    return TensorSlice<TensorType, SubDim - 1, PtrTraits>(
        tensor_, data_ + (int64_t)i * (int64_t)tensor_.getStride(-SubDim));
  }
  DataType* operator&() { return data_; }
  const DataType* operator&() const { return data_; }
  DataPtrType data() { return data_; }
  const DataPtrType data() const { return data_; }
  template <typename U>
  typename PtrTraits<U>::PtrType dataAs() {
    return reinterpret_cast<typename PtrTraits<U>::PtrType>(data_);
  }
  template <typename U>
  typename PtrTraits<const U>::PtrType dataAs() const {
    return reinterpret_cast<typename PtrTraits<const U>::PtrType>(data_);
  }

 protected:
  TensorType& tensor_;
  DataPtrType data_;
};

template <typename TensorType, template <typename U> class PtrTraits>
class TensorSlice<TensorType, 0, PtrTraits> {
 public:
  typedef typename TensorType::DataType DataType;
  typedef typename TensorType::DataPtrType DataPtrType;

  explicit TensorSlice(TensorType& t) : tensor_(t), data_(t.data()) {}
  TensorSlice(TensorType& t, DataPtrType data) : tensor_(t), data_(data) {}

  TensorSlice& operator=(DataType v) {
    *data_ = v;
    return *this;
  }
  operator DataType&() { return *data_; }
  operator const DataType&() const { return *data_; }
  DataType* operator&() { return data_; }
  const DataType* operator&() const { return data_; }
  DataPtrType data() { return data_; }
  const DataPtrType data() const { return data_; }
  template <typename U>
  typename PtrTraits<U>::PtrType dataAs() {
    return reinterpret_cast<typename PtrTraits<U>::PtrType>(data_);
  }
  template <typename U>
  typename PtrTraits<const U>::PtrType dataAs() const {
    return reinterpret_cast<typename PtrTraits<const U>::PtrType>(data_);
  }

 protected:
  TensorType& tensor_;
  DataPtrType data_;
};

/// Tensor with storage: owned (allocated in a MemorySpace on the current
/// device, or on `device` when given, freed on destruction) or aliased
/// (caller memory, never freed).
template <typename T, int Dim, typename IndexT = int,
          template <typename U> class PtrTraits = DefaultPtrTraits>
class DeviceTensor : public Tensor<T, Dim, IndexT, PtrTraits> {
 public:
  typedef Tensor<T, Dim, IndexT, PtrTraits> Base;
  typedef T DataType;
  typedef IndexT IndexType;
  typedef typename PtrTraits<T>::PtrType DataPtrType;

  DeviceTensor() : Base(), own_(false), space_(MemorySpace::Device) {}
  ~DeviceTensor() { release(); }
  DeviceTensor(DeviceTensor&& o) noexcept : Base(), own_(false), space_(MemorySpace::Device) {
    *this = std::move(o);
  }
  DeviceTensor& operator=(DeviceTensor&& o) noexcept {
    if (this != &o) {
      release();
      Base::operator=(std::move(o));
      own_ = o.own_;
      space_ = o.space_;
      o.own_ = false;
    }
    return *this;
  }
  DeviceTensor(const DeviceTensor&) = delete;
  DeviceTensor& operator=(const DeviceTensor&) = delete;

  /// Allocates the given sizes in `space` (device_tensor.h:56-72).
  DeviceTensor(const IndexType sizes[Dim], MemorySpace space = MemorySpace::Device, int device = -1)
      : Base(nullptr, sizes), own_(false), space_(space) {
    allocate(device);
  }
  DeviceTensor(std::initializer_list<IndexType> sizes, MemorySpace space = MemorySpace::Device,
               int device = -1)
      : Base(nullptr, sizes), own_(false), space_(space) {
    allocate(device);
  }

  /// Aliases `data` (device_tensor.h:73-97).
  DeviceTensor(DataPtrType data, const IndexType sizes[Dim], MemorySpace space)
      : Base(data, sizes), own_(false), space_(space) {}
  DeviceTensor(DataPtrType data, std::initializer_list<IndexType> sizes, MemorySpace space)
      : Base(data, sizes), own_(false), space_(space) {}
  DeviceTensor(DataPtrType data, const IndexType sizes[Dim], const IndexType strides[Dim],
               MemorySpace space)
      : Base(data, sizes, strides), own_(false), space_(space) {}

  /// Aliases `data` if it lives on the current device, else allocates device
  /// memory and copies it in on `stream` (device_tensor.h:99-135).
  DeviceTensor(DataPtrType data, const IndexType sizes[Dim], Stream stream = nullptr)
      : Base(data, sizes), own_(false), space_(MemorySpace::Device) {
    adoptOrCopy(data, stream);
  }
  DeviceTensor(DataPtrType data, std::initializer_list<IndexType> sizes, Stream stream = nullptr)
      : Base(data, sizes), own_(false), space_(MemorySpace::Device) {
    adoptOrCopy(data, stream);
  }
  DeviceTensor(DataPtrType data, const IndexType sizes[Dim], const IndexType strides[Dim],
               Stream stream = nullptr)
      : Base(data, sizes, strides), own_(false), space_(MemorySpace::Device) {
    host_assert(this->isContiguous());
    adoptOrCopy(data, stream);
  }

  /// Deep copy of `other` into new storage in `space` (device_tensor.h:137-146).
  DeviceTensor(const Base& other, MemorySpace space)
      : Base(nullptr, other.sizes()), own_(false), space_(space) {
    host_assert(other.isContiguous());
    allocate(-1);
    this->copyFrom(other);
  }

  /// Device of the data, -1 for host memory.
  int device() const { return getDeviceForAddress(this->data_); }
  MemorySpace space() const { return space_; }
  bool owner() const { return own_; }

  /// Zero fill, ordered on `stream`.
  DeviceTensor& zero(Stream stream = nullptr) {
    if (this->data_ && this->getNumElements()) {
      host_assert(this->isContiguous());
      checkStatus(rplsh_memset_async(this->data_, 0, this->getDataMemSize(), stream), "DeviceTensor::zero");
    }
    return *this;
  }
  /// Elements into host memory `out`, ordered on `stream`.
  void toHost(T* out, Stream stream = nullptr) const {
    host_assert(this->data_ && out);
    this->copyTo(out, stream);
  }

 private:
  void allocate(int device) {
    this->data_ = nullptr;
    if (this->getNumElements() > 0)
      allocMemory(this->data_, this->getNumElements(), space_, device < 0 ? getCurrentDevice() : device);
    own_ = true;
  }
  void adoptOrCopy(DataPtrType data, Stream stream) {
    const int src = getDeviceForAddress(data);
    if (src != getCurrentDevice()) {
      allocate(-1);
      if (this->getNumElements()) copyMemory(this->data_, static_cast<const T*>(data), this->getNumElements(), stream);
    }
  }
  void release() {
    if (own_ && this->data_) freeMemory(this->data_);
    if (own_) this->data_ = nullptr;
    own_ = false;
  }

  bool own_;
  MemorySpace space_;
};

}  // namespace curplsh
