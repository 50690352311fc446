// Memory spaces and raw allocation (reference: include/memory_space.h:11-42),
// over the C ABI: hipMalloc / hipMallocManaged and stream-ordered copies.
#pragma once

#include <cstddef>

#include "internal/assertions.h"
#include "rplsh.h"

namespace curplsh {

/// An opaque HIP stream (hipStream_t); nullptr is the null stream.
typedef void* Stream;

enum class MemorySpace {
  /// Manually managed device memory (hipMalloc / hipFree)
  Device = 0,
  /// Unified memory (hipMallocManaged / hipFree)
  Unified = 1,
};

/// Allocates num elements of T on `device` (allocMemory, memory_space.h:26-37).
template <typename T>
inline void allocMemory(T*& ptr, size_t num, MemorySpace space, int device = 0) {
  void* p = nullptr;
  checkStatus(rplsh_malloc(device, num * sizeof(T), (int)space, &p), "allocMemory");
  ptr = static_cast<T*>(p);
}

inline void freeMemory(void* ptr) { checkStatus(rplsh_free(ptr), "freeMemory"); }

/// Stream-ordered copy in any direction (copyMemory, memory_space.h:39-41).
template <typename T>
inline void copyMemory(T* dst, const T* src, size_t num, Stream stream = nullptr) {
  checkStatus(rplsh_memcpy_async(dst, src, num * sizeof(T), stream), "copyMemory");
}

/// Device of a pointer, -1 for host memory (getDeviceForAddress).
inline int getDeviceForAddress(const void* p) {
  int d = -1;
  checkStatus(rplsh_pointer_device(p, &d), "getDeviceForAddress");
  return d;
}

inline int getNumDevices() {
  int n = 0;
  checkStatus(rplsh_device_count(&n), "getNumDevices");
  return n;
}

inline int getCurrentDevice() {
  int d = 0;
  checkStatus(rplsh_get_device(&d), "getCurrentDevice");
  return d;
}

inline void setCurrentDevice(int device) {
  checkStatus(rplsh_set_device(device), "setCurrentDevice");
}

inline void streamSynchronize(Stream s) { checkStatus(rplsh_stream_sync(s), "streamSynchronize"); }

}  // namespace curplsh
