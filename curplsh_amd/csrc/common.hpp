// Shared host-side helpers of libcurplsh_amd: status/error plumbing and HIP
// call checking. Every C-ABI entry point converts failures into an
// rplsh_status plus a thread-local message instead of aborting (the reference
// aborts: include/internal/assertions.h:7-14, cuda_utils.h:991-1006); the C++
// Index classes turn a failure back into the reference's abort.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "rplsh.h"

namespace rplsh {

// Error carried from deep inside the library to the C-ABI edge.
struct Error : std::runtime_error {
  rplsh_status status;
  Error(rplsh_status s, const std::string& msg) : std::runtime_error(msg), status(s) {}
};

[[noreturn]] void fail(rplsh_status s, const char* fmt, ...);

void set_last_error(const std::string& msg);
void clear_last_error();

#define RPLSH_HIP(call)                                                              \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      ::rplsh::fail(e_ == hipErrorOutOfMemory ? RPLSH_ERR_OUT_OF_MEMORY : RPLSH_ERR_HIP, \
                    "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__,   \
                    __LINE__);                                                       \
    }                                                                                \
  } while (0)

#define RPLSH_REQUIRE(cond, ...)                                    \
  do {                                                              \
    if (!(cond)) ::rplsh::fail(RPLSH_ERR_INVALID_ARG, __VA_ARGS__); \
  } while (0)

// Wraps the body of a C-ABI function: exceptions become status codes.
template <typename F>
rplsh_status guard(F&& f) {
  try {
    clear_last_error();
    f();
    return RPLSH_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.status;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return RPLSH_ERR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return RPLSH_ERR_STATE;
  }
}

inline int64_t div_up(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline int64_t round_up(int64_t a, int64_t b) { return div_up(a, b) * b; }

// Current device, and an RAII scope that switches to another and back
// (DeviceScope, include/device_resources.h:13-32).
int current_device();
struct DeviceScope {
  int prev = -1;
  explicit DeviceScope(int device);
  ~DeviceScope();
  DeviceScope(const DeviceScope&) = delete;
  DeviceScope& operator=(const DeviceScope&) = delete;
};

// Device of a pointer, -1 for (unregistered) host memory.
int pointer_device(const void* p);

}  // namespace rplsh
