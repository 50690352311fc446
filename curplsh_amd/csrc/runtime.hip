// Runtime layer of libcurplsh_amd: error plumbing, devices, streams
// (DeviceResources, include/device_resources.h:60-263), memory
// (include/memory_space.h:26-41) and events. HIP streams and events replace
// the reference's cuBLAS handle and cuRAND generator.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "rplsh.h"

namespace rplsh {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string& msg) { g_last_error = msg; }
void clear_last_error() { g_last_error.clear(); }

void fail(rplsh_status s, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  throw Error(s, buf);
}

int current_device() {
  int d = 0;
  RPLSH_HIP(hipGetDevice(&d));
  return d;
}

DeviceScope::DeviceScope(int device) {
  int cur = current_device();
  if (cur != device) {
    RPLSH_HIP(hipSetDevice(device));
    prev = cur;
  }
}

DeviceScope::~DeviceScope() {
  if (prev >= 0) (void)hipSetDevice(prev);
}

int pointer_device(const void* p) {
  if (!p) return -1;
  hipPointerAttribute_t attr;
  hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // plain host pointer: clear the sticky error
    return -1;
  }
  if (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged) return attr.device;
  return -1;
}

}  // namespace rplsh

using namespace rplsh;

// Per-device stream set, created lazily by initializeForDevice.
struct DeviceStreams {
  hipStream_t def = nullptr;       // created default stream
  hipStream_t user_def = nullptr;  // user override (not owned)
  bool has_user = false;
  hipStream_t copy = nullptr;
  std::vector<hipStream_t> alt;
};

struct rplsh_resources {
  std::map<int, DeviceStreams> devs;
};

struct rplsh_event {
  hipEvent_t ev;
};

static DeviceStreams& init_device(rplsh_resources* res, int device) {
  auto it = res->devs.find(device);
  if (it != res->devs.end()) return it->second;
  int count = 0;
  RPLSH_HIP(hipGetDeviceCount(&count));
  RPLSH_REQUIRE(device >= 0 && device < count, "device %d out of range [0, %d)", device, count);
  DeviceScope scope(device);
  DeviceStreams s;
  // Every stream at the default priority, non-blocking w.r.t. the null stream.
  RPLSH_HIP(hipStreamCreateWithFlags(&s.def, hipStreamNonBlocking));
  RPLSH_HIP(hipStreamCreateWithFlags(&s.copy, hipStreamNonBlocking));
  for (int i = 0; i < 2; ++i) {
    hipStream_t a;
    RPLSH_HIP(hipStreamCreateWithFlags(&a, hipStreamNonBlocking));
    s.alt.push_back(a);
  }
  return res->devs.emplace(device, s).first->second;
}

extern "C" {

const char* rplsh_last_error(void) { return g_last_error.c_str(); }
const char* rplsh_version(void) { return "curplsh_amd 0.1 (gfx950)"; }

rplsh_status rplsh_device_count(int* count) {
  return guard([&] {
    RPLSH_REQUIRE(count, "count is NULL");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      c = 0;
    }
    *count = c;
  });
}

rplsh_status rplsh_get_device(int* device) {
  return guard([&] {
    RPLSH_REQUIRE(device, "device is NULL");
    *device = current_device();
  });
}

rplsh_status rplsh_set_device(int device) {
  return guard([&] { RPLSH_HIP(hipSetDevice(device)); });
}

rplsh_status rplsh_pointer_device(const void* ptr, int* device) {
  return guard([&] {
    RPLSH_REQUIRE(device, "device is NULL");
    *device = pointer_device(ptr);
  });
}

rplsh_status rplsh_resources_create(rplsh_resources** out) {
  return guard([&] {
    RPLSH_REQUIRE(out, "out is NULL");
    *out = new rplsh_resources();
  });
}

rplsh_status rplsh_resources_destroy(rplsh_resources* res) {
  return guard([&] {
    if (!res) return;
    for (auto& kv : res->devs) {
      DeviceScope scope(kv.first);
      (void)hipStreamSynchronize(kv.second.def);
      (void)hipStreamDestroy(kv.second.def);
      (void)hipStreamDestroy(kv.second.copy);
      for (auto a : kv.second.alt) (void)hipStreamDestroy(a);
    }
    delete res;
  });
}

rplsh_status rplsh_resources_init_device(rplsh_resources* res, int device) {
  return guard([&] {
    RPLSH_REQUIRE(res, "resources is NULL");
    init_device(res, device);
  });
}

rplsh_status rplsh_resources_default_stream(rplsh_resources* res, int device, void** stream) {
  return guard([&] {
    RPLSH_REQUIRE(res && stream, "NULL argument");
    auto& s = init_device(res, device);
    *stream = s.has_user ? (void*)s.user_def : (void*)s.def;
  });
}

rplsh_status rplsh_resources_set_default_stream(rplsh_resources* res, int device, void* stream) {
  return guard([&] {
    RPLSH_REQUIRE(res, "resources is NULL");
    auto& s = init_device(res, device);
    s.has_user = stream != nullptr;
    s.user_def = (hipStream_t)stream;
  });
}

rplsh_status rplsh_resources_alternate_streams(rplsh_resources* res, int device, void** streams,
                                               int* count) {
  return guard([&] {
    RPLSH_REQUIRE(res && count, "NULL argument");
    auto& s = init_device(res, device);
    int n = (int)s.alt.size();
    if (streams) {
      for (int i = 0; i < n && i < *count; ++i) streams[i] = s.alt[i];
    }
    *count = n;
  });
}

rplsh_status rplsh_resources_async_copy_stream(rplsh_resources* res, int device, void** stream) {
  return guard([&] {
    RPLSH_REQUIRE(res && stream, "NULL argument");
    *stream = init_device(res, device).copy;
  });
}

rplsh_status rplsh_resources_sync_default_stream(rplsh_resources* res, int device) {
  return guard([&] {
    RPLSH_REQUIRE(res, "resources is NULL");
    auto& s = init_device(res, device);
    DeviceScope scope(device);
    RPLSH_HIP(hipStreamSynchronize(s.has_user ? s.user_def : s.def));
  });
}

rplsh_status rplsh_malloc(int device, size_t bytes, int memory_space, void** out) {
  return guard([&] {
    RPLSH_REQUIRE(out, "out is NULL");
    RPLSH_REQUIRE(memory_space == 0 || memory_space == 1, "bad memory space %d", memory_space);
    DeviceScope scope(device);
    *out = nullptr;
    if (bytes == 0) return;
    if (memory_space == 0) {
      RPLSH_HIP(hipMalloc(out, bytes));
    } else {
      RPLSH_HIP(hipMallocManaged(out, bytes));
    }
  });
}

rplsh_status rplsh_free(void* ptr) {
  return guard([&] {
    if (ptr) RPLSH_HIP(hipFree(ptr));
  });
}

rplsh_status rplsh_memcpy_async(void* dst, const void* src, size_t bytes, void* stream) {
  return guard([&] {
    if (bytes == 0) return;
    RPLSH_REQUIRE(dst && src, "NULL pointer");
    RPLSH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, (hipStream_t)stream));
  });
}

rplsh_status rplsh_memset_async(void* dst, int value, size_t bytes, void* stream) {
  return guard([&] {
    if (bytes == 0) return;
    RPLSH_REQUIRE(dst, "NULL pointer");
    RPLSH_HIP(hipMemsetAsync(dst, value, bytes, (hipStream_t)stream));
  });
}

rplsh_status rplsh_stream_sync(void* stream) {
  return guard([&] { RPLSH_HIP(hipStreamSynchronize((hipStream_t)stream)); });
}

rplsh_status rplsh_device_sync(void) {
  return guard([&] { RPLSH_HIP(hipDeviceSynchronize()); });
}

rplsh_status rplsh_host_alloc(size_t bytes, void** out) {
  return guard([&] {
    RPLSH_REQUIRE(out, "out is NULL");
    *out = nullptr;
    if (bytes) RPLSH_HIP(hipHostMalloc(out, bytes, hipHostMallocDefault));
  });
}

rplsh_status rplsh_host_free(void* ptr) {
  return guard([&] {
    if (ptr) RPLSH_HIP(hipHostFree(ptr));
  });
}

rplsh_status rplsh_mem_info(int device, size_t* free_bytes, size_t* total_bytes) {
  return guard([&] {
    RPLSH_REQUIRE(free_bytes && total_bytes, "NULL argument");
    DeviceScope scope(device);
    RPLSH_HIP(hipMemGetInfo(free_bytes, total_bytes));
  });
}

rplsh_status rplsh_event_create(rplsh_event** out) {
  return guard([&] {
    RPLSH_REQUIRE(out, "out is NULL");
    auto* e = new rplsh_event();
    hipError_t err = hipEventCreate(&e->ev);
    if (err != hipSuccess) {
      delete e;
      RPLSH_HIP(err);
    }
    *out = e;
  });
}

rplsh_status rplsh_event_destroy(rplsh_event* ev) {
  return guard([&] {
    if (!ev) return;
    (void)hipEventDestroy(ev->ev);
    delete ev;
  });
}

rplsh_status rplsh_event_record(rplsh_event* ev, void* stream) {
  return guard([&] {
    RPLSH_REQUIRE(ev, "event is NULL");
    RPLSH_HIP(hipEventRecord(ev->ev, (hipStream_t)stream));
  });
}

rplsh_status rplsh_event_elapsed_ms(rplsh_event* start, rplsh_event* stop, float* ms) {
  return guard([&] {
    RPLSH_REQUIRE(start && stop && ms, "NULL argument");
    RPLSH_HIP(hipEventElapsedTime(ms, start->ev, stop->ev));
  });
}

rplsh_status rplsh_event_sync(rplsh_event* ev) {
  return guard([&] {
    RPLSH_REQUIRE(ev, "event is NULL");
    RPLSH_HIP(hipEventSynchronize(ev->ev));
  });
}

}  // extern "C"
