// searchL2Distance without allocation: the variant of rplsh_search_l2_dev for
// callers that cannot allocate device memory (code under stream capture,
// tight serving loops, harnesses that forbid hipMalloc). Every temporary --
// norms when the caller gives none, rows padded to 16 bytes when x is not,
// the search workspaces of every retry depth -- is carved from `scratch`, a
// file-scope device array of kScratchBytes; no hipMalloc, hipFree, host
// allocation or event creation happens on this path.
//
// `scratch` as named in host code is a host-side stand-in, not a device
// address: its device address is taken once per device with
// hipGetSymbolAddress (with that device current) and only that address is
// used. Every call on a device shares that device's copy: concurrent calls on
// one device race (callers serialise them); contents persist between calls,
// and everything the search needs zeroed is cleared by its own fill kernels.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "common.hpp"
#include "engine.hpp"
#include "rplsh.h"

namespace {

constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;

__device__ __align__(256) char scratch[kScratchBytes];

std::mutex g_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's scratch (the current device).
char* scratch_for_current_device() {
  const int dev = rplsh::current_device();
  RPLSH_REQUIRE(dev >= 0 && dev < kMaxDevices, "device %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    RPLSH_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    RPLSH_REQUIRE(ptr != nullptr, "no device address for the scratch array");
    g_scratch_ptr[dev] = ptr;
  }
  return static_cast<char*>(g_scratch_ptr[dev]);
}

}  // namespace

extern "C" {

size_t rplsh_search_l2_noalloc_scratch_bytes(void) { return kScratchBytes; }

rplsh_status rplsh_search_l2_dev_noalloc(const float* x, int64_t n, int d, const float* xnorms,
                                         const float* q, int64_t nq, int k, int32_t* indices,
                                         float* distances, int exact, void* stream) {
  using namespace rplsh;
  return guard([&] {
    RPLSH_REQUIRE(d > 0 && filter_dpad(Elem::F32, d) > 0, "dimension %d unsupported (1..%d)", d, kMaxKLoopDim);
    RPLSH_REQUIRE(n >= 0 && nq >= 0 && k > 0, "bad sizes");
    const bool aligned = (d % 4 == 0) && ((uintptr_t)x % 16 == 0);
    const size_t need = search_l2_device_bytes(n, d, xnorms != nullptr, aligned, nq, k);
    RPLSH_REQUIRE(need <= kScratchBytes,
                  "n = %lld, d = %d, nq = %lld, k = %d needs %zu bytes of scratch; the no-allocation "
                  "entry point holds %zu (split the queries, or use rplsh_search_l2_dev)",
                  (long long)n, d, (long long)nq, k, need, kScratchBytes);
    Arena arena;
    arena.base = scratch_for_current_device();
    arena.cap = kScratchBytes;
    L2Scratch sc{DevBuf(&arena), DevBuf(&arena), DevBuf(&arena), {}};
    sc.wss.reserve(4);
    for (int i = 0; i < 4; ++i) sc.wss.emplace_back(&arena);
    search_l2_device(x, n, d, xnorms, q, nq, k, indices, distances, exact, (hipStream_t)stream, sc);
  });
}

}  // extern "C"
