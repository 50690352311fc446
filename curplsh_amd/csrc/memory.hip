// HBM budget with reusable device blocks: the implementation behind
// curplsh::DeviceMemory / DeviceMemoryReservation (reference interface:
// include/device_memory.h:12-89) and the admission budget of the serving
// loop.
//
// A budget bounds the bytes reserved at any time. Reservations with memory
// (rplsh_memory_reserve) take a cached free block of a fitting size (at least
// the request, at most twice it) or allocate a new one, so steady-state
// reservations do not call hipMalloc; cached blocks count against the budget
// too, and are freed (smallest first) when a new allocation would push
// reserved + cached past it. Accounting-only reservations (the server's batch
// workspaces and graph caches, which own their buffers) take budget bytes
// without memory. Releasing a block waits for the stream it was reserved on,
// so work still reading it has finished before anyone else gets it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>

#include "common.hpp"
#include "engine.hpp"
#include "rplsh.h"

using namespace rplsh;

struct rplsh_memory {
  int device = 0;
  size_t budget = 0;
  std::mutex mu;
  std::condition_variable cv;
  size_t reserved = 0;       // bytes held by reservations (with or without memory)
  size_t peak_reserved = 0;
  size_t cached = 0;         // bytes of free blocks kept for reuse
  size_t allocated = 0;      // bytes currently obtained from hipMalloc (live + cached)
  size_t high_water = 0;     // peak of `allocated`
  int64_t reservations = 0, waits = 0, mallocs = 0;
  struct Block {
    size_t bytes;       // block size
    size_t charged;     // bytes charged to the budget (the block size)
    hipStream_t stream;
  };
  std::multimap<size_t, void*> free_blocks;  // size -> block
  std::unordered_map<void*, Block> live;
};

namespace {

constexpr size_t kGranule = 256;

size_t round_bytes(size_t b) { return (std::max<size_t>(b, 1) + kGranule - 1) / kGranule * kGranule; }

// Frees cached blocks, smallest first, until reserved + cached + extra fits
// the budget (lock held).
void trim_cache(rplsh_memory* m, size_t extra) {
  while (!m->free_blocks.empty() && m->reserved + m->cached + extra > m->budget) {
    auto it = m->free_blocks.begin();
    (void)hipFree(it->second);
    m->cached -= it->first;
    m->allocated -= it->first;
    m->free_blocks.erase(it);
  }
}

// Waits until `bytes` more fit in the budget (lock held); false on timeout.
bool wait_for_room(rplsh_memory* m, std::unique_lock<std::mutex>& lk, size_t bytes, double timeout_s) {
  auto fits = [&] { return m->reserved + bytes <= m->budget; };
  if (fits()) return true;
  if (timeout_s == 0.0) return false;
  m->waits += 1;
  if (timeout_s < 0) {
    m->cv.wait(lk, fits);
    return true;
  }
  return m->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), fits);
}

}  // namespace

namespace rplsh {

rplsh_memory* memory_create(int device, size_t budget) {
  auto* m = new rplsh_memory();
  m->device = device;
  m->budget = budget;
  return m;
}

void memory_destroy(rplsh_memory* m) {
  if (!m) return;
  DeviceScope scope(m->device);
  std::lock_guard<std::mutex> lk(m->mu);
  for (auto& f : m->free_blocks) (void)hipFree(f.second);
  for (auto& l : m->live) {
    if (l.second.stream) (void)hipStreamSynchronize(l.second.stream);
    (void)hipFree(l.first);
  }
  m->free_blocks.clear();
  m->live.clear();
  delete m;
}

bool memory_try_charge(rplsh_memory* m, size_t bytes) {
  std::lock_guard<std::mutex> lk(m->mu);
  if (m->reserved + bytes > m->budget) return false;
  m->reserved += bytes;
  m->peak_reserved = std::max(m->peak_reserved, m->reserved);
  return true;
}

void memory_force_charge(rplsh_memory* m, size_t bytes) {
  std::lock_guard<std::mutex> lk(m->mu);
  m->reserved += bytes;
  m->peak_reserved = std::max(m->peak_reserved, m->reserved);
}

void memory_uncharge(rplsh_memory* m, size_t bytes) {
  {
    std::lock_guard<std::mutex> lk(m->mu);
    m->reserved -= std::min(bytes, m->reserved);
  }
  m->cv.notify_all();
}

size_t memory_budget(const rplsh_memory* m) { return m->budget; }
int memory_device(const rplsh_memory* m) { return m->device; }

size_t memory_reserved(rplsh_memory* m) {
  std::lock_guard<std::mutex> lk(m->mu);
  return m->reserved;
}

size_t memory_peak(rplsh_memory* m) {
  std::lock_guard<std::mutex> lk(m->mu);
  return m->peak_reserved;
}

void memory_reset_peak(rplsh_memory* m) {
  std::lock_guard<std::mutex> lk(m->mu);
  m->peak_reserved = m->reserved;
}

}  // namespace rplsh

extern "C" {

rplsh_status rplsh_memory_create(int device, size_t budget_bytes, rplsh_memory** out) {
  return guard([&] {
    RPLSH_REQUIRE(out, "out is NULL");
    int count = 0;
    rplsh_device_count(&count);
    RPLSH_REQUIRE(device >= 0 && device < count, "device %d out of range [0, %d)", device, count);
    if (budget_bytes == 0) {
      DeviceScope scope(device);
      size_t fr = 0, tot = 0;
      RPLSH_HIP(hipMemGetInfo(&fr, &tot));
      budget_bytes = fr / 4;
    }
    *out = memory_create(device, budget_bytes);
  });
}

rplsh_status rplsh_memory_destroy(rplsh_memory* m) {
  return guard([&] {
    if (!m) return;
    {
      std::lock_guard<std::mutex> lk(m->mu);
      if (!m->live.empty()) fail(RPLSH_ERR_STATE, "%zu reservations still held", m->live.size());
    }
    memory_destroy(m);
  });
}

rplsh_status rplsh_memory_reserve(rplsh_memory* m, void* stream, size_t bytes, double timeout_s,
                                  void** ptr) {
  return guard([&] {
    RPLSH_REQUIRE(m && ptr, "NULL argument");
    *ptr = nullptr;
    const size_t b = round_bytes(bytes);
    if (b > m->budget)
      fail(RPLSH_ERR_BUDGET, "reservation of %zu bytes exceeds the budget of %zu", bytes, m->budget);
    std::unique_lock<std::mutex> lk(m->mu);
    if (!wait_for_room(m, lk, b, timeout_s))
      fail(RPLSH_ERR_BUDGET, "%zu bytes not available within %.3f s (%zu of %zu reserved)", bytes,
           timeout_s, m->reserved, m->budget);
    // A cached block of size in [b, 2b]: no allocation.
    void* p = nullptr;
    size_t got = 0;
    auto it = m->free_blocks.lower_bound(b);
    if (it != m->free_blocks.end() && it->first <= 2 * b && m->reserved + it->first <= m->budget) {
      p = it->second;
      got = it->first;
      m->cached -= got;
      m->free_blocks.erase(it);
    } else {
      trim_cache(m, b);
      DeviceScope scope(m->device);
      const hipError_t e = hipMalloc(&p, b);
      if (e != hipSuccess) {
        // Everything cached goes back to the device, then one more try.
        (void)hipGetLastError();
        trim_cache(m, m->budget);
        RPLSH_HIP(hipMalloc(&p, b));
      }
      got = b;
      m->allocated += b;
      m->mallocs += 1;
      m->high_water = std::max(m->high_water, m->allocated);
    }
    m->reserved += got;
    m->peak_reserved = std::max(m->peak_reserved, m->reserved);
    m->reservations += 1;
    m->live[p] = rplsh_memory::Block{got, got, (hipStream_t)stream};
    *ptr = p;
  });
}

rplsh_status rplsh_memory_release(rplsh_memory* m, void* ptr) {
  return guard([&] {
    RPLSH_REQUIRE(m, "NULL argument");
    if (!ptr) return;
    hipStream_t st = nullptr;
    {
      std::lock_guard<std::mutex> lk(m->mu);
      auto it = m->live.find(ptr);
      RPLSH_REQUIRE(it != m->live.end(), "%p is not a reservation of this budget", ptr);
      st = it->second.stream;
    }
    // Work ordered on the reservation's stream may still use the block.
    {
      DeviceScope scope(m->device);
      RPLSH_HIP(hipStreamSynchronize(st));
    }
    {
      std::lock_guard<std::mutex> lk(m->mu);
      auto it = m->live.find(ptr);
      const rplsh_memory::Block blk = it->second;
      m->live.erase(it);
      m->reserved -= blk.charged;
      m->free_blocks.emplace(blk.bytes, ptr);
      m->cached += blk.bytes;
      trim_cache(m, 0);
    }
    m->cv.notify_all();
  });
}

rplsh_status rplsh_memory_get_stats(rplsh_memory* m, rplsh_memory_stats* out) {
  return guard([&] {
    RPLSH_REQUIRE(m && out, "NULL argument");
    std::lock_guard<std::mutex> lk(m->mu);
    std::memset(out, 0, sizeof(*out));
    out->device = m->device;
    out->budget_bytes = m->budget;
    out->reserved_bytes = m->reserved;
    out->peak_reserved_bytes = m->peak_reserved;
    out->cached_bytes = m->cached;
    size_t largest = m->free_blocks.empty() ? 0 : m->free_blocks.rbegin()->first;
    out->available_without_malloc = std::min(largest, m->budget - std::min(m->budget, m->reserved));
    out->high_water_malloc_bytes = m->high_water;
    out->reservations = m->reservations;
    out->waits = m->waits;
    out->mallocs = m->mallocs;
  });
}

}  // extern "C"
