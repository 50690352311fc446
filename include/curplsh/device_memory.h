// DeviceMemory / DeviceMemoryReservation (reference:
// include/device_memory.h:12-89): temporary device memory ordered with a
// stream, handed out by a manager and returned when the reservation dies.
// The interface is the reference's (cudaStream_t becomes curplsh::Stream);
// BudgetDeviceMemory implements it over the library's HBM budget
// (rplsh_memory_*): reservations come from reusable blocks, never exceed the
// budget (getMemory waits for room), and can share one budget with a serving
// loop (rplsh_server_memory).
#pragma once

#include <cstdio>
#include <string>
#include <utility>

#include "memory_space.h"

namespace curplsh {

class DeviceMemoryReservation;

class DeviceMemory {
 public:
  virtual ~DeviceMemory() = default;
  /// Device the memory belongs to.
  virtual int getDevice() const = 0;
  /// A temporary allocation of `size` bytes whose use is ordered on `stream`.
  virtual DeviceMemoryReservation getMemory(Stream stream, size_t size) = 0;
  /// Bytes a reservation can get right now without a new device allocation.
  virtual size_t getSizeAvailable() const = 0;
  virtual std::string toString() const = 0;
  /// Peak bytes obtained from the device allocator.
  virtual size_t getHighWaterCudaMalloc() const = 0;

 protected:
  friend class DeviceMemoryReservation;
  virtual void returnAllocation(DeviceMemoryReservation& m) = 0;
};

class DeviceMemoryReservation {
 public:
  DeviceMemoryReservation() : state_(nullptr), device_(0), data_(nullptr), size_(0), stream_(nullptr) {}
  DeviceMemoryReservation(DeviceMemory* state, int device, void* data, size_t size, Stream stream)
      : state_(state), device_(device), data_(data), size_(size), stream_(stream) {}
  DeviceMemoryReservation(DeviceMemoryReservation&& o) noexcept : DeviceMemoryReservation() {
    *this = std::move(o);
  }
  DeviceMemoryReservation& operator=(DeviceMemoryReservation&& o) noexcept {
    if (this != &o) {
      giveBack();
      state_ = o.state_;
      device_ = o.device_;
      data_ = o.data_;
      size_ = o.size_;
      stream_ = o.stream_;
      o.data_ = nullptr;
    }
    return *this;
  }
  DeviceMemoryReservation(const DeviceMemoryReservation&) = delete;
  DeviceMemoryReservation& operator=(const DeviceMemoryReservation&) = delete;
  ~DeviceMemoryReservation() { giveBack(); }

  int device() const { return device_; }
  void* get() const { return data_; }
  size_t size() const { return size_; }
  Stream stream() const { return stream_; }

 private:
  void giveBack() {
    if (data_) {
      host_assert(state_);
      state_->returnAllocation(*this);
    }
    data_ = nullptr;
  }
  DeviceMemory* state_;
  int device_;
  void* data_;
  size_t size_;
  Stream stream_;
};

/// DeviceMemory over an HBM budget: its own (`budgetBytes` on `device`; 0 = a
/// quarter of free memory) or an existing one, e.g. a serving loop's.
class BudgetDeviceMemory : public DeviceMemory {
 public:
  BudgetDeviceMemory(int device, size_t budgetBytes) : own_(true) {
    checkStatus(rplsh_memory_create(device, budgetBytes, &mem_), "BudgetDeviceMemory");
  }
  explicit BudgetDeviceMemory(rplsh_memory* shared) : mem_(shared), own_(false) {}
  ~BudgetDeviceMemory() override {
    if (own_) checkStatus(rplsh_memory_destroy(mem_), "~BudgetDeviceMemory");
  }
  BudgetDeviceMemory(const BudgetDeviceMemory&) = delete;
  BudgetDeviceMemory& operator=(const BudgetDeviceMemory&) = delete;

  int getDevice() const override { return stats().device; }
  /// Waits while other reservations hold the budget; aborts if `size`
  /// exceeds the whole budget.
  DeviceMemoryReservation getMemory(Stream stream, size_t size) override {
    void* p = nullptr;
    checkStatus(rplsh_memory_reserve(mem_, stream, size, -1.0, &p), "DeviceMemory::getMemory");
    return DeviceMemoryReservation(this, getDevice(), p, size, stream);
  }
  size_t getSizeAvailable() const override { return stats().available_without_malloc; }
  std::string toString() const override {
    const rplsh_memory_stats s = stats();
    char buf[256];
    std::snprintf(buf, sizeof(buf),
                  "device %d: budget %zu, reserved %zu (peak %zu), cached %zu, high water %zu, "
                  "%lld reservations, %lld waits, %lld allocations",
                  s.device, s.budget_bytes, s.reserved_bytes, s.peak_reserved_bytes, s.cached_bytes,
                  s.high_water_malloc_bytes, (long long)s.reservations, (long long)s.waits,
                  (long long)s.mallocs);
    return buf;
  }
  size_t getHighWaterCudaMalloc() const override { return stats().high_water_malloc_bytes; }
  rplsh_memory_stats stats() const {
    rplsh_memory_stats s;
    checkStatus(rplsh_memory_get_stats(mem_, &s), "DeviceMemory::stats");
    return s;
  }
  rplsh_memory* handle() { return mem_; }

 protected:
  void returnAllocation(DeviceMemoryReservation& m) override {
    checkStatus(rplsh_memory_release(mem_, m.get()), "DeviceMemory::returnAllocation");
  }

 private:
  rplsh_memory* mem_ = nullptr;
  bool own_ = false;
};

}  // namespace curplsh
