// Timers (reference: include/timer.h:10-95): HostTimer measures host wall
// time, DeviceTimer waits for the device first, EventTimer brackets work on
// one stream with HIP events (the reference's CUDATimer). Each prints its
// elapsed time on destruction, as the reference's do.
#pragma once

#include <chrono>
#include <cstdio>

#include "memory_space.h"

namespace curplsh {

inline void printTime(const char* name, float ms) { std::printf("[INFO] %s: %fms elapsed.\n", name, ms); }

class HostTimer {
 public:
  HostTimer() : HostTimer("") {}
  explicit HostTimer(const char* name) : name_(name), start_(Clock::now()) {}
  ~HostTimer() { printTime(name_, elapsed()); }
  HostTimer(const HostTimer&) = delete;
  HostTimer& operator=(const HostTimer&) = delete;
  /// Milliseconds since construction (after the device is idle, for a
  /// DeviceTimer).
  float elapsed() const {
    if (syncDevice_) checkStatus(rplsh_device_sync(), "DeviceTimer");
    return std::chrono::duration<float, std::milli>(Clock::now() - start_).count();
  }

 protected:
  typedef std::chrono::steady_clock Clock;
  const char* name_;
  Clock::time_point start_;
  bool syncDevice_ = false;
};

class DeviceTimer : public HostTimer {
 public:
  DeviceTimer() : DeviceTimer("") {}
  explicit DeviceTimer(const char* name) : HostTimer(name) { syncDevice_ = true; }
};

class EventTimer {
 public:
  EventTimer() : EventTimer("") {}
  explicit EventTimer(const char* name, Stream stream = nullptr) : name_(name), stream_(stream) {
    checkStatus(rplsh_event_create(&start_), "EventTimer");
    checkStatus(rplsh_event_create(&stop_), "EventTimer");
    start();
  }
  ~EventTimer() {
    printTime(name_, elapsed());
    rplsh_event_destroy(start_);
    rplsh_event_destroy(stop_);
  }
  EventTimer(const EventTimer&) = delete;
  EventTimer& operator=(const EventTimer&) = delete;
  void start() { checkStatus(rplsh_event_record(start_, stream_), "EventTimer::start"); }
  float elapsed() const {
    float ms = 0.0f;
    checkStatus(rplsh_event_record(stop_, stream_), "EventTimer::elapsed");
    checkStatus(rplsh_event_sync(stop_), "EventTimer::elapsed");
    checkStatus(rplsh_event_elapsed_ms(start_, stop_, &ms), "EventTimer::elapsed");
    return ms;
  }

 private:
  const char* name_;
  Stream stream_;
  rplsh_event* start_ = nullptr;
  rplsh_event* stop_ = nullptr;
};

}  // namespace curplsh
