// DeviceResources (reference: include/device_resources.h:13-263): per-device
// default, async-copy and alternate streams, created lazily. The reference's
// cuBLAS handle and cuRAND generator have no counterpart: the contraction is
// this library's own HIP kernel and nothing draws random numbers on device.
#pragma once

#include <initializer_list>
#include <vector>

#include "memory_space.h"

namespace curplsh {

/// RAII: make `device` current, restore the previous one on destruction
/// (device_resources.h:13-32).
class DeviceScope {
 public:
  explicit DeviceScope(int device) {
    prevDevice_ = getCurrentDevice();
    if (prevDevice_ != device) {
      setCurrentDevice(device);
    } else {
      prevDevice_ = -1;
    }
  }
  ~DeviceScope() {
    if (prevDevice_ != -1) setCurrentDevice(prevDevice_);
  }
  DeviceScope(const DeviceScope&) = delete;
  DeviceScope& operator=(const DeviceScope&) = delete;

 private:
  int prevDevice_;
};

class DeviceResources {
 public:
  DeviceResources();
  explicit DeviceResources(int device);
  DeviceResources(std::initializer_list<int> devices);
  ~DeviceResources();
  DeviceResources(const DeviceResources&) = delete;
  DeviceResources& operator=(const DeviceResources&) = delete;

  /// Pre-allocates the streams of a device (device_resources.h:113-165).
  void initializeForDevice(int device);
  /// Replaces the default stream of a device by a caller-owned one
  /// (userDefaultStreams_, device_resources.h:248); nullptr restores it.
  void setDefaultStream(int device, Stream stream);

  Stream getDefaultStream(int device);
  std::vector<Stream> getAlternateStreams(int device);
  Stream getAsyncCopyStream(int device);
  Stream getDefaultStreamCurrentDevice() { return getDefaultStream(getCurrentDevice()); }
  std::vector<Stream> getAlternateStreamsCurrentDevice() {
    return getAlternateStreams(getCurrentDevice());
  }
  Stream getAsyncCopyStreamCurrentDevice() { return getAsyncCopyStream(getCurrentDevice()); }
  void syncDefaultStream(int device);
  void syncDefaultStreamCurrentDevice() { syncDefaultStream(getCurrentDevice()); }

  /// The C-ABI handle (for code that binds the C layer directly).
  rplsh_resources* handle() { return handle_; }

 private:
  rplsh_resources* handle_ = nullptr;
};

}  // namespace curplsh
