// Constants of the reference's internal headers (include/internal/constants.h)
// with this build's values: a wavefront is 64 lanes on CDNA, not a 32-lane
// warp, so kWarpSize is 64 and the full lane mask 64 bits wide. The search's
// sentinels are the same: kFloatMax as the empty key, -1 as the empty label.
#pragma once

#include <cstdint>
#include <limits>

namespace curplsh {

typedef unsigned int uint;
typedef unsigned short ushort;

constexpr int kWarpSize = 64;
constexpr int kMaxThreadsPerBlock = 1024;

constexpr uint64_t kFullMask = ~0ull;

constexpr float kFloatMin = std::numeric_limits<float>::lowest();
constexpr float kFloatMax = std::numeric_limits<float>::max();

constexpr int kIntMin = std::numeric_limits<int>::lowest();
constexpr int kIntMax = std::numeric_limits<int>::max();

constexpr uint kUintMin = std::numeric_limits<uint>::lowest();
constexpr uint kUintMax = std::numeric_limits<uint>::max();

}  // namespace curplsh
