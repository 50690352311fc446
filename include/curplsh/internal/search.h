// searchL2Distance (reference: include/internal/search.h:8-11) on device
// tensors, plain FP32 path. basesNorm may be nullptr (then computed).
// needRealDistances adds ||q||^2 (exact distances), otherwise the relative
// distances ||x||^2 - 2 q.x are returned. Synchronises the default stream of
// the current device.
#pragma once

#include "../device_resources.h"
#include "../tensor.h"

namespace curplsh {

void searchL2Distance(DeviceResources* resources, const Tensor<float, 2>& bases,
                      Tensor<float, 1>* basesNorm, const Tensor<float, 2>& queries, int k,
                      Tensor<int, 2>& indices, Tensor<float, 2>& distances,
                      bool needRealDistances);

}  // namespace curplsh
