// computeL2Norm (reference: include/internal/norm.h:7-8) on device tensors.
#pragma once

#include "../tensor.h"

namespace curplsh {

void computeL2Norm(const Tensor<float, 2>& vectors, Tensor<float, 1>& norms, bool squared,
                   Stream stream = nullptr);

}  // namespace curplsh
