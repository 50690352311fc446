// broadcastAlongRowsSum (reference: include/internal/broadcast.h:7-8):
// output[r][c] += input[r], enqueued on `stream`.
#pragma once

#include "../tensor.h"

namespace curplsh {

void broadcastAlongRowsSum(Tensor<float, 1>& input, Tensor<float, 2>& output, Stream stream);

}  // namespace curplsh
