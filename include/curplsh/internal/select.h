// l2Select and blockSelect (reference: include/internal/select.h:8-34) on
// device tensors. Results are ascending (descending when selectMax), ties
// broken by the lower label; k larger than a row pads with FLT_MAX (-FLT_MAX
// when selectMax) and -1. k <= 4096. They return when the results are in place.
#pragma once

#include "../tensor.h"

namespace curplsh {

/// Blockwise k-selection; the labels are the column indices.
void blockSelect(Tensor<float, 2>& in, Tensor<float, 2>& outK, Tensor<int, 2>& outV, int k,
                 bool selectMax, Stream stream);

/// Blockwise k-selection with given labels.
void blockSelect(Tensor<float, 2>& inK, Tensor<int, 2>& inV, Tensor<float, 2>& outK,
                 Tensor<int, 2>& outV, int k, bool selectMax, Stream stream);

/// k smallest baseNorms[i] + productDistances[r][i] per row, with indices.
void l2Select(Tensor<float, 2>& productDistances, Tensor<float, 1>& baseNorms,
              Tensor<float, 2>& distances, Tensor<int, 2>& indices, int k, Stream stream);

}  // namespace curplsh
