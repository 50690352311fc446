// DeviceTensor (reference: include/device_tensor.h:8-186). Declared with
// Tensor in tensor.h; this header exists so code that includes
// "device_tensor.h", as the reference's drivers do, compiles unchanged.
#pragma once

#include "tensor.h"
