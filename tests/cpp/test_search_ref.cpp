// The body of the reference's own driver (test/test_search.cpp:17-64),
// compiled against include/curplsh/ with the include path as the only build
// change. Lines that are CUDA-only are replaced, one for one, and marked:
//   getCudaDevice(argc, argv)   (a CUDA sample helper)  -> the first argument
//   min(float, float)           (CUDA's global overload) -> std::min
//   cudaMemPrefetchAsync x4     (unified-memory hints)   -> removed
// After the reference's printout, the closed-form answer is checked: queries
// of ones against rows of min(i, 10) give relative distances
// 128 m^2 - 256 m (m = min(i, 10)), so the top 10 are rows 1, 0, 2, ..., 9
// (rows 0 and 2 tie at 0: the lower index first), all exact in FP32.
#include <algorithm>
#include <cstdlib>
#include <initializer_list>
#include <iostream>

#include "device_resources.h"
#include "device_tensor.h"
#include "timer.h"

#include "internal/norm.h"
#include "internal/search.h"
#include "internal/constants.h"

int main(int argc, const char** argv) {
  int device = argc > 1 ? std::atoi(argv[1]) : 0;  // CUDA-only: getCudaDevice(argc, argv)
  curplsh::DeviceScope scope(device);
  curplsh::DeviceResources resources;

  const int dim = 128;
  const int numBases = 1000;
  const int numQueries = 100;
  const int k = 10;

  curplsh::DeviceTensor<float, 2> bases({numBases, dim},
                                        curplsh::MemorySpace::Unified);
  curplsh::DeviceTensor<float, 1> baseNorms({numBases},
                                            curplsh::MemorySpace::Unified);
  curplsh::DeviceTensor<float, 2> queries({numQueries, dim},
                                          curplsh::MemorySpace::Unified);
  curplsh::DeviceTensor<int, 2> indices({numQueries, k},
                                        curplsh::MemorySpace::Unified);
  curplsh::DeviceTensor<float, 2> distances({numQueries, k},
                                            curplsh::MemorySpace::Unified);

  for (int d = 0; d < dim; ++d) {
    for (int i = 0; i < queries.getSize(0); ++i) queries[i][d] = 1.f;
    for (int i = 0; i < bases.getSize(0); ++i) bases[i][d] = std::min((float)i, 10.f);  // CUDA-only: min
  }

  // CUDA-only: the four cudaMemPrefetchAsync hints are removed.

  {
    curplsh::DeviceTimer timer;
    curplsh::computeL2Norm(bases, baseNorms, true);
  }

  {
    curplsh::DeviceTimer timer;
    resources.initializeForDevice(device);
  }

  for (int i = 0; i < 3; ++i) {
    curplsh::DeviceTimer timer;
    curplsh::searchL2Distance(&resources, bases, &baseNorms, queries, k, indices,
                              distances, false);
  }

  for (int i = 0; i < indices.getSize(0); ++i) {
    for (int j = 0; j < indices.getSize(1); ++j) {
      std::cout << '(' << indices[i][j] << ',' << distances[i][j] << ')' << '\t';
    }
    std::cout << std::endl;
  }

  // Not in the reference: the closed-form check.
  const int expect[10] = {1, 0, 2, 3, 4, 5, 6, 7, 8, 9};
  int bad = 0;
  for (int i = 0; i < numQueries; ++i) {
    for (int j = 0; j < k; ++j) {
      const int r = expect[j];
      const float m = std::min((float)r, 10.f);
      const float rel = 128.f * m * m - 256.f * m;
      if ((int)indices[i][j] != r || (float)distances[i][j] != rel) ++bad;
    }
    if ((float)baseNorms[i] != 128.f * std::min((float)i, 10.f) * std::min((float)i, 10.f)) ++bad;
  }
  if (bad) {
    std::cout << "test_search: " << bad << " wrong entries" << std::endl;
    return 1;
  }
  std::cout << "test_search: closed form ok" << std::endl;
  return 0;
}
