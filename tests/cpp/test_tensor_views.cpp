// Host-only checks of curplsh::Tensor / HostTensor views (reference:
// include/tensor.h:16-449, include/host_tensor.h): strides, operator[] down
// to elements, narrow along any dimension, view, cast / isCastable,
// isContiguous, and 64-bit element offsets past 2^31. No GPU needed.
#include <cstdint>
#include <cstdio>
#include <vector>

#include "host_tensor.h"

#define CHECK(c)                                                          \
  do {                                                                    \
    if (!(c)) {                                                           \
      std::fprintf(stderr, "FAIL %s at %s:%d\n", #c, __FILE__, __LINE__); \
      return 1;                                                           \
    }                                                                     \
  } while (0)

using namespace curplsh;

int main() {
  // 3 x 4 x 5, filled through operator[] and read back through data().
  HostTensor<float, 3> t({3, 4, 5});
  CHECK(t.owner() && t.getNumElements() == 60 && t.getDataMemSize() == 240);
  CHECK(t.getStride(0) == 20 && t.getStride(1) == 5 && t.getStride(2) == 1 && t.getStride(-1) == 1);
  CHECK(t.getSize(-1) == 5 && t.isContiguous());
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 5; ++k) t[i][j][k] = (float)(100 * i + 10 * j + k);
  for (int e = 0; e < 60; ++e) CHECK(t.data()[e] == (float)(100 * (e / 20) + 10 * (e / 5 % 4) + e % 5));
  CHECK(&t[2][3] == t.data() + 55);
  const HostTensor<float, 3>& ct = t;
  CHECK((float)ct[1][2][3] == 123.0f);

  // narrow along the middle dimension keeps the strides: not contiguous.
  Tensor<float, 3> n1 = t.narrow(1, 1, 2);
  CHECK(n1.getSize(0) == 3 && n1.getSize(1) == 2 && n1.getSize(2) == 5);
  CHECK(n1.getStride(0) == 20 && n1.data() == t.data() + 5 && !n1.isContiguous());
  CHECK((float)n1[2][1][4] == 224.0f);
  CHECK(!n1.isContiguous(0) && n1.isContiguous(1) && n1.isContiguous(2));
  // narrow along the outermost dimension stays contiguous.
  Tensor<float, 3> n0 = t.narrow(0, 1, 2);
  CHECK(n0.isContiguous() && (float)n0[0][0][0] == 100.0f && n0.getNumElements() == 40);
  CHECK(n0.isSame(t.narrowOutermost(1, 2)) && n0.isSameSizes(t.narrow(0, 0, 2)));

  // view: same elements, new shape.
  Tensor<float, 2> v = t.view<2>({12, 5});
  CHECK(v.getSize(0) == 12 && v.getStride(0) == 5 && (float)v[7][3] == 133.0f);
  Tensor<float, 1> flat = t.view<1>({60});
  CHECK((float)flat[59] == 234.0f);

  // cast float -> float2 (innermost size 6 = 3 pairs) and back-checks.
  HostTensor<float, 2> p({4, 6});
  for (int i = 0; i < 24; ++i) p.data()[i] = (float)i;
  struct alignas(8) F2 {
    float x, y;
  };
  CHECK(p.isCastable<F2>());
  Tensor<F2, 2> p2 = p.cast<F2>();
  CHECK(p2.getSize(0) == 4 && p2.getSize(1) == 3 && p2.getStride(0) == 3);
  CHECK(((F2&)p2[2][1]).x == 14.0f && ((F2&)p2[2][1]).y == 15.0f);
  CHECK(!t.isCastable<F2>());  // innermost size 5 is not a whole number of pairs
  Tensor<float, 2> odd = p.narrow(1, 1, 4);  // misaligned start
  CHECK(!odd.isCastable<F2>());

  // Aliasing: no ownership, no copy.
  std::vector<int> buf(10, 7);
  HostTensor<int, 2> a(buf.data(), {2, 5});
  CHECK(!a.owner() && a.data() == buf.data());
  a[1][4] = 3;
  CHECK(buf[9] == 3);
  Tensor<int, 2> moved(std::move(a));
  CHECK(moved.data() == buf.data() && a.data() == nullptr && a.getSize(0) == 0);

  // 64-bit element offsets: strides of 2^20 x index 3000 pass 2^31 elements.
  const int sizes[2] = {4000, 1 << 20};
  Tensor<float, 2> big(reinterpret_cast<float*>(0x1000), sizes);
  CHECK(&big[3000][5] == reinterpret_cast<float*>(0x1000) + (int64_t)3000 * (1 << 20) + 5);
  CHECK(big.getNumElements() == (size_t)4000 << 20);
  std::printf("tensor views ok\n");
  return 0;
}
