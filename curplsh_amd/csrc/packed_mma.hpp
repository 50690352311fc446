// MFMA operand fragments of the packed candidate copies (32-row tiles in
// operand order) and the contraction step per K-chunk, shared by the packed
// filter kernels (filter_packed.hip) and the persistent serving kernel
// (persist.hip): both compute a tile's accumulators with exactly these
// instructions in this order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace rplsh {

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

template <Elem E>
struct PackTraits;
template <>
struct PackTraits<Elem::BF16> {
  static constexpr int kBytes = 2;
  static constexpr int kLaneBytes = 16;  // 8 bf16
  static constexpr int kChunkElems = 16;
};
template <>
struct PackTraits<Elem::FP8> {
  static constexpr int kBytes = 1;
  static constexpr int kLaneBytes = 32;  // 32 e4m3
  static constexpr int kChunkElems = 64;
};

template <int LB>
struct Frag {
  i32x4 v[LB / 16];
};

template <int LB>
__device__ inline Frag<LB> load_frag(const char* p) {
  Frag<LB> f;
#pragma unroll
  for (int i = 0; i < LB / 16; ++i) f.v[i] = *reinterpret_cast<const i32x4*>(p + 16 * i);
  return f;
}

// The same load through a global (address space 1) pointer: for callers whose
// pointers were read from memory, where the compiler would otherwise emit flat
// loads that wait on the LDS counter too.
template <int LB>
__device__ inline Frag<LB> load_frag_global(const char* p) {
  typedef const __attribute__((address_space(1))) i32x4* gp;
  Frag<LB> f;
#pragma unroll
  for (int i = 0; i < LB / 16; ++i) f.v[i] = *(gp)(p + 16 * i);
  return f;
}

template <int LB>
__device__ inline Frag<LB> zero_frag() {
  Frag<LB> f;
#pragma unroll
  for (int i = 0; i < LB / 16; ++i) f.v[i] = i32x4{0, 0, 0, 0};
  return f;
}

template <Elem E>
__device__ inline f32x16 mfma_chunk(const Frag<PackTraits<E>::kLaneBytes>& a,
                                    const Frag<PackTraits<E>::kLaneBytes>& b, f32x16 c, int sa,
                                    int sb);

template <>
__device__ inline f32x16 mfma_chunk<Elem::BF16>(const Frag<16>& a, const Frag<16>& b, f32x16 c,
                                                int, int) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a.v[0]),
                                                  __builtin_bit_cast(bf16x8, b.v[0]), c, 0, 0, 0);
}

template <>
__device__ inline f32x16 mfma_chunk<Elem::FP8>(const Frag<32>& a, const Frag<32>& b, f32x16 c,
                                               int sa, int sb) {
  i32x8 av, bv;
  av.lo = a.v[0];
  av.hi = a.v[1];
  bv.lo = b.v[0];
  bv.hi = b.v[1];
  return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, bv, c, 0, 0, 0, sa, 0, sb);
}

__device__ inline int e8m0_of(float s) { return (int)((__float_as_uint(s) >> 23) & 0xff); }

template <Elem E, int DPAD>
struct Geometry {
  static constexpr int LB = PackTraits<E>::kLaneBytes;
  static constexpr int CHUNK = 64 * LB;                       // bytes of one K-chunk of a tile
  static constexpr int NC = DPAD / PackTraits<E>::kChunkElems;  // K-chunks per row
  static constexpr int TILE = NC * CHUNK;                     // bytes of a 32-row tile
  static constexpr int ROWB = DPAD * PackTraits<E>::kBytes;   // bytes of a query row
};

}  // namespace

}  // namespace rplsh
