// Buffer loads for the neighbour gathers of the site kernels: a resource
// whose base may sit below the array (so that the scalar offset, a shifted
// neighbour delta, is non-negative) and 16-byte record loads.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace misa {
namespace site {

typedef __amdgpu_buffer_rsrc_t Rsrc;

__device__ __forceinline__ Rsrc rsrcAt(uintptr_t base) {
  return __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void*>(base), 0, 0x7fffffff, 0x00020000);
}
// Resource whose base sits `lo` elements (lo <= 0) below array `p`.
template <typename T>
__device__ __forceinline__ Rsrc rsrcBelow(const T* p, int lo) {
  return rsrcAt(reinterpret_cast<uintptr_t>(p) + static_cast<intptr_t>(lo) * static_cast<intptr_t>(sizeof(T)));
}

__device__ __forceinline__ double ldD(Rsrc r, uint32_t v, uint32_t s) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, v, s, 0));
}
__device__ __forceinline__ float ldF(Rsrc r, uint32_t v, uint32_t s) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, v, s, 0));
}
__device__ __forceinline__ int8_t ldB(Rsrc r, uint32_t v, uint32_t s) {
  return static_cast<int8_t>(__builtin_amdgcn_raw_buffer_load_b8(r, v, s, 0));
}
template <typename T>
__device__ __forceinline__ T ldT(Rsrc r, uint32_t v, uint32_t s);
template <>
__device__ __forceinline__ double ldT<double>(Rsrc r, uint32_t v, uint32_t s) {
  return ldD(r, v, s);
}
template <>
__device__ __forceinline__ float ldT<float>(Rsrc r, uint32_t v, uint32_t s) {
  return ldF(r, v, s);
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// First N values of a record at byte offset `o` of resource `r`.  The 16-byte
// load is reinterpreted as a whole (float4 / double2): bit-casting single
// elements of the loaded vector is miscompiled by this toolchain (only lane 0
// is kept).
template <typename T, int N>
__device__ __forceinline__ void ldRec(Rsrc r, uint32_t o, T (&out)[N]) {
  constexpr int per = 16 / sizeof(T);
  constexpr int full = N / per;
#pragma unroll
  for (int q = 0; q < full; q++) {
    const u32x4 w = __builtin_amdgcn_raw_buffer_load_b128(r, o + 16 * q, 0, 0);
    if constexpr (sizeof(T) == 4) {
      const float4 f = __builtin_bit_cast(float4, w);
      out[4 * q + 0] = f.x;
      out[4 * q + 1] = f.y;
      out[4 * q + 2] = f.z;
      out[4 * q + 3] = f.w;
    } else {
      const double2 d = __builtin_bit_cast(double2, w);
      out[2 * q + 0] = d.x;
      out[2 * q + 1] = d.y;
    }
  }
#pragma unroll
  for (int q = full * per; q < N; q++) out[q] = ldT<T>(r, o + q * static_cast<uint32_t>(sizeof(T)), 0);
}

typedef unsigned int u32x3 __attribute__((ext_vector_type(3)));
typedef float f32x3 __attribute__((ext_vector_type(3)));

// Three floats (12 bytes) at byte offset `o`: a load whose destination
// registers are all used, so none of them is reused while it is in flight.
__device__ __forceinline__ f32x3 ld3(Rsrc r, uint32_t o) {
  return __builtin_bit_cast(f32x3, __builtin_amdgcn_raw_buffer_load_b96(r, o, 0, 0));
}

__device__ __forceinline__ float4 ldRec4(Rsrc r, uint32_t v, uint32_t s) {
  return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, v, s, 0));
}
__device__ __forceinline__ double2 ldRec64(Rsrc r, uint32_t v, uint32_t s) {
  return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, v, s, 0));
}

}  // namespace site
}  // namespace misa
