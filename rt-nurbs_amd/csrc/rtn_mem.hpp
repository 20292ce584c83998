// Read-only device loads shared by the trace and seed-network code: control
// nets, boxes and node data are read through plain global pointers, or
// through the constant address space when every lane of a wave reads the same
// address; then the compiler fetches them with scalar loads into scalar
// registers (the trace never writes them).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace rtn {

template <class T>
using cptr = const __attribute__((address_space(4))) T*;

template <class T>
__device__ __forceinline__ cptr<T> as_const(const T* p) {
  return (cptr<T>)(p);
}

typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float4 ld4(const float4* p, int64_t i) { return p[i]; }
__device__ __forceinline__ float4 ld4(cptr<float4> p, int64_t i) {
  const f4v v = ((cptr<f4v>)(p))[i];
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float ld1(const float* p, int64_t i) { return p[i]; }
__device__ __forceinline__ float ld1(cptr<float> p, int64_t i) { return p[i]; }

}  // namespace rtn
