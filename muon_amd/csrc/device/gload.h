// Loads from the scene arrays (and other device global memory reached
// through pointers read from memory) as global-address-space loads.
#ifndef MUONX_DEVICE_GLOAD_H_
#define MUONX_DEVICE_GLOAD_H_

#include <hip/hip_runtime.h>

namespace muonx {

// A load through a pointer of the global address space. A pointer read
// from memory (a scene view in the engine's scene table) is generic to the
// compiler, which then issues flat loads; the wave counts those as LDS
// operations too, so a wait for an LDS traversal-stack entry also waits for
// the node or primitive fetch in flight. Every scene array is in device
// global memory, so its loads go through this. The 16 B forms load the
// builtin vector type: copying a float4 / uint4 object out of a global
// lvalue would pass through its copy constructor's generic reference and
// lose the address space again.
typedef float GlobalF4 __attribute__((ext_vector_type(4)));
typedef unsigned int GlobalU4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ldg(const float4 *p) {
  const GlobalF4 v = *(const __attribute__((address_space(1))) GlobalF4 *)p;
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint4 ldg(const uint4 *p) {
  const GlobalU4 v = *(const __attribute__((address_space(1))) GlobalU4 *)p;
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ int4 ldg(const int4 *p) {
  const GlobalU4 v = *(const __attribute__((address_space(1))) GlobalU4 *)p;
  return make_int4(int(v.x), int(v.y), int(v.z), int(v.w));
}
template <class T>
__device__ __forceinline__ T ldg(const T *p) {
  static_assert(__is_arithmetic(T), "scalar loads only");
  return *(const __attribute__((address_space(1))) T *)p;
}

// The 12 entries of a column-major mat4 that xform_point / xform_vec read
// (columns 0-3, rows 0-2) into m[4 c + r]; row 3 is left unset.
__device__ __forceinline__ void load_xform(const float *xf, float m[16]) {
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    m[4 * c] = ldg(xf + 4 * c);
    m[4 * c + 1] = ldg(xf + 4 * c + 1);
    m[4 * c + 2] = ldg(xf + 4 * c + 2);
  }
}

// atomicAdd (relaxed, device scope) through a global-address-space pointer,
// for counters and film sums reached through pointers read from memory.
__device__ __forceinline__ void atomic_add_global(unsigned long long *p, unsigned long long v) {
  __hip_atomic_fetch_add((__attribute__((address_space(1))) unsigned long long *)p, v, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace muonx

#endif
