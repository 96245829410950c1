// Device helpers shared by the end-engine kernels (line kernel, fused kernels).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "super.h"

namespace pe {

// Z_4^4 addition of packed local rotations (two bits per quadrant).
__device__ __forceinline__ int local_add(int a, int b) {
  return ((((a & 0x33) + (b & 0x33)) & 0x33) + (((a & 0xcc) + (b & 0xcc)) & 0xcc)) & 0xff;
}
__device__ __forceinline__ int dot4(const int a[4], const int b[4]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
}
// C-order strides of a 4-D shape.
__device__ __forceinline__ void strides4(const int s[4], int st[4]) {
  st[3] = 1; st[2] = s[3]; st[1] = s[2] * s[3]; st[0] = s[1] * s[2] * s[3];
}
// One super (32 bytes) as two 16-byte loads / stores.
__device__ __forceinline__ S4 ld(const uint64_t* p) {
  const ulonglong2* q = reinterpret_cast<const ulonglong2*>(p);
  const ulonglong2 a = q[0], b = q[1];
  S4 r; r.w[0] = a.x; r.w[1] = a.y; r.w[2] = b.x; r.w[3] = b.y;
  return r;
}
__device__ __forceinline__ void st(uint64_t* p, S4 v) {
  ulonglong2* q = reinterpret_cast<ulonglong2*>(p);
  q[0] = make_ulonglong2(v.w[0], v.w[1]);
  q[1] = make_ulonglong2(v.w[2], v.w[3]);
}

}  // namespace pe
