// Device atomics with the reference's names and contracts (core/atomic.hpp:24-51):
// cas returns whether the swap happened, write_min whether it lowered the
// value, write_add adds.  They map to native gfx950 atomics: 32-bit CAS and
// min, f32 min on bit patterns of non-negative values, f64 add.
#ifndef GG_CORE_ATOMIC_HPP
#define GG_CORE_ATOMIC_HPP

#include <hip/hip_runtime.h>

#include "type.hpp"

template <class T>
__device__ inline bool cas(T *ptr, T old_val, T new_val) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "cas: 4 or 8 byte types");
  if constexpr (sizeof(T) == 4) {
    unsigned o = __builtin_bit_cast(unsigned, old_val), n = __builtin_bit_cast(unsigned, new_val);
    return atomicCAS((unsigned *)ptr, o, n) == o;
  } else {
    unsigned long long o = __builtin_bit_cast(unsigned long long, old_val);
    unsigned long long n = __builtin_bit_cast(unsigned long long, new_val);
    return atomicCAS((unsigned long long *)ptr, o, n) == o;
  }
}

__device__ inline bool write_min(VertexId *ptr, VertexId val) {
  return atomicMin(ptr, val) > val;
}
// Distances are non-negative, so IEEE bit patterns order like the values.
__device__ inline bool write_min(float *ptr, float val) {
  if (!(val >= 0.f)) {  // general case: compare-and-swap loop
    float cur = *ptr;
    while (cur > val) {
      unsigned o = __float_as_uint(cur);
      unsigned got = atomicCAS((unsigned *)ptr, o, __float_as_uint(val));
      if (got == o) return true;
      cur = __uint_as_float(got);
    }
    return false;
  }
  return __uint_as_float(atomicMin((unsigned *)ptr, __float_as_uint(val))) > val;
}

template <class T>
__device__ inline void write_add(T *ptr, T val) {
  atomicAdd(ptr, val);
}
template <>
__device__ inline void write_add<double>(double *ptr, double val) {
  unsafeAtomicAdd(ptr, val);
}

#endif
