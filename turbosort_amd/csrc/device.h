// Small wave64 helpers shared by the kernel files.
#pragma once

#include <hip/hip_runtime.h>

#include "keys.h"

namespace ts {

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

// Inclusive prefix sum over the 64 lanes of a wave.
template <class T> __device__ __forceinline__ T wave_inclusive_add(T v) {
  const unsigned lane = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T o = __shfl_up(v, d, 64);
    if (lane >= unsigned(d)) v += o;
  }
  return v;
}

// Emits nothing, but hides x's value from the optimiser, so that values
// derived from x before this point are recomputed after it instead of being
// kept live in registers.
template <class U> __device__ __forceinline__ void opaque(U& x) {
  if constexpr (sizeof(U) == 8) {
    uint64_t t = uint64_t(x);
    asm("" : "+v"(t));
    x = U(t);
  } else {
    uint32_t t = uint32_t(x);
    asm("" : "+v"(t));
    x = U(t);
  }
}

// Forward key transform of N keys under one (wave-uniform) mode test.
template <class U, int N> __device__ __forceinline__ void to_keys(U (&k)[N], int mode) {
  if (mode == kSigned) {
#pragma unroll
    for (int j = 0; j < N; ++j) k[j] = to_key(k[j], kSigned);
  } else if (mode == kFloat) {
#pragma unroll
    for (int j = 0; j < N; ++j) k[j] = to_key(k[j], kFloat);
  }
}

}  // namespace ts
