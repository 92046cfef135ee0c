// Key model shared by the host dispatcher and the gfx950 kernels.
//
// turbosort maps every sealed SortableKey type to an unsigned key of the same
// width whose unsigned order is the type's total order (key.rs:115-198).  The
// kernels never materialise those keys in HBM: the forward map is applied to
// the first load of a sort and the inverse to its last store.
#pragma once

#include <stddef.h>
#include <stdint.h>

#ifndef __HIPCC__
#define __host__
#define __device__
#define __forceinline__ inline
#endif

namespace ts {

// How a dtype's bits become a radix key.
enum KeyMode : int {
  kIdentity = 0,  // u8..u64 (key.rs:115-131)
  kSigned = 1,    // i8..i64: flip the sign bit (key.rs:135-158)
  kFloat = 2,     // f32/f64: IEEE-754 totalOrder (key.rs:162-198)
};

template <class U> __host__ __device__ __forceinline__ constexpr U top_bit() {
  return U(U(1) << (sizeof(U) * 8 - 1));
}

// Forward map.  A float with its sign clear gains the top bit (it lands above
// every negative); one with the sign set has every bit complemented (negative
// magnitudes reverse).  NaNs follow their sign: +NaN last, -NaN first.
template <class U> __host__ __device__ __forceinline__ U to_key(U bits, int mode) {
  const U top = top_bit<U>();
  if (mode == kSigned) return U(bits ^ top);
  if (mode == kFloat) {
    const U spread = (bits & top) ? U(~U(0)) : U(0);
    return U(bits ^ U(spread | top));
  }
  return bits;
}

// Inverse map.  A key with its top bit set came from a sign-clear float.
template <class U> __host__ __device__ __forceinline__ U from_key(U key, int mode) {
  const U top = top_bit<U>();
  if (mode == kSigned) return U(key ^ top);
  if (mode == kFloat) {
    const U spread = (key & top) ? U(0) : U(~U(0));
    return U(key ^ U(spread | top));
  }
  return key;
}

// A 64-bit finaliser (the SplitMix64 mixer) used by the multiset checksum.
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__host__ __device__ __forceinline__ uint64_t splitmix64_at(uint64_t seed, uint64_t i) {
  return mix64(seed + (i + 1) * 0x9E3779B97F4A7C15ull);
}

// ts_dtype -> (key width in bytes, mode).
__host__ __device__ __forceinline__ int dtype_bytes(int dt) {
  switch (dt) {
    case 0: case 4: return 1;
    case 1: case 5: return 2;
    case 2: case 6: case 8: return 4;
    case 3: case 7: case 9: return 8;
    default: return 0;
  }
}
__host__ __device__ __forceinline__ int dtype_mode(int dt) {
  if (dt >= 4 && dt <= 7) return kSigned;
  if (dt == 8 || dt == 9) return kFloat;
  return kIdentity;
}

// Device bookkeeping of one sort: the all-digit histogram, the digit bases
// (exclusive prefix of each digit's histogram), the tile tickets of each pass
// and the list of live digits, i.e. digits on which the keys are not all
// equal.  Passes are decided on the device, so a sort needs no host round trip.
struct SortMeta {
  unsigned long long hist[8][256];
  unsigned long long base[8][256];
  uint32_t tickets[8];
  uint32_t n_live;
  uint32_t live[8];
  uint32_t error;  // non-zero: a look-back wait hit its bound (never expected)
  uint32_t pad[14];
  unsigned long long check[4];  // checksum scratch: sum, xor, unsorted pairs
};

}  // namespace ts
