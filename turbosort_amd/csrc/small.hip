// Small-input tier: one workgroup sorts up to kSmallMax keys in LDS with a
// bitonic network.
//
// This replaces the reference's tiny and small tiers (tiny.rs:15-18,
// small.rs:11-14 -> arch/x86_64/avx2.rs:19-418): register sorting networks for
// up to 128 keys and bitonic merges of 128-key blocks up to 512, padded with
// the maximum key (avx2.rs:311-313, load_8_keys :502-513).  Here the network
// runs over a power-of-two LDS array padded with the all-ones key, which sorts
// last and is never written back; the transform is fused into the load and
// the store as in avx2.rs:452-551.  For the by-key path each key carries its
// position and pairs compare as (key, position), which makes the network's
// result the unique stable order.
#include <hip/hip_runtime.h>

#include "radix.h"

namespace ts {
namespace {

constexpr int kSmallThreads = 1024;

template <class U, bool ORDER>
__global__ __launch_bounds__(kSmallThreads) void k_small(U* keys, size_t n, int mode, const unsigned char* rec,
                                                         size_t stride, size_t key_off, uint32_t* order) {
  __shared__ U s_k[kSmallMax];
  __shared__ uint32_t s_p[ORDER ? kSmallMax : 1];
  const unsigned tid = threadIdx.x;
  unsigned N = 2;
  while (N < n) N <<= 1;
  for (unsigned i = tid; i < N; i += kSmallThreads) {
    if (i < n) {
      const U x = rec ? *reinterpret_cast<const U*>(rec + size_t(i) * stride + key_off) : keys[i];
      s_k[i] = to_key(x, mode);
      if (ORDER) s_p[i] = i;
    } else {
      s_k[i] = U(~U(0));
      if (ORDER) s_p[i] = 0xFFFFFFFFu;
    }
  }
  __syncthreads();
  for (unsigned size = 2; size <= N; size <<= 1) {
    for (unsigned half = size >> 1; half > 0; half >>= 1) {
      for (unsigned t = tid; t < N / 2; t += kSmallThreads) {
        const unsigned lo = 2 * t - (t & (half - 1));
        const unsigned hi = lo + half;
        const bool up = (lo & size) == 0;
        const U a = s_k[lo], b = s_k[hi];
        bool greater;
        if (ORDER) {
          const uint32_t pa = s_p[lo], pb = s_p[hi];
          greater = a > b || (a == b && pa > pb);
          if (greater == up) {
            s_p[lo] = pb;
            s_p[hi] = pa;
          }
        } else {
          greater = a > b;
        }
        if (greater == up) {
          s_k[lo] = b;
          s_k[hi] = a;
        }
      }
      __syncthreads();
    }
  }
  for (unsigned i = tid; i < n; i += kSmallThreads) {
    if (ORDER)
      order[i] = s_p[i];
    else
      keys[i] = from_key(s_k[i], mode);
  }
}

template <class U>
hipError_t launch_small(int mode, void* keys, size_t n, const unsigned char* rec, size_t stride, size_t key_off,
                        uint32_t* order, hipStream_t s) {
  if (order)
    hipLaunchKernelGGL((k_small<U, true>), dim3(1), dim3(kSmallThreads), 0, s, static_cast<U*>(keys), n, mode, rec,
                       stride, key_off, order);
  else
    hipLaunchKernelGGL((k_small<U, false>), dim3(1), dim3(kSmallThreads), 0, s, static_cast<U*>(keys), n, mode, rec,
                       stride, key_off, order);
  return hipGetLastError();
}

}  // namespace

hipError_t small_sort(int key_bytes, int mode, void* keys, size_t n, const unsigned char* records, size_t stride,
                      size_t key_offset, uint32_t* order, hipStream_t s, KernelClock* clk) {
  if (n > kSmallMax) return hipErrorInvalidValue;
  if (clk) clk->begin(s, "k_small", double(n) * key_bytes * 2);
  hipError_t e;
  switch (key_bytes) {
    case 1: e = launch_small<uint8_t>(mode, keys, n, records, stride, key_offset, order, s); break;
    case 2: e = launch_small<uint16_t>(mode, keys, n, records, stride, key_offset, order, s); break;
    case 4: e = launch_small<uint32_t>(mode, keys, n, records, stride, key_offset, order, s); break;
    case 8: e = launch_small<uint64_t>(mode, keys, n, records, stride, key_offset, order, s); break;
    default: e = hipErrorInvalidValue;
  }
  if (clk) clk->end(s);
  return e;
}

}  // namespace ts
