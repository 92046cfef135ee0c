// Element gather (move_into_order / apply_order), synthetic key generation
// and the order-independent output check.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "radix.h"

namespace ts {
namespace {

constexpr int kThreads = 256;

inline unsigned grid_for(size_t n, size_t per_thread = 1) {
  size_t b = (n + kThreads * per_thread - 1) / (kThreads * per_thread);
  return unsigned(std::max<size_t>(1, std::min<size_t>(b, 65536)));
}

// by_key.rs:315-334: out[i] = src[order[i]].  Coalesced order read and
// output write; the source read is random.  The order is u32 positions, or
// (Packed) the 16-byte (key, u64 position) records of a sort beyond
// u32::MAX, whose words 2 and 3 hold the position.
struct Packed {
  uint4 v;
};
__device__ __forceinline__ size_t order_at(const uint32_t* __restrict__ o, size_t i) { return o[i]; }
__device__ __forceinline__ size_t order_at(const Packed* __restrict__ o, size_t i) {
  const uint4 p = o[i].v;
  return size_t(p.z) | (size_t(p.w) << 32);
}

template <class E, class O>
__global__ void k_gather(const O* __restrict__ order, const E* __restrict__ src, E* __restrict__ dst, size_t n) {
  const size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x, gs = size_t(gridDim.x) * blockDim.x;
  for (size_t i = g; i < n; i += gs) dst[i] = src[order_at(order, i)];
}

template <class O>
__global__ void k_gather_bytes(const O* __restrict__ order, const unsigned char* __restrict__ src,
                               unsigned char* __restrict__ dst, size_t n, size_t eb) {
  const size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x, gs = size_t(gridDim.x) * blockDim.x;
  for (size_t i = g; i < n; i += gs) {
    const unsigned char* s = src + order_at(order, i) * eb;
    unsigned char* d = dst + i * eb;
    for (size_t b = 0; b < eb; ++b) d[b] = s[b];
  }
}

template <class O>
__global__ void k_gather_words(const O* __restrict__ order, const uint32_t* __restrict__ src,
                               uint32_t* __restrict__ dst, size_t n, size_t words) {
  const size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x, gs = size_t(gridDim.x) * blockDim.x;
  for (size_t i = g; i < n; i += gs) {
    const uint32_t* s = src + order_at(order, i) * words;
    uint32_t* d = dst + i * words;
    for (size_t b = 0; b < words; ++b) d[b] = s[b];
  }
}

// The (raw key bits, position) record of each element for a by-key sort
// beyond u32::MAX: words 0-1 the key (zero-extended), words 2-3 its index.
template <class U>
__global__ void k_pack_key_pos(const unsigned char* __restrict__ keys, size_t stride, size_t key_off,
                               uint4* __restrict__ out, size_t n) {
  const size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x, gs = size_t(gridDim.x) * blockDim.x;
  for (size_t i = g; i < n; i += gs) {
    const uint64_t k = uint64_t(*reinterpret_cast<const U*>(keys + i * stride + key_off));
    out[i] = make_uint4(uint32_t(k), uint32_t(k >> 32), uint32_t(i), uint32_t(uint64_t(i) >> 32));
  }
}

template <class O>
hipError_t gather_by(const O* order, const void* src, void* dst, size_t n, size_t eb, hipStream_t s,
                     KernelClock* clk) {
  if (n == 0) return hipSuccess;
  const uintptr_t al = reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst);
  const unsigned grid = grid_for(n);
  if (clk) clk->begin(s, "k_gather", double(n) * (double(sizeof(O)) + 2.0 * eb));
  if (eb == 16 && al % 16 == 0)
    hipLaunchKernelGGL((k_gather<uint4, O>), dim3(grid), dim3(kThreads), 0, s, order, static_cast<const uint4*>(src),
                       static_cast<uint4*>(dst), n);
  else if (eb == 8 && al % 8 == 0)
    hipLaunchKernelGGL((k_gather<uint64_t, O>), dim3(grid), dim3(kThreads), 0, s, order,
                       static_cast<const uint64_t*>(src), static_cast<uint64_t*>(dst), n);
  else if (eb == 4 && al % 4 == 0)
    hipLaunchKernelGGL((k_gather<uint32_t, O>), dim3(grid), dim3(kThreads), 0, s, order,
                       static_cast<const uint32_t*>(src), static_cast<uint32_t*>(dst), n);
  else if (eb == 2 && al % 2 == 0)
    hipLaunchKernelGGL((k_gather<uint16_t, O>), dim3(grid), dim3(kThreads), 0, s, order,
                       static_cast<const uint16_t*>(src), static_cast<uint16_t*>(dst), n);
  else if (eb == 1)
    hipLaunchKernelGGL((k_gather<uint8_t, O>), dim3(grid), dim3(kThreads), 0, s, order,
                       static_cast<const uint8_t*>(src), static_cast<uint8_t*>(dst), n);
  else if (eb % 4 == 0 && al % 4 == 0)
    hipLaunchKernelGGL(k_gather_words<O>, dim3(grid), dim3(kThreads), 0, s, order, static_cast<const uint32_t*>(src),
                       static_cast<uint32_t*>(dst), n, eb / 4);
  else
    hipLaunchKernelGGL(k_gather_bytes<O>, dim3(grid), dim3(kThreads), 0, s, order,
                       static_cast<const unsigned char*>(src), static_cast<unsigned char*>(dst), n, eb);
  if (clk) clk->end(s);
  return hipGetLastError();
}

__global__ void k_splitmix(void* out, size_t n, uint64_t seed, int kind) {
  const size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x, gs = size_t(gridDim.x) * blockDim.x;
  for (size_t i = g; i < n; i += gs) {
    const uint64_t z = splitmix64_at(seed, i);
    if (kind == 0)
      static_cast<uint32_t*>(out)[i] = uint32_t(z >> 32);
    else if (kind == 1)
      static_cast<uint64_t*>(out)[i] = z;
    else
      static_cast<uint64_t*>(out)[i] = z >> 32;
  }
}

__global__ void k_check_zero(unsigned long long* out) {
  if (threadIdx.x < 3) out[threadIdx.x] = 0;
}

// Sum and xor of fingerprint_of(bits) over the keys (a multiset fingerprint
// that any permutation keeps) and the number of adjacent pairs out of order.
// Each lane takes kCheckPer keys of a wave's 64 * kCheckPer-key stretch per
// step, all loads in flight together; a key's successor comes from the next
// lane (a shuffle), and only the wave's last lane loads the key after the
// stretch.
constexpr int kCheckPer = 8;

template <class U>
__global__ __launch_bounds__(kThreads) void k_check(const U* __restrict__ x, size_t n, int mode,
                                                    unsigned long long* out) {
  const unsigned lane = threadIdx.x & 63;
  const size_t wave = (size_t(blockIdx.x) * kThreads + threadIdx.x) >> 6;
  const size_t waves = (size_t(gridDim.x) * kThreads) >> 6;
  constexpr size_t kStretch = size_t(64) * kCheckPer;
  unsigned long long sum = 0, xr = 0, bad = 0;
  for (size_t b = wave * kStretch; b < n; b += waves * kStretch) {
    U a[kCheckPer];
#pragma unroll
    for (int r = 0; r < kCheckPer; ++r) a[r] = x[min(b + size_t(r) * 64 + lane, n - 1)];
    const size_t after = min(b + kStretch, n - 1);
    const U tail = x[after];  // the key after the stretch (lane 63 uses it)
#pragma unroll
    for (int r = 0; r < kCheckPer; ++r) {
      const size_t i = b + size_t(r) * 64 + lane;
      // The next key: the next lane's, or the next row's first (lane 0).
      U nx = __shfl_down(a[r], 1, 64);
      const U first_next = r + 1 < kCheckPer ? __shfl(a[r + 1 < kCheckPer ? r + 1 : r], 0, 64) : tail;
      if (lane == 63) nx = first_next;
      if (i < n) {
        const uint64_t h = fingerprint_of(uint64_t(a[r]));
        sum += h;
        xr ^= h;
        if (i + 1 < n && to_key(a[r], mode) > to_key(nx, mode)) ++bad;
      }
    }
  }
  for (int d = 32; d >= 1; d >>= 1) {
    sum += __shfl_xor(sum, d, 64);
    xr ^= __shfl_xor(xr, d, 64);
    bad += __shfl_xor(bad, d, 64);
  }
  // One set of atomics per workgroup: the grid is a few workgroups per CU,
  // so the three global words see little contention.
  __shared__ unsigned long long part[3][kThreads / 64];
  const unsigned w = threadIdx.x >> 6;
  if (lane == 0) {
    part[0][w] = sum;
    part[1][w] = xr;
    part[2][w] = bad;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (unsigned i = 1; i < kThreads / 64; ++i) {
      sum += part[0][i];
      xr ^= part[1][i];
      bad += part[2][i];
    }
    atomicAdd(&out[0], sum);
    atomicXor(&out[1], xr);
    atomicAdd(&out[2], bad);
  }
}

}  // namespace

hipError_t gather(const uint32_t* order, const void* src, void* dst, size_t n, size_t eb, hipStream_t s,
                  KernelClock* clk) {
  return gather_by(order, src, dst, n, eb, s, clk);
}

hipError_t gather_packed(const void* packed, const void* src, void* dst, size_t n, size_t eb, hipStream_t s,
                         KernelClock* clk) {
  return gather_by(static_cast<const Packed*>(packed), src, dst, n, eb, s, clk);
}

hipError_t pack_key_pos(int key_bytes, const void* keys, size_t stride, size_t key_offset, void* out, size_t n,
                        hipStream_t s) {
  if (n == 0) return hipSuccess;
  const unsigned grid = grid_for(n);
  const unsigned char* k = static_cast<const unsigned char*>(keys);
  uint4* o = static_cast<uint4*>(out);
  switch (key_bytes) {
    case 1: hipLaunchKernelGGL(k_pack_key_pos<uint8_t>, dim3(grid), dim3(kThreads), 0, s, k, stride, key_offset, o, n); break;
    case 2: hipLaunchKernelGGL(k_pack_key_pos<uint16_t>, dim3(grid), dim3(kThreads), 0, s, k, stride, key_offset, o, n); break;
    case 4: hipLaunchKernelGGL(k_pack_key_pos<uint32_t>, dim3(grid), dim3(kThreads), 0, s, k, stride, key_offset, o, n); break;
    case 8: hipLaunchKernelGGL(k_pack_key_pos<uint64_t>, dim3(grid), dim3(kThreads), 0, s, k, stride, key_offset, o, n); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t fill_splitmix64(void* out, size_t n, uint64_t seed, int kind, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_splitmix, dim3(grid_for(n, 4)), dim3(kThreads), 0, s, out, n, seed, kind);
  return hipGetLastError();
}

hipError_t fingerprint_zero(unsigned long long* out, hipStream_t s) {
  hipLaunchKernelGGL(k_check_zero, dim3(1), dim3(64), 0, s, out);
  return hipGetLastError();
}

hipError_t checksum(int key_bytes, int mode, const void* data, size_t n, unsigned long long* out, hipStream_t s,
                    KernelClock* clk) {
  hipLaunchKernelGGL(k_check_zero, dim3(1), dim3(64), 0, s, out);
  if (n == 0) return hipGetLastError();
  const unsigned grid = std::min(grid_for(n, 8), 2048u);
  if (clk) clk->begin(s, "k_check", double(n) * key_bytes);
  switch (key_bytes) {
    case 1:
      hipLaunchKernelGGL(k_check<uint8_t>, dim3(grid), dim3(kThreads), 0, s, static_cast<const uint8_t*>(data), n,
                         mode, out);
      break;
    case 2:
      hipLaunchKernelGGL(k_check<uint16_t>, dim3(grid), dim3(kThreads), 0, s, static_cast<const uint16_t*>(data), n,
                         mode, out);
      break;
    case 4:
      hipLaunchKernelGGL(k_check<uint32_t>, dim3(grid), dim3(kThreads), 0, s, static_cast<const uint32_t*>(data), n,
                         mode, out);
      break;
    case 8:
      hipLaunchKernelGGL(k_check<uint64_t>, dim3(grid), dim3(kThreads), 0, s, static_cast<const uint64_t*>(data), n,
                         mode, out);
      break;
    default: return hipErrorInvalidValue;
  }
  if (clk) clk->end(s);
  return hipGetLastError();
}

}  // namespace ts
