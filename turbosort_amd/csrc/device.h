// Small wave64 helpers shared by the kernel files.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

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

// Inclusive prefix sum of a 32-bit value over the 64 lanes of a wave in
// registers (DPP): shifts by 1, 2, 4 and 8 lanes inside each row of 16 (lanes
// shifted in from outside the row read 0), then the last lane of each row
// broadcast into the rows above it (row_bcast 15 into rows 1 and 3, then
// row_bcast 31 into rows 2 and 3).  The value of wave_inclusive_add without
// the LDS-crossbar shuffles.
__device__ __forceinline__ uint32_t wave_inclusive_add_dpp(uint32_t v) {
  v += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x111, 0xf, 0xf, true));   // row_shr:1
  v += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x112, 0xf, 0xf, true));   // row_shr:2
  v += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x114, 0xf, 0xf, true));   // row_shr:4
  v += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x118, 0xf, 0xf, true));   // row_shr:8
  v += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x142, 0xa, 0xf, false));  // row_bcast:15
  v += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x143, 0xc, 0xf, false));  // row_bcast:31
  return v;
}

// One step of a DPP wave scan: each lane combines its value with the one
// that CTRL moves to it; lanes whose source is outside its row, and rows
// outside the row mask RM, take `identity`.  64-bit values move as two
// halves under the same control.
template <int CTRL, int RM, class U, class Op>
__device__ __forceinline__ U dpp_combine(U v, U identity, Op op) {
  if constexpr (sizeof(U) == 8) {
    const uint32_t lo = uint32_t(
        __builtin_amdgcn_update_dpp(int(uint32_t(identity)), int(uint32_t(v)), CTRL, RM, 0xf, false));
    const uint32_t hi = uint32_t(__builtin_amdgcn_update_dpp(int(uint32_t(uint64_t(identity) >> 32)),
                                                             int(uint32_t(uint64_t(v) >> 32)), CTRL, RM, 0xf, false));
    return op(v, U((uint64_t(hi) << 32) | lo));
  } else {
    return op(v, U(uint32_t(
                     __builtin_amdgcn_update_dpp(int(uint32_t(identity)), int(uint32_t(v)), CTRL, RM, 0xf, false))));
  }
}

// An associative op with `identity` over the 64 lanes of a wave by the DPP
// scan of wave_inclusive_add_dpp: lane 63 ends with the whole wave's value,
// which every lane gets.  Every lane of the wave must be active.
template <class U, class Op> __device__ __forceinline__ U wave_reduce_dpp(U v, U identity, Op op) {
  v = dpp_combine<0x111, 0xf>(v, identity, op);  // row_shr:1
  v = dpp_combine<0x112, 0xf>(v, identity, op);  // row_shr:2
  v = dpp_combine<0x114, 0xf>(v, identity, op);  // row_shr:4
  v = dpp_combine<0x118, 0xf>(v, identity, op);  // row_shr:8
  v = dpp_combine<0x142, 0xa>(v, identity, op);  // row_bcast:15
  v = dpp_combine<0x143, 0xc>(v, identity, op);  // row_bcast:31
  if constexpr (sizeof(U) == 8) {
    const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(v)), 63));
    const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(uint64_t(v) >> 32)), 63));
    return U((uint64_t(hi) << 32) | lo);
  } else {
    return U(uint32_t(__builtin_amdgcn_readlane(int(uint32_t(v)), 63)));
  }
}
template <class U> __device__ __forceinline__ U wave_max_dpp(U v) {
  return wave_reduce_dpp(v, U(0), [](U a, U b) { return a > b ? a : b; });
}
template <class U> __device__ __forceinline__ U wave_min_dpp(U v) {
  return wave_reduce_dpp(v, U(~U(0)), [](U a, U b) { return a < b ? a : b; });
}
template <class U> __device__ __forceinline__ U wave_or_dpp(U v) {
  return wave_reduce_dpp(v, U(0), [](U a, U b) { return U(a | b); });
}

// Exclusive scan of one value per lane over the first 256 lanes of the
// workgroup (every lane of the workgroup must call it; s_wave holds four
// slots, one per wave of those lanes).
template <class T> __device__ __forceinline__ T block256_exclusive_add(T v, T* s_wave) {
  const unsigned tid = threadIdx.x, w = tid >> 6;
  T inc;
  if constexpr (std::is_same<T, uint32_t>::value)
    inc = wave_inclusive_add_dpp(v);
  else
    inc = wave_inclusive_add(v);
  if ((tid & 63) == 63 && w < 4) s_wave[w] = inc;
  __syncthreads();
  T off = 0;
  for (unsigned i = 0; i < w && i < 4; ++i) off += s_wave[i];
  __syncthreads();
  return off + inc - v;
}

// block256_exclusive_add for 32-bit values with the wave scan in registers
// and no closing barrier: the caller writes s_wave again only after a later
// barrier.
__device__ __forceinline__ uint32_t block256_exclusive_add_once(uint32_t v, uint32_t* s_wave) {
  const unsigned tid = threadIdx.x, w = tid >> 6;
  const uint32_t inc = wave_inclusive_add_dpp(v);
  if ((tid & 63) == 63 && w < 4) s_wave[w] = inc;
  __syncthreads();
  uint32_t off = 0;
  for (unsigned i = 0; i < w && i < 4; ++i) off += s_wave[i];
  return off + inc - v;
}

// The byte digit of a key at bit `shift` (a multiple of 8, uniform).
template <class U> __device__ __forceinline__ uint32_t digit_of(U k, unsigned shift) {
  if constexpr (sizeof(U) == 8) {
    const uint32_t part = (shift & 32u) ? uint32_t(uint64_t(k) >> 32) : uint32_t(k);
    return __builtin_amdgcn_ubfe(part, shift & 31u, 8u);
  } else {
    return __builtin_amdgcn_ubfe(uint32_t(k), shift, 8u);
  }
}

// x & ~(vote ^ m): keep the lanes whose vote equals this lane's bit, as one
// v_bitop3_b32 (truth table 0x90 over x, vote, m).  A builtin, not inline
// asm, so the compiler schedules the wait states a VALU read of the freshly
// written vote (an SGPR) needs.
__device__ __forceinline__ uint32_t keep_equal(uint32_t x, uint32_t vote, uint32_t m) {
  return __builtin_amdgcn_bitop3_b32(x, vote, m, 0x90);
}

// Stable rank of each of a lane's IPT keys among the keys of its wave that
// carry the same digit, counted into the wave's LDS histogram.  Lanes with
// equal digits are found with eight ballots: lane l keeps the lanes whose
// vote on every digit bit equals its own, i.e. ~(vote ^ mybit) per bit.
template <class U, int IPT, bool FULL>
__device__ __forceinline__ void rank_in_wave(const U (&k)[IPT], uint32_t (&rk)[(IPT + 1) / 2], unsigned shift,
                                             uint32_t* whist, uint32_t lb, uint32_t valid) {
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t d = digit_of(k[j], shift);
    uint32_t lo = 0xFFFFFFFFu, hi = 0xFFFFFFFFu;
    bool ok = true;
    if (!FULL) {
      ok = lb + uint32_t(j) * 64 < valid;
      const uint64_t okm = __ballot(ok);
      lo = uint32_t(okm);
      hi = uint32_t(okm >> 32);
    }
    // All eight votes first, then the combines: a vote (an SGPR) read by the
    // VALU right after the compare that wrote it costs wait states.
    uint32_t m[8];
    uint64_t vote[8];
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      m[b] = uint32_t(__builtin_amdgcn_sbfe(int32_t(d), uint32_t(b), 1u));  // 0 or ~0
      // Emits nothing; keeps the compiler from re-deriving the vote from d
      // with a second shift instead of comparing m.
      asm("" : "+v"(m[b]));
      vote[b] = __ballot(m[b] != 0);
    }
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      lo = keep_equal(lo, uint32_t(vote[b]), m[b]);
      hi = keep_equal(hi, uint32_t(vote[b] >> 32), m[b]);
    }
    const uint32_t below = __builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0u));
    const uint32_t before = whist[d];
    // Every lane of a digit group writes the same new count, so no branch is
    // needed.  A masked-off lane of the last tile holds a copy of a real key;
    // its group holds only the valid lanes sharing that digit, so it writes
    // their count (or the unchanged count when there are none).
    whist[d] = before + uint32_t(__popc(lo) + __popc(hi));
    // Ranks are below 64 * IPT, so two share a register.
    if (j & 1)
      rk[j >> 1] |= (before + below) << 16;
    else
      rk[j >> 1] = before + below;
    (void)ok;
  }
}

// The same ranks as rank_in_wave from LDS atomic adds on the wave's digit
// counters: each add returns the count before it.  The adds of one wave
// instruction to one counter take effect in lane order (a property of the
// LDS that the context checks once per device before using this ranking),
// and a wave's instructions in program order, so ranks follow the key order
// j * 64 + lane, as stability needs.  Lanes past the tile's end add nothing.
// Adds to one counter from one instruction are applied one after another, so
// a digit that at least 16 of the wave's keys share (presorted or low-entropy
// data) is counted first by one add of its lane count, and its lanes take
// their ranks from that base and their lane order among themselves.
template <class U, int IPT, bool FULL>
__device__ __forceinline__ void rank_in_wave_atomic(const U (&k)[IPT], uint32_t (&rk)[(IPT + 1) / 2], unsigned shift,
                                                    uint32_t* whist, uint32_t lb, uint32_t valid) {
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t d = digit_of(k[j], shift);
    const bool in = FULL || lb + uint32_t(j) * 64 < valid;
    const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
    const uint64_t same = __ballot(in && d == d0);
    uint32_t r = 0;
    if (__popcll(same) >= 16) {
      const uint32_t leader = uint32_t(__builtin_ctzll(same));
      uint32_t base = 0;
      if (lane_id() == leader) base = atomicAdd(&whist[d0], uint32_t(__popcll(same)));
      base = __builtin_amdgcn_readlane(base, leader);
      if ((same >> lane_id()) & 1u)
        r = base + __builtin_amdgcn_mbcnt_hi(uint32_t(same >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(same), 0u));
      else if (in)
        r = atomicAdd(&whist[d], 1u);
    } else if (in) {
      r = atomicAdd(&whist[d], 1u);
    }
    if (j & 1)
      rk[j >> 1] |= r << 16;
    else
      rk[j >> 1] = r;
  }
}

// A window of `count` 32-bit words at `base` for range-checked buffer
// accesses (gfx9 buffer resource, stride 0): a load at a byte offset past the
// window returns 0 and a store there is dropped by the hardware.  The offset
// is a VGPR; base and count are wave-uniform, so a lane's address for
// several windows of one array is one register.
struct WordWindow {
  __amdgpu_buffer_rsrc_t r;
  // base and count must be the same in every lane; they are read from the
  // first active lane, so the resource sits in scalar registers.
  __device__ __forceinline__ WordWindow(const void* base, uint32_t count) : r(make(base, count)) {}
  static __device__ __forceinline__ __amdgpu_buffer_rsrc_t make(const void* base, uint32_t count) {
    const uint64_t b = reinterpret_cast<uint64_t>(base);
    const uint64_t ub = uint64_t(uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(b))))) |
                        (uint64_t(uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(b >> 32))))) << 32);
    const int bytes = __builtin_amdgcn_readfirstlane(int(count * 4u));
    return __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void*>(ub), short(0), bytes, 0x00020000);
  }
  __device__ __forceinline__ uint32_t load(uint32_t byte_off) const {
    return __builtin_amdgcn_raw_buffer_load_b32(r, byte_off, 0, 0);
  }
  __device__ __forceinline__ void store(uint32_t x, uint32_t byte_off) const {
    __builtin_amdgcn_raw_buffer_store_b32(x, r, byte_off, 0, 0);
  }
};

// The same window for 32- or 64-bit elements (E = uint32_t or uint64_t);
// offsets are in bytes.
template <class E> struct ElemWindow {
  static_assert(sizeof(E) == 4 || sizeof(E) == 8, "32- or 64-bit elements");
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ ElemWindow(const void* base, uint32_t count)
      : r(WordWindow::make(base, count * uint32_t(sizeof(E) / 4))) {}
  __device__ __forceinline__ E load(uint32_t byte_off) const {
    if constexpr (sizeof(E) == 4) {
      return E(__builtin_amdgcn_raw_buffer_load_b32(r, byte_off, 0, 0));
    } else {
      return __builtin_bit_cast(E, __builtin_amdgcn_raw_buffer_load_b64(r, byte_off, 0, 0));
    }
  }
  __device__ __forceinline__ void store(E x, uint32_t byte_off) const {
    if constexpr (sizeof(E) == 4) {
      __builtin_amdgcn_raw_buffer_store_b32(uint32_t(x), r, byte_off, 0, 0);
    } else {
      using V = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V, x), r, byte_off, 0, 0);
    }
  }
};

// The radix key transform and its inverse as two uniform masks (no per-key
// branch): key = bits ^ (top | spread), where spread is every bit for a
// negative float (forward: taken from the bits; inverse: from the key).
template <class U> struct KeyMasks {
  U top, spread;
  __device__ __forceinline__ explicit KeyMasks(int mode)
      : top(mode != kIdentity ? top_bit<U>() : U(0)), spread(mode == kFloat ? U(~U(0)) : U(0)) {}
  __device__ __forceinline__ U sign_fill(U x) const {
    using S = typename std::conditional<sizeof(U) == 8, int64_t, int32_t>::type;
    return U(S(x) >> (sizeof(U) * 8 - 1));
  }
  __device__ __forceinline__ U to(U b) const { return U(b ^ (top | (spread & sign_fill(b)))); }
  __device__ __forceinline__ U from(U k) const { return U(k ^ (top | (spread & U(~sign_fill(k))))); }
};

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
