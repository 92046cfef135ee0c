// Batched segmented sort: every segment [offsets[s], offsets[s+1]) of a key
// array is sorted stably by key on its own, u32 values moving with their keys.
// Per segment this is sort_by_key_with_buffer keyed by the element itself
// (by_key.rs:175-235) followed by apply_order on the values
// (by_key.rs:136-144); the closest reference usage is
// chunks_exact_mut(size).for_each(turbosort::sort) (benches/sort_bench.rs:229-250).
//
// A binning kernel sorts the segment ids into size classes on the device; each
// class then runs one persistent kernel over its list:
//   * up to 64 keys: a bitonic network across the lanes of one wave (two or
//     four segments per wave below 33 and 17 keys) on (key, position) pairs.
//     The pairs are unique, so the network's result is the stable order.
//   * up to the LDS capacity of one workgroup: an LSD radix sort inside LDS
//     with 4-bit digits, skipping every digit on which the segment's keys
//     agree.  Each lane holds a run of consecutive keys; counts of earlier
//     equal digits come from 16 nibble counters packed in one 64-bit
//     register, and a workgroup scan over (digit, lane) counters gives each
//     key its stable destination.  Other shapes of these classes: byte
//     digits (k_seg_radix8), one MSD pass for keys only (k_seg_msd), and for
//     4-byte keys with values one MSD pass with stable ranks inside the bins
//     (k_seg_pairs_msd, the default there).
//   * longer segments are left in the device class for the caller, which
//     sorts them with the device-wide radix sort.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <type_traits>

#include "device.h"
#include "radix.h"

namespace ts {
namespace {

struct SegArgs {
  void* keys;
  uint32_t* vals;  // null: keys only
  const unsigned long long* offsets;
  const unsigned long long* ends;  // null: segment s ends at offsets[s + 1]
  size_t n;
  size_t n_segments;
  uint32_t* counts;  // per class, then the invalid-offsets flag
  uint32_t* lists;   // [kSegClasses][n_segments] segment ids
  uint32_t caps[kSegClasses];
  int mode;
  int atomic_rank;  // byte-digit ranks by LDS atomic adds (rank_in_wave_atomic), else ballots
  uint32_t* error;  // sticky error word, or null
  unsigned long long max_len;  // the caller's bound on segment lengths (0: none)
};

__device__ __forceinline__ unsigned long long seg_end(const SegArgs& a, size_t s) {
  return a.ends ? a.ends[s] : a.offsets[s + 1];
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
}

template <class U> __device__ __forceinline__ U wave_or(U v) { return wave_or_dpp(v); }

// Zeroes the class counters (a kernel rather than a memset so that the whole
// sort can be captured in a graph).
__global__ void k_seg_clear(uint32_t* counts) {
  if (threadIdx.x < kSegClasses + 2) counts[threadIdx.x] = 0;
}

// Appends every segment of two or more keys to its class list.  A workgroup
// takes kClassifyItems segments per lane, counts each class in LDS (a wave
// adds all its segments of one class with one LDS atomic) and reserves its
// slots in the global lists with one atomic per class.
constexpr int kClassifyThreads = 256;
constexpr int kClassifyItems = 2;

__global__ __launch_bounds__(kClassifyThreads) void k_seg_classify(SegArgs a) {
  __shared__ uint32_t lcount[kSegClasses], lbase[kSegClasses];
  const unsigned tid = threadIdx.x, lane = lane_id();
  if (tid < kSegClasses) lcount[tid] = 0;
  __syncthreads();
  const size_t chunk = size_t(blockIdx.x) * kClassifyThreads * kClassifyItems;
  int cls[kClassifyItems];
  uint32_t slot[kClassifyItems];
#pragma unroll
  for (int r = 0; r < kClassifyItems; ++r) {
    const size_t s = chunk + size_t(r) * kClassifyThreads + tid;
    int c = -1;
    if (s < a.n_segments) {
      const unsigned long long lo = a.offsets[s], hi = seg_end(a, s);
      if (hi < lo || hi > a.n) {
        a.counts[kSegClasses] = 1;
        if (a.error) atomicOr(a.error, kErrSegOffsets);
      } else if (hi - lo >= 2) {
        const unsigned long long len = hi - lo;
        if (a.max_len && len > a.max_len && a.error) atomicOr(a.error, kErrSegHint);
        c = kSegDevice;
        for (int x = 0; x < kSegDevice; ++x) {
          if (len <= a.caps[x]) {
            c = x;
            break;
          }
        }
      }
    }
    cls[r] = c;
    slot[r] = 0;
    uint64_t todo = __ballot(c >= 0);
    while (todo) {
      const int leader = __ffsll((long long)todo) - 1;
      const int cl = __shfl(c, leader, 64);
      const uint64_t m = __ballot(c == cl);
      uint32_t b = 0;
      if (int(lane) == leader) b = atomicAdd(&lcount[cl], uint32_t(__popcll(m)));
      b = __shfl(b, leader, 64);
      if (c == cl) slot[r] = b + lanes_below(m);
      todo &= ~m;
    }
  }
  __syncthreads();
  if (tid < kSegClasses) lbase[tid] = lcount[tid] ? atomicAdd(&a.counts[tid], lcount[tid]) : 0u;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kClassifyItems; ++r) {
    if (cls[r] >= 0) {
      const size_t s = chunk + size_t(r) * kClassifyThreads + tid;
      a.lists[size_t(cls[r]) * a.n_segments + lbase[cls[r]] + slot[r]] = uint32_t(s);
    }
  }
}

// ---------------------------------------------------------------------------
// Wave bitonic network: W lanes per segment, 64 / W segments per wave.

template <class U, int W, bool VALS> __global__ __launch_bounds__(256) void k_seg_bitonic(SegArgs a, int cls) {
  constexpr uint32_t G = 64 / W;
  using C = std::conditional_t<(sizeof(U) < 4), uint32_t, U>;
  const uint32_t count = a.counts[cls];
  const uint32_t* __restrict__ list = a.lists + size_t(cls) * a.n_segments;
  U* __restrict__ keys = static_cast<U*>(a.keys);
  const unsigned lane = lane_id(), g = lane / W, i = lane % W;
  const uint32_t wave = blockIdx.x * 4 + (threadIdx.x >> 6), waves = gridDim.x * 4;
  for (uint32_t base = wave * G; base < count; base += waves * G) {
    const uint32_t job = base + g;
    size_t lo = 0;
    uint32_t len = 0;
    if (job < count) {
      const uint32_t seg = list[job];
      lo = a.offsets[seg];
      len = uint32_t(seg_end(a, seg) - lo);
    }
    C k = C(U(~U(0)));
    uint32_t v = 0;
    if (i < len) {
      k = C(to_key<U>(keys[lo + i], a.mode));
      if (VALS) v = a.vals[lo + i];
    }
    uint32_t q = i;
#pragma unroll
    for (int kk = 2; kk <= W; kk <<= 1) {
#pragma unroll
      for (int j = kk >> 1; j > 0; j >>= 1) {
        const C ok = __shfl_xor(k, j, 64);
        const uint32_t oq = __shfl_xor(q, j, 64);
        const bool less = ok < k || (ok == k && oq < q);
        // The lower lane of an ascending pair and the upper lane of a
        // descending pair keep the smaller element.
        const bool keep_min = ((i & kk) == 0) == ((i & j) == 0);
        if (less == keep_min) {
          k = ok;
          q = oq;
        }
      }
    }
    uint32_t vo = 0;
    if (VALS) vo = __shfl(v, int(g * W + q), 64);
    if (i < len) {
      keys[lo + i] = from_key<U>(U(k), a.mode);
      if (VALS) a.vals[lo + i] = vo;
    }
  }
}

// ---------------------------------------------------------------------------
// LDS radix sort: one workgroup of T lanes per segment, IPT keys per lane.

template <class U, int T, int IPT, bool VALS> struct SegRadixLds {
  static constexpr int CAP = T * IPT;
  // Arrays indexed through lds_pad: one unused slot after every 32, so that
  // lanes reading runs of 16 or 32 consecutive slots hit distinct banks.
  alignas(16) U keys[CAP + CAP / 32];
  alignas(16) uint32_t vals[VALS ? CAP + CAP / 32 : 1];
  uint32_t cnt[8 * T + 8 * T / 32];  // (digit & 7, lane) counters; digit >> 3 picks the 16-bit half
  uint32_t wsum[T / 64];
  uint32_t any[2];
};

template <class U> __device__ __forceinline__ uint32_t nib(U k, int sh) { return uint32_t(k >> sh) & 15u; }

__device__ __forceinline__ uint32_t lds_pad(uint32_t i) { return i + (i >> 5); }

template <class U, int T, int IPT, bool VALS> __global__ __launch_bounds__(T) void k_seg_radix(SegArgs a, int cls) {
  static_assert(IPT <= 16 || IPT == 32, "nibble counters cover blocks of 16 keys");
  static_assert(T * IPT < 65536, "per-half counts are 16-bit");
  constexpr int NW = T / 64;
  constexpr int NB = IPT > 16 ? 2 : 1;
  constexpr int BI = IPT > 16 ? 16 : IPT;  // keys per nibble block
  constexpr int NIB = 2 * int(sizeof(U));
  __shared__ SegRadixLds<U, T, IPT, VALS> L;
  const unsigned tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const uint32_t count = a.counts[cls];
  const uint32_t* __restrict__ list = a.lists + size_t(cls) * a.n_segments;
  // Padded slots: runs of IPT (a divisor of 32) consecutive positions never
  // cross a pad, and T is a multiple of 32, so every padded index is a base
  // plus a constant.
  constexpr uint32_t TP = T + T / 32;             // lds_pad(i + T) - lds_pad(i)
  const uint32_t sb = lds_pad(tid);               // striped slot j: sb + j * TP
  const uint32_t bb = lds_pad(tid * IPT);         // blocked slot j: bb + j
  const uint32_t rb = lds_pad(tid * 8);           // raking slot x: rb + x
  static_assert(32 % IPT == 0 || IPT == 32, "runs of IPT slots must not cross a pad");
  static_assert(T % 32 == 0, "");

  for (uint32_t job = blockIdx.x; job < count; job += gridDim.x) {
    const uint32_t seg = list[job];
    const size_t lo = a.offsets[seg];
    const uint32_t len = uint32_t(seg_end(a, seg) - lo);
    U* __restrict__ keys = static_cast<U*>(a.keys) + lo;
    uint32_t* __restrict__ vals = VALS ? a.vals + lo : nullptr;

    // Striped, coalesced load (indices clamped so that every load is issued
    // at once); padding keys are the largest key and, placed after every
    // real key, stay last.
    {
      U kin[IPT];
      uint32_t vin[VALS ? IPT : 1];
#pragma unroll
      for (int j = 0; j < IPT; ++j) kin[j] = keys[min(uint32_t(j) * T + tid, len - 1)];
      if (VALS) {
#pragma unroll
        for (int j = 0; j < IPT; ++j) vin[j] = vals[min(uint32_t(j) * T + tid, len - 1)];
      }
      // Keys and values in flight together: one memory round trip.
      __builtin_amdgcn_sched_barrier(0);
      to_keys(kin, a.mode);
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const uint32_t i = uint32_t(j) * T + tid;
        L.keys[sb + j * TP] = i < len ? kin[j] : U(~U(0));
        if (VALS) L.vals[sb + j * TP] = vin[j];
      }
    }
    if (tid == 0) {
      L.any[0] = 0;
      L.any[1] = 0;
    }
    __syncthreads();

    // Blocked arrangement: lane tid holds positions tid*IPT .. tid*IPT+IPT-1.
    U k[IPT];
    uint32_t v[VALS ? IPT : 1];
    const U k0 = L.keys[0];
    U vary = 0;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint32_t p = tid * IPT + j;
      k[j] = L.keys[bb + j];
      if (VALS) v[j] = L.vals[bb + j];
      if (p < len) vary |= U(k[j] ^ k0);
    }
    {
      const uint64_t x = wave_or<uint64_t>(uint64_t(vary));
      if (lane == 0 && x) {
        atomicOr(&L.any[0], uint32_t(x));
        if (sizeof(U) == 8) atomicOr(&L.any[1], uint32_t(x >> 32));
      }
    }
    __syncthreads();
    const uint64_t varying = uint64_t(L.any[0]) | (uint64_t(L.any[1]) << 32);
    // With every key equal the stable order is the input order: nothing moves.
    if (varying == 0) {
      __syncthreads();  // every lane has read L.any before the next segment clears it
      continue;
    }

#pragma unroll 1
    for (int p = 0; p < NIB; ++p) {
      const int sh = 4 * p;
      if (((varying >> sh) & 15u) == 0) continue;  // every key agrees on this digit

      // Earlier keys of this lane with the same digit (bytes packed four to a
      // register), and per-digit totals in nibble counters.
      uint32_t before[(IPT + 3) / 4];
      uint64_t pk[NB];
      bool same[NB];
      uint32_t dsame[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        uint64_t c = 0;
        uint32_t last = 0;
#pragma unroll
        for (int jj = 0; jj < BI; ++jj) {
          const int j = b * 16 + jj;
          const uint32_t d = nib(k[j], sh);
          const uint32_t e = uint32_t(c >> (4 * d)) & 15u;
          if (j % 4 == 0) before[j / 4] = e;
          else before[j / 4] |= e << (8 * (j % 4));
          c += uint64_t(1) << (4 * d);
          last = e;
        }
        pk[b] = c;
        // A nibble holds at most 15, so sixteen equal digits wrap it; then the
        // last key saw 15 equal digits before it and the block is that digit.
        same[b] = BI == 16 && last == 15u;
        dsame[b] = nib(k[b * 16 + BI - 1], sh);
      }
      if (NB == 2) {
        // Keys of the second block also follow the equal digits of the first.
#pragma unroll
        for (int jj = 0; jj < 16; ++jj) {
          const int j = 16 + jj;
          const uint32_t d = nib(k[j], sh);
          const uint32_t t0 = same[0] ? (d == dsame[0] ? 16u : 0u) : uint32_t(pk[0] >> (4 * d)) & 15u;
          before[j / 4] += t0 << (8 * (j % 4));
        }
      }
#pragma unroll
      for (int j = 0; j < IPT; ++j) opaque(k[j]);  // recompute digits below rather than keep them
      // Counter words: low half digit dl, high half digit dl + 8.
#pragma unroll
      for (int dl = 0; dl < 8; ++dl) {
        uint32_t word = 0;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          const uint32_t lo16 = uint32_t(pk[b] >> (4 * dl)) & 15u;
          const uint32_t hi16 = uint32_t(pk[b] >> (4 * dl + 32)) & 15u;
          const uint32_t fix = (dsame[b] & 7u) == uint32_t(dl) ? (16u << (16 * (dsame[b] >> 3))) : 0u;
          word += same[b] ? fix : (lo16 | (hi16 << 16));
        }
        L.cnt[dl * TP + sb] = word;
      }
      __syncthreads();

      // Exclusive scan over the counters in (digit, lane) order: each lane
      // takes eight consecutive words.  Both halves scan at once; their
      // totals stay below 2^16, so no carry crosses between them.
      uint32_t r[8];
      uint32_t sum = 0;
#pragma unroll
      for (int x = 0; x < 8; ++x) {
        r[x] = L.cnt[rb + x];
        sum += r[x];
      }
      const uint32_t inc = wave_inclusive_add_dpp(sum);
      if (lane == 63) L.wsum[w] = inc;
      __syncthreads();
      uint32_t wpre = 0, all = 0;
#pragma unroll
      for (int x = 0; x < NW; ++x) {
        const uint32_t s = L.wsum[x];
        all += s;
        wpre += x < int(w) ? s : 0u;
      }
      uint32_t run = wpre + inc - sum;
#pragma unroll
      for (int x = 0; x < 8; ++x) {
        L.cnt[rb + x] = run;
        run += r[x];
      }
      const uint32_t below8 = all & 0xFFFFu;  // keys whose digit is below 8
      __syncthreads();

      // Stable destinations; scatter into LDS, then read back blocked.  The
      // counter reads of eight keys go out before their writes (a write may
      // alias a later read for all the compiler knows).
      constexpr int CH = IPT < 8 ? IPT : 8;
#pragma unroll
      for (int j0 = 0; j0 < IPT; j0 += CH) {
        uint32_t wd[CH];
#pragma unroll
        for (int jj = 0; jj < CH; ++jj) wd[jj] = L.cnt[(nib(k[j0 + jj], sh) & 7u) * TP + sb];
#pragma unroll
        for (int jj = 0; jj < CH; ++jj) {
          const int j = j0 + jj;
          const uint32_t d = nib(k[j], sh);
          const uint32_t base = (d & 8u) ? (wd[jj] >> 16) + below8 : (wd[jj] & 0xFFFFu);
          const uint32_t dst = base + ((before[j / 4] >> (8 * (j % 4))) & 0xFFu);
          L.keys[lds_pad(dst)] = k[j];
          if (VALS) L.vals[lds_pad(dst)] = v[j];
        }
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        k[j] = L.keys[bb + j];
        if (VALS) v[j] = L.vals[bb + j];
      }
    }

    // L.keys / L.vals hold the sorted segment; write it out striped.
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint32_t i = uint32_t(j) * T + tid;
      if (i < len) {
        keys[i] = from_key<U>(L.keys[sb + j * TP], a.mode);
        if (VALS) vals[i] = L.vals[sb + j * TP];
      }
    }
    __syncthreads();  // the next segment reuses the LDS
  }
}

// ---------------------------------------------------------------------------
// LDS radix sort with byte digits for the longer classes: each wave ranks its
// keys stably by digit with ballots (rank_in_wave), a scan over (digit, wave)
// counts places them, and the pass scatters inside LDS.  Wave w holds the
// positions w*64*IPT .. (w+1)*64*IPT-1, lane-minor, so the order (wave, step,
// lane) is the position order and the pass is stable.  Half the passes of the
// nibble kernel, for about 40 VALU operations per key and pass.

// 4-byte keys with values are staged as (key, value) pairs, so that a key and
// its value move with one 8-byte LDS access; other widths use two arrays.
template <class U, int T, int IPT, bool VALS> struct SegByteLds {
  static constexpr int CAP = T * IPT;
  static constexpr bool KV = VALS && sizeof(U) == 4;
  alignas(16) U keys[KV ? 1 : CAP];
  alignas(16) uint32_t vals[VALS && !KV ? CAP : 1];
  alignas(16) uint2 kv[KV ? CAP : 1];
  uint32_t whist[T / 64][256];  // per-wave digit counts, then bases
  uint32_t wsum[4];
  uint32_t any[2];
  __device__ __forceinline__ void put(uint32_t i, U k, uint32_t v) {
    if constexpr (KV) {
      kv[i] = make_uint2(uint32_t(k), v);
    } else {
      keys[i] = k;
      if (VALS) vals[i] = v;
    }
  }
  __device__ __forceinline__ void get(uint32_t i, U& k, uint32_t& v) const {
    if constexpr (KV) {
      const uint2 x = kv[i];
      k = U(x.x);
      v = x.y;
    } else {
      k = keys[i];
      if (VALS) v = vals[i];
    }
  }
  __device__ __forceinline__ U key(uint32_t i) const {
    if constexpr (KV)
      return U(kv[i].x);
    else
      return keys[i];
  }
};

// Exclusive scan of one value per lane over all T lanes of the workgroup.
template <int T> __device__ __forceinline__ uint32_t block_exclusive_add(uint32_t v, uint32_t* s_wave) {
  const unsigned tid = threadIdx.x, w = tid >> 6;
  const uint32_t inc = wave_inclusive_add_dpp(v);
  if ((tid & 63) == 63) s_wave[w] = inc;
  __syncthreads();
  uint32_t off = 0;
#pragma unroll
  for (int i = 0; i < T / 64; ++i)
    if (uint32_t(i) < w) off += s_wave[i];
  __syncthreads();
  return off + inc - v;
}

// The LSD byte passes over the digits set in `varying`, inside LDS.  The keys
// and values are in registers in the wave-blocked layout (wave w, step j,
// lane l holds position w*64*IPT + 64 j + l) and are back there after each
// pass; after the last pass the LDS also holds the sorted segment in order.
// `L` provides put/get of a (key, value) slot and the per-wave counters.
template <class U, int T, int IPT, bool VALS, class Lds>
__device__ __forceinline__ void lds_byte_passes(Lds& L, U (&k)[IPT], uint32_t (&v)[IPT], uint64_t varying,
                                                int atomic_rank) {
  static_assert(T % 64 == 0 && (T >= 256 || 256 % T == 0), "lanes share the 256 digits evenly");
  static_assert(64 * IPT <= 65536, "wave ranks are 16-bit");
  constexpr int NW = T / 64;
  constexpr int P = int(sizeof(U));
  const unsigned tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const uint32_t wb = w * 64 * IPT + lane;
#pragma unroll 1
  for (int p = 0; p < P; ++p) {
    const unsigned shift = 8u * unsigned(p);
    if (((varying >> shift) & 0xFFu) == 0) continue;  // every key agrees on this digit
    for (unsigned x = lane; x < 256; x += 64) L.whist[w][x] = 0;  // this wave's row only
    uint32_t rk[(IPT + 1) / 2];
    if (atomic_rank)
      rank_in_wave_atomic<U, IPT, true>(k, rk, shift, L.whist[w], 0u, 0u);
    else
      rank_in_wave<U, IPT, true>(k, rk, shift, L.whist[w], 0u, 0u);
#pragma unroll
    for (int j = 0; j < IPT; ++j) opaque(k[j]);  // recompute digits below rather than keep them
    __syncthreads();
    // Digit bases: digits in order, then waves in order.
    if constexpr (T >= 256) {
      uint32_t cnt = 0;
      if (tid < 256) {
#pragma unroll
        for (int ww = 0; ww < NW; ++ww) cnt += L.whist[ww][tid];
      }
      uint32_t run = block256_exclusive_add(cnt, L.wsum);
      if (tid < 256) {
#pragma unroll
        for (int ww = 0; ww < NW; ++ww) {
          const uint32_t c = L.whist[ww][tid];
          L.whist[ww][tid] = run;
          run += c;
        }
      }
    } else {
      constexpr int DPL = 256 / T;  // consecutive digits per lane
      uint32_t cnt = 0;
#pragma unroll
      for (int i = 0; i < DPL; ++i)
#pragma unroll
        for (int ww = 0; ww < NW; ++ww) cnt += L.whist[ww][tid * DPL + i];
      uint32_t run = block_exclusive_add<T>(cnt, L.wsum);
#pragma unroll
      for (int i = 0; i < DPL; ++i)
#pragma unroll
        for (int ww = 0; ww < NW; ++ww) {
          const uint32_t c = L.whist[ww][tid * DPL + i];
          L.whist[ww][tid * DPL + i] = run;
          run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint32_t d = digit_of(k[j], shift);
      const uint32_t dst = L.whist[w][d] + ((rk[j >> 1] >> (16 * (j & 1))) & 0xFFFFu);
      L.put(dst, k[j], v[j]);
      if ((j & 7) == 7) __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < IPT; ++j) L.get(wb + 64u * j, k[j], v[j]);
  }
}

template <class U, int T, int IPT, bool VALS>
__global__ __launch_bounds__(T) void k_seg_radix8(SegArgs a, int cls) {
  __shared__ SegByteLds<U, T, IPT, VALS> L;
  const unsigned tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const uint32_t wb = w * 64 * IPT + lane;  // this lane's slot for step j: wb + 64 j
  const uint32_t count = a.counts[cls];
  const uint32_t* __restrict__ list = a.lists + size_t(cls) * a.n_segments;

  for (uint32_t job = blockIdx.x; job < count; job += gridDim.x) {
    const uint32_t seg = list[job];
    const size_t lo = a.offsets[seg];
    const uint32_t len = uint32_t(seg_end(a, seg) - lo);
    U* __restrict__ keys = static_cast<U*>(a.keys) + lo;
    uint32_t* __restrict__ vals = VALS ? a.vals + lo : nullptr;

    // Striped, coalesced load (indices clamped so all loads issue at once);
    // padding keys are the largest key and, after every real key, stay last.
    {
      U kin[IPT];
      uint32_t vin[VALS ? IPT : 1];
#pragma unroll
      for (int j = 0; j < IPT; ++j) kin[j] = keys[min(uint32_t(j) * T + tid, len - 1)];
      if (VALS) {
#pragma unroll
        for (int j = 0; j < IPT; ++j) vin[j] = vals[min(uint32_t(j) * T + tid, len - 1)];
      }
      // Keys and values in flight together: one memory round trip.
      __builtin_amdgcn_sched_barrier(0);
      to_keys(kin, a.mode);
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const uint32_t i = uint32_t(j) * T + tid;
        L.put(i, i < len ? kin[j] : U(~U(0)), VALS ? vin[j] : 0u);
      }
    }
    if (tid == 0) {
      L.any[0] = 0;
      L.any[1] = 0;
    }
    __syncthreads();

    U k[IPT];
    uint32_t v[IPT];
    const U k0 = L.key(0);
    U vary = 0;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint32_t p = wb + 64u * j;
      L.get(p, k[j], v[j]);
      if (p < len) vary |= U(k[j] ^ k0);
    }
    {
      const uint64_t x = wave_or<uint64_t>(uint64_t(vary));
      if (lane == 0 && x) {
        atomicOr(&L.any[0], uint32_t(x));
        if (sizeof(U) == 8) atomicOr(&L.any[1], uint32_t(x >> 32));
      }
    }
    __syncthreads();
    const uint64_t varying = uint64_t(L.any[0]) | (uint64_t(L.any[1]) << 32);
    // With every key equal the stable order is the input order: nothing moves.
    if (varying == 0) {
      __syncthreads();  // every lane has read L.any before the next segment clears it
      continue;
    }
    lds_byte_passes<U, T, IPT, VALS>(L, k, v, varying, a.atomic_rank);

    // The LDS holds the sorted segment; write it out striped.
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint32_t i = uint32_t(j) * T + tid;
      if (i < len) {
        U kk;
        uint32_t vv;
        L.get(i, kk, vv);
        keys[i] = from_key<U>(kk, a.mode);
        if (VALS) vals[i] = vv;
      }
    }
    __syncthreads();  // the next segment reuses the LDS
  }
}

// ---------------------------------------------------------------------------
// Keys-only segments by one most-significant-digit pass.  Without values no
// order among equal keys is observable (equal keys are equal bits), so the
// passes need not be stable: the keys (in registers) are counted by the top
// kMsdBits bits of their offset from the segment's smallest key, the counts
// scanned, and every key scattered into LDS at its digit's next free slot
// (an LDS atomic add); every key then finds its place in its short run by
// counting the run's smaller keys.  A segment with a run longer than
// kMsdRunMax keys (a cluster of distinct keys: the count is quadratic in the
// run) is finished by the byte-digit LSD passes of k_seg_radix8 instead, over
// the bytes in which its keys differ.  The run counters are 16-bit halves of
// LDS words (segments of at most 65535 keys), two bins per word.
#ifndef TS_MSD_BITS
#define TS_MSD_BITS 13
#endif
constexpr int kMsdBits = TS_MSD_BITS, kMsdBins = 1 << kMsdBits, kMsdRunMax = 64;

__device__ __forceinline__ uint32_t msd_count(const uint32_t* cnt, uint32_t d) {
  return (cnt[d >> 1] >> ((d & 1u) * 16u)) & 0xFFFFu;
}
// Adds one to bin d's counter and returns its old value.
__device__ __forceinline__ uint32_t msd_bump(uint32_t* cnt, uint32_t d) {
  const uint32_t sh = (d & 1u) * 16u;
  return (atomicAdd(&cnt[d >> 1], 1u << sh) >> sh) & 0xFFFFu;
}

template <class U, int T, int IPT> struct SegMsdLds {
  static constexpr int CAP = T * IPT;
  alignas(16) U keys[CAP];
  union {
    uint32_t cnt[kMsdBins / 2];  // run counts, then starts, then ends (16-bit halves)
    uint32_t whist[T / 64][256];  // the LSD passes' per-wave digit counts
  };
  uint32_t wsum[T / 64];
  uint32_t flags[3];  // bits in which the keys differ (low, high word), longest run
  U kmin, kmax;
  // Keys-only slots for lds_byte_passes (no values).
  __device__ __forceinline__ void put(uint32_t i, U k, uint32_t) { keys[i] = k; }
  __device__ __forceinline__ void get(uint32_t i, U& k, uint32_t&) const { k = keys[i]; }
};

template <class U, int T, int IPT>
__global__ __launch_bounds__(T) void k_seg_msd(SegArgs a, int cls) {
  static_assert(T >= 256 && T % 64 == 0 && kMsdBins % (2 * T) == 0, "lanes must divide the bins");
  static_assert(T * IPT < 65536, "16-bit run counters");
  constexpr int NW = T / 64, PER = kMsdBins / T;
  constexpr int P = int(sizeof(U));
  constexpr int CAP = T * IPT;
  __shared__ SegMsdLds<U, T, IPT> L;
  const unsigned tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const uint32_t count = a.counts[cls];
  const uint32_t* __restrict__ list = a.lists + size_t(cls) * a.n_segments;

  for (uint32_t job = blockIdx.x; job < count; job += gridDim.x) {
    const uint32_t seg = list[job];
    const size_t lo = a.offsets[seg];
    const uint32_t len = uint32_t(seg_end(a, seg) - lo);
    U* __restrict__ keys = static_cast<U*>(a.keys) + lo;
    U k[IPT];
#pragma unroll
    for (int j = 0; j < IPT; ++j) k[j] = keys[min(uint32_t(j) * T + tid, len - 1)];
    to_keys(k, a.mode);
    const U k0 = to_key(keys[0], a.mode);
    U vary = 0, mn = k0, mx_k = k0;
#pragma unroll
    for (int j = 0; j < IPT; ++j)
      if (uint32_t(j) * T + tid < len) {
        vary |= U(k[j] ^ k0);
        mn = k[j] < mn ? k[j] : mn;
        mx_k = k[j] > mx_k ? k[j] : mx_k;
      }
    for (unsigned i = tid; i < unsigned(kMsdBins / 2); i += T) L.cnt[i] = 0;
    if (tid == 0) {
      L.flags[0] = 0;
      L.flags[1] = 0;
      L.flags[2] = 0;
      L.kmin = k0;
      L.kmax = k0;
    }
    __syncthreads();
    {
      const uint64_t x = wave_or<uint64_t>(uint64_t(vary));
      const uint64_t a0 = wave_min_dpp(uint64_t(mn)), a1 = wave_max_dpp(uint64_t(mx_k));
      if (lane == 0 && x) {
        atomicOr(&L.flags[0], uint32_t(x));
        if (P == 8) atomicOr(&L.flags[1], uint32_t(x >> 32));
        atomicMin(&L.kmin, U(a0));
        atomicMax(&L.kmax, U(a1));
      }
    }
    __syncthreads();
    const uint64_t varying = uint64_t(L.flags[0]) | (uint64_t(L.flags[1]) << 32);
    if (varying == 0) {  // every key equal: the segment is sorted as it is
      __syncthreads();
      continue;
    }
    // Digit: the top kMsdBits bits of the key's offset from the segment's
    // smallest key, so the bins split the keys' actual range.
    const U kmin = L.kmin;
    const uint64_t span = uint64_t(L.kmax - kmin);
    const int bits = 64 - __clzll((long long)span);
    const unsigned sh = bits > kMsdBits ? unsigned(bits - kMsdBits) : 0u;
    auto dig = [&](U x) { return uint32_t(uint64_t(U(x - kmin)) >> sh); };
#pragma unroll
    for (int j = 0; j < IPT; ++j)
      if (uint32_t(j) * T + tid < len) msd_bump(L.cnt, dig(k[j]));
    __syncthreads();
    // Run starts (digits in order) and the longest run.
    uint32_t c[PER], sum = 0, mx = 0;
#pragma unroll
    for (int q = 0; q < PER / 2; ++q) {
      const uint32_t wv = L.cnt[tid * (PER / 2) + q];
      c[2 * q] = wv & 0xFFFFu;
      c[2 * q + 1] = wv >> 16;
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      sum += c[q];
      mx = max(mx, c[q]);
    }
    mx = wave_max_dpp(mx);
    if (lane == 0) atomicMax(&L.flags[2], mx);
    uint32_t run = block_exclusive_add<T>(sum, L.wsum);
#pragma unroll
    for (int q = 0; q < PER / 2; ++q) {
      const uint32_t lo = run;
      run += c[2 * q];
      L.cnt[tid * (PER / 2) + q] = lo | (run << 16);
      run += c[2 * q + 1];
    }
    __syncthreads();
    // Scatter: a key's slot is its digit's next free one (any order inside a run).
#pragma unroll
    for (int j = 0; j < IPT; ++j)
      if (uint32_t(j) * T + tid < len) L.keys[msd_bump(L.cnt, dig(k[j]))] = k[j];
    __syncthreads();  // bin d's counter is now the end of run d
    if (L.flags[2] <= uint32_t(kMsdRunMax)) {
      // Each key's final slot: its run's start plus the keys of the run that
      // are smaller, or equal and earlier (independent reads, no chains);
      // then the keys are read and all move at once.
      uint32_t to[IPT];
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const uint32_t p = uint32_t(j) * T + tid;
        to[j] = p;
        if (p < len) {
          const U x = L.keys[p];
          const uint32_t d = dig(x);
          const uint32_t b = d ? msd_count(L.cnt, d - 1) : 0u, e = msd_count(L.cnt, d);
          uint32_t r = b;
          for (uint32_t q = b; q < e; ++q) {
            const U y = L.keys[q];
            r += (y < x || (y == x && q < p)) ? 1u : 0u;
          }
          to[j] = r;
        }
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < IPT; ++j) k[j] = L.keys[uint32_t(j) * T + tid];
      __syncthreads();
#pragma unroll
      for (int j = 0; j < IPT; ++j)
        if (uint32_t(j) * T + tid < len) L.keys[to[j]] = k[j];
    } else {
      // LSD byte passes over the bytes in which the keys differ; padding
      // slots hold the largest key and stay last.
      __syncthreads();  // every lane has read L.cnt and L.flags before whist reuses them
      for (unsigned i = len + tid; i < unsigned(CAP); i += T) L.keys[i] = U(~U(0));
      __syncthreads();
      const uint32_t wb = w * 64 * IPT + lane;
#pragma unroll
      for (int j = 0; j < IPT; ++j) k[j] = L.keys[wb + 64u * j];
#pragma unroll 1
      for (int p = 0; p < P; ++p) {
        const unsigned shift = 8u * unsigned(p);
        if (((varying >> shift) & 0xFFu) == 0) continue;
        for (unsigned x = lane; x < 256; x += 64) L.whist[w][x] = 0;
        uint32_t rk[(IPT + 1) / 2];
        if (a.atomic_rank)
          rank_in_wave_atomic<U, IPT, true>(k, rk, shift, L.whist[w], 0u, 0u);
        else
          rank_in_wave<U, IPT, true>(k, rk, shift, L.whist[w], 0u, 0u);
        __syncthreads();
        uint32_t cnt = 0;
        if (tid < 256) {
#pragma unroll
          for (int ww = 0; ww < NW; ++ww) cnt += L.whist[ww][tid];
        }
        uint32_t base = block256_exclusive_add(cnt, L.wsum);
        if (tid < 256) {
#pragma unroll
          for (int ww = 0; ww < NW; ++ww) {
            const uint32_t cc = L.whist[ww][tid];
            L.whist[ww][tid] = base;
            base += cc;
          }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
          const uint32_t d = digit_of(k[j], shift);
          L.keys[L.whist[w][d] + ((rk[j >> 1] >> (16 * (j & 1))) & 0xFFFFu)] = k[j];
          if ((j & 7) == 7) __builtin_amdgcn_sched_barrier(0);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < IPT; ++j) k[j] = L.keys[wb + 64u * j];
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint32_t i = uint32_t(j) * T + tid;
      if (i < len) keys[i] = from_key<U>(L.keys[i], a.mode);
    }
    __syncthreads();  // the next segment reuses the LDS
  }
}

// ---------------------------------------------------------------------------
// 4-byte keys with values by one most-significant-digit pass and a stable
// rank inside each bin.  The keys, in registers, are counted by the top BB
// bits of their offset from the segment's smallest key (2^BB bins, at least
// one per key slot; 16-bit counters, two to an LDS word).  The counter's
// atomic add gives each key a slot in its bin, where it writes one word: the
// key's offset bits below the bin over its position in the segment.  The
// offset has at most 32 - BB bits below the bin and the position at most BB
// bits, so the word fits 32 bits, and within a bin the words compare as
// (key, position) pairs do.  A key's final place is therefore its bin's
// start plus the words of its bin below its own, which is the stable order
// (by_key.rs:175-235 per segment).  Keys and values then move once, as 8-byte
// pairs, into LDS in their final order and leave in a striped store.  A
// segment with a bin of more than kMsdRunMax keys (the rank costs the square
// of the bin's size) takes the byte-digit LSD passes instead.
// threadIdx.x through an empty asm: values derived from it are recomputed
// where they are used rather than kept live across phases.
__device__ __forceinline__ unsigned fresh_lane() {
  unsigned t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

template <int T, int IPT, int BB> struct SegPairMsdLds {
  static constexpr int CAP = T * IPT, BINS = 1 << BB;
  union {
    struct {
      uint32_t word[CAP];      // per slot, in bin order: (offset bits below the bin, position)
      uint32_t cnt[BINS / 2];  // bin counts, then bin starts (16-bit halves)
    } r;
    uint2 kv[CAP];  // (key, value) pairs in their final order; the LSD passes' staging
  };
  uint32_t whist[T / 64][256];  // the LSD passes' per-wave digit counts
  uint32_t wsum[T / 64];
  uint32_t flags[2];  // bits in which the keys differ, largest bin
  uint32_t kmin, kmax;
  __device__ __forceinline__ void put(uint32_t i, uint32_t k, uint32_t v) { kv[i] = make_uint2(k, v); }
  __device__ __forceinline__ void get(uint32_t i, uint32_t& k, uint32_t& v) const {
    const uint2 x = kv[i];
    k = x.x;
    v = x.y;
  }
};

// The radix key transform of a 4-byte key and its inverse as uniform masks
// (no per-key branch): key = bits ^ (top | spread(bits)), where spread is
// every bit for a negative float; the inverse takes spread from the key.
struct KeyMasks32 {
  uint32_t top, spread;
  __device__ __forceinline__ explicit KeyMasks32(int mode)
      : top(mode != kIdentity ? 0x80000000u : 0u), spread(mode == kFloat ? ~0u : 0u) {}
  __device__ __forceinline__ uint32_t to(uint32_t b) const { return b ^ (top | (spread & uint32_t(int32_t(b) >> 31))); }
  __device__ __forceinline__ uint32_t from(uint32_t k) const {
    return k ^ (top | (spread & ~uint32_t(int32_t(k) >> 31)));
  }
};

// At least four waves per SIMD (128 VGPRs): two 512-lane or four 256-lane
// workgroups per CU, as their LDS allows.  Each workgroup walks its segments
// as a pipeline in which every memory wait falls on data that was requested
// one phase earlier: a segment's keys are requested before the previous
// segment's stores, its values after its bins are counted, the next
// segment's bounds before its values.  Loop turns issue the same memory
// operations in the same order (an empty window stores nothing), so the
// in-order memory counter lets the keys' wait leave the stores in flight.
template <int T, int IPT, int BB>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(4))) void k_seg_pairs_msd(SegArgs a, int cls) {
  using U = uint32_t;
  using L_t = SegPairMsdLds<T, IPT, BB>;
  constexpr int CAP = L_t::CAP, BINS = L_t::BINS, PER = BINS / T;
  constexpr int PB = __builtin_ctz(unsigned(CAP));
  static_assert((CAP & (CAP - 1)) == 0 && PB <= BB, "positions must fit below the bin's offset bits");
  static_assert(PER >= 2 && PER % 2 == 0 && BINS % T == 0, "each lane scans whole counter words");
  static_assert(CAP < 65536 && BB <= 16, "16-bit bin counters and starts");
  __shared__ L_t L;
  const unsigned tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const uint32_t count = a.counts[cls];
  const uint32_t* __restrict__ list = a.lists + size_t(cls) * a.n_segments;
  const KeyMasks32 km(a.mode);
  const uint32_t G = gridDim.x;
  uint32_t job = blockIdx.x;
  if (job >= count) return;
  auto rfl = [](uint32_t x) { return uint32_t(__builtin_amdgcn_readfirstlane(int(x))); };
  auto seg_hi = [&](uint32_t seg) { return a.ends ? a.ends[seg] : a.offsets[seg + 1]; };
  // The low word of a segment's end (lengths fit 32 bits): a 32-bit load,
  // so no half of a loaded register is dead while the load is in flight.
  auto seg_hi_low = [&](uint32_t seg) {
    return reinterpret_cast<const uint32_t*>(a.ends ? a.ends + seg : a.offsets + seg + 1)[0];
  };
  if (tid == 0) {
    L.kmin = ~0u;
    L.kmax = 0;
    L.flags[0] = 0;
    L.flags[1] = 0;
  }
  // Range-checked windows over a segment: positions past its end load 0
  // (never used) and store nothing.  Positions derived from the lane id are
  // recomputed in each phase (from a copy the compiler cannot see through)
  // rather than held in registers across the phases.
  size_t cur_lo, prev_lo = 0;
  uint32_t cur_len, prev_len = 0;  // prev_len 0: nothing to store
  {
    const uint32_t seg = rfl(list[job]);
    cur_lo = size_t(a.offsets[seg]);
    cur_len = rfl(uint32_t(seg_hi(seg) - cur_lo));
    cur_lo = (size_t(rfl(uint32_t(cur_lo >> 32))) << 32) | rfl(uint32_t(cur_lo));
  }
  U k[IPT];
  {
    const WordWindow kw(static_cast<U*>(a.keys) + cur_lo, cur_len);
    const unsigned t = fresh_lane();
#pragma unroll
    for (int j = 0; j < IPT; ++j) k[j] = kw.load((uint32_t(j) * T + t) * 4u);
  }
  uint32_t nseg_v = list[min(job + G, count - 1)];  // the next segment's id, in flight

  for (;;) {
    __syncthreads();  // the previous segment's pairs are in LDS in their final order
    {
      const WordWindow kw(static_cast<U*>(a.keys) + prev_lo, prev_len), vw(a.vals + prev_lo, prev_len);
      const unsigned t = fresh_lane();
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const uint32_t p = uint32_t(j) * T + t;
        const uint2 x = L.kv[p];
        kw.store(km.from(x.x), p * 4u);
        vw.store(x.y, p * 4u);
      }
      if (tid == 0) {  // every lane has read the previous segment's bounds and flags
        L.kmin = ~0u;
        L.kmax = 0;
        L.flags[0] = 0;
        L.flags[1] = 0;
      }
    }
    __syncthreads();  // the LDS is free for this segment
    if (job >= count) break;
    const uint32_t len = cur_len;
#pragma unroll
    for (int j = 0; j < IPT; ++j) k[j] = km.to(k[j]);
    {
      const unsigned t = fresh_lane();
      U mn = ~0u, mx = 0;
#pragma unroll
      for (int j = 0; j < IPT; ++j)
        if (uint32_t(j) * T + t < len) {
          mn = min(mn, k[j]);
          mx = max(mx, k[j]);
        }
      mn = wave_min_dpp(mn);
      mx = wave_max_dpp(mx);
      if (lane == 0) {
        atomicMin(&L.kmin, mn);
        atomicMax(&L.kmax, mx);
      }
    }
    for (unsigned i = tid; i < unsigned(BINS / 2); i += T) L.r.cnt[i] = 0;
    __syncthreads();
    const U kmin = L.kmin, kmax = L.kmax;
    // The next segment's bounds, in flight while this segment is sorted.
    const bool more = job + G < count;
    const uint32_t nseg = rfl(nseg_v);
    const unsigned long long nlo_v = a.offsets[nseg];
    const uint32_t nhi_v = seg_hi_low(nseg);
    // With every key equal the stable order is the input order: nothing moves.
    const bool moves = kmin != kmax;
    if (moves) {
      const int bits = 32 - __clz(int(kmax - kmin));
      const unsigned sh = bits > BB ? unsigned(bits - BB) : 0u;
      const uint32_t low_mask = (1u << sh) - 1u;  // sh <= 32 - BB < 32
      const WordWindow vw(a.vals + cur_lo, len);
      // Count the bins.
      {
        const unsigned t = fresh_lane();
#pragma unroll
        for (int j = 0; j < IPT; ++j)
          if (uint32_t(j) * T + t < len) msd_bump(L.r.cnt, uint32_t(k[j] - kmin) >> sh);
      }
      __syncthreads();
      // Bin starts (bins in order) and the largest bin; the lane's counter
      // words are read again after the scan rather than held in registers.
      {
        const unsigned t = fresh_lane();
        uint32_t sum = 0, big = 0;
#pragma unroll
        for (int q = 0; q < PER / 2; ++q) {
          const uint32_t wv = L.r.cnt[t * (PER / 2) + q];
          sum += (wv & 0xFFFFu) + (wv >> 16);
          big = max(big, max(wv & 0xFFFFu, wv >> 16));
        }
        big = wave_max_dpp(big);
        if (lane == 0) atomicMax(&L.flags[1], big);
        uint32_t run = block_exclusive_add<T>(sum, L.wsum);
#pragma unroll
        for (int q = 0; q < PER / 2; ++q) {
          const uint32_t wv = L.r.cnt[t * (PER / 2) + q];
          const uint32_t s0 = run;
          run += wv & 0xFFFFu;
          L.r.cnt[t * (PER / 2) + q] = s0 | (run << 16);
          run += wv >> 16;
        }
      }
      // The values are needed only for the final move: their loads fly
      // while the keys are placed and ranked.
      uint32_t v[IPT];
      {
        const unsigned t = fresh_lane();
#pragma unroll
        for (int j = 0; j < IPT; ++j) v[j] = vw.load((uint32_t(j) * T + t) * 4u);
      }
      __syncthreads();
      if (L.flags[1] <= uint32_t(kMsdRunMax)) {
        // Scatter the words into their bins, each at its bin's next free
        // slot (any order inside a bin); the counters end as the bins' ends.
        {
          const unsigned t = fresh_lane();
#pragma unroll
          for (int j = 0; j < IPT; ++j) {
            const uint32_t p = uint32_t(j) * T + t;
            if (p < len) {
              const uint32_t off = k[j] - kmin;
              L.r.word[msd_bump(L.r.cnt, off >> sh)] = ((off & low_mask) << PB) | p;
            }
          }
        }
        __syncthreads();
        // Each key's final place: its bin's start plus the words below its
        // own (independent reads; bins hold a few keys).
        uint32_t to[(IPT + 1) / 2];
        {
          const unsigned t = fresh_lane();
#pragma unroll
          for (int j = 0; j < IPT; ++j) {
            const uint32_t p = uint32_t(j) * T + t;
            uint32_t r = 0;
            if (p < len) {
              const uint32_t off = k[j] - kmin, d = off >> sh;
              const uint32_t mine = ((off & low_mask) << PB) | p;
              const uint32_t b = d ? msd_count(L.r.cnt, d - 1) : 0u, e = msd_count(L.r.cnt, d);
              r = b;
              for (uint32_t q = b; q < e; ++q) r += L.r.word[q] < mine ? 1u : 0u;
            }
            if (j & 1)
              to[j >> 1] |= r << 16;
            else
              to[j >> 1] = r;
            __builtin_amdgcn_sched_barrier(0);  // one key's rank at a time: bounded live ranges
          }
        }
        __syncthreads();  // the words and counters are read; the pairs take their place
        {
          const unsigned t = fresh_lane();
#pragma unroll
          for (int j = 0; j < IPT; ++j)
            if (uint32_t(j) * T + t < len) L.kv[(to[j >> 1] >> (16 * (j & 1))) & 0xFFFFu] = make_uint2(k[j], v[j]);
        }
      } else {
        // A crowded bin: LSD byte passes over the bytes in which the keys
        // differ, from the wave-blocked layout; padding slots hold the
        // largest key and, coming after every real key, stay last.
        {
          const unsigned t = fresh_lane();
          U vary = 0;
#pragma unroll
          for (int j = 0; j < IPT; ++j) {
            const uint32_t p = uint32_t(j) * T + t;
            if (p < len) vary |= k[j] ^ kmin;
            L.kv[p] = p < len ? make_uint2(k[j], v[j]) : make_uint2(~0u, 0u);
          }
          vary = wave_or_dpp(vary);
          if (lane == 0) atomicOr(&L.flags[0], vary);
        }
        __syncthreads();
        const uint32_t varying = L.flags[0];
        const uint32_t wb = w * 64 * IPT + lane;
#pragma unroll
        for (int j = 0; j < IPT; ++j) L.get(wb + 64u * j, k[j], v[j]);
        lds_byte_passes<U, T, IPT, true>(L, k, v, uint64_t(varying), a.atomic_rank);
      }
    }
    // Hand over: this segment is stored at the top of the next turn; the
    // next segment's keys (an empty window past the last) and the id of the
    // one after it are requested now.
    prev_lo = cur_lo;
    prev_len = moves ? len : 0u;
    {
      const size_t nlo = (size_t(rfl(uint32_t(nlo_v >> 32))) << 32) | rfl(uint32_t(nlo_v));
      cur_len = more ? rfl(nhi_v - uint32_t(nlo_v)) : 0u;
      cur_lo = nlo;
      const WordWindow nw(static_cast<U*>(a.keys) + cur_lo, cur_len);
      const unsigned t = fresh_lane();
#pragma unroll
      for (int j = 0; j < IPT; ++j) k[j] = nw.load((uint32_t(j) * T + t) * 4u);
    }
    nseg_v = list[min(job + 2 * G, count - 1)];
    job += G;
  }
}

// ---------------------------------------------------------------------------
// Keys-only segments by the MSD pass of k_seg_msd, as a pipeline like
// k_seg_pairs_msd's: the previous segment is stored at the top of each turn,
// after the next segment's keys were requested, so their latency hides
// behind the stores; the bounds of the segment after next are requested one
// phase ahead; the segment's smallest and largest keys need no reference
// load (every key equal: min = max).
template <class U, int T, int IPT>
__global__ __launch_bounds__(T) void k_seg_msd_pipe(SegArgs a, int cls) {
  static_assert(T >= 64 && T % 64 == 0 && kMsdBins % (2 * T) == 0, "lanes must divide the bins");
  static_assert(T * IPT < 65536, "16-bit run counters and final slots");
  constexpr int PER = kMsdBins / T;
  __shared__ SegMsdLds<U, T, IPT> L;
  const unsigned tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const uint32_t count = a.counts[cls];
  const uint32_t* __restrict__ list = a.lists + size_t(cls) * a.n_segments;
  const KeyMasks<U> km(a.mode);
  const uint32_t G = gridDim.x;
  uint32_t job = blockIdx.x;
  if (job >= count) return;
  auto rfl = [](uint32_t x) { return uint32_t(__builtin_amdgcn_readfirstlane(int(x))); };
  auto seg_hi_low = [&](uint32_t seg) {
    return reinterpret_cast<const uint32_t*>(a.ends ? a.ends + seg : a.offsets + seg + 1)[0];
  };
  auto rfl64 = [&](unsigned long long x) {
    return (size_t(rfl(uint32_t(x >> 32))) << 32) | rfl(uint32_t(x));
  };
  if (tid == 0) {
    L.kmin = U(~U(0));
    L.kmax = U(0);
    L.flags[0] = L.flags[1] = L.flags[2] = 0;
  }
  size_t cur_lo, prev_lo = 0;
  uint32_t cur_len, prev_len = 0;  // prev_len 0: nothing to store
  {
    const uint32_t seg = rfl(list[job]);
    const unsigned long long lo = a.offsets[seg];
    cur_len = rfl(seg_hi_low(seg) - uint32_t(lo));
    cur_lo = rfl64(lo);
  }
  U k[IPT];
  {
    const ElemWindow<U> kw(static_cast<U*>(a.keys) + cur_lo, cur_len);
    const unsigned t = fresh_lane();
#pragma unroll
    for (int j = 0; j < IPT; ++j) k[j] = kw.load((uint32_t(j) * T + t) * uint32_t(sizeof(U)));
  }
  uint32_t nseg_v = list[min(job + G, count - 1)];  // the next segment's id, in flight

  for (;;) {
    __syncthreads();  // the previous segment is sorted in LDS
    {
      const ElemWindow<U> kw(static_cast<U*>(a.keys) + prev_lo, prev_len);
      const unsigned t = fresh_lane();
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const uint32_t p = uint32_t(j) * T + t;
        kw.store(km.from(L.keys[p]), p * uint32_t(sizeof(U)));
      }
      if (tid == 0) {  // every lane has read the previous segment's bounds and flags
        L.kmin = U(~U(0));
        L.kmax = U(0);
        L.flags[0] = L.flags[1] = L.flags[2] = 0;
      }
    }
    __syncthreads();  // the LDS is free for this segment
    if (job >= count) break;
    const uint32_t len = cur_len;
#pragma unroll
    for (int j = 0; j < IPT; ++j) k[j] = km.to(k[j]);
    {
      const unsigned t = fresh_lane();
      U mn = U(~U(0)), mx = U(0);
#pragma unroll
      for (int j = 0; j < IPT; ++j)
        if (uint32_t(j) * T + t < len) {
          mn = k[j] < mn ? k[j] : mn;
          mx = k[j] > mx ? k[j] : mx;
        }
      mn = wave_min_dpp(mn);
      mx = wave_max_dpp(mx);
      if (lane == 0) {
        atomicMin(&L.kmin, mn);
        atomicMax(&L.kmax, mx);
      }
    }
    for (unsigned i = tid; i < unsigned(kMsdBins / 2); i += T) L.cnt[i] = 0;
    __syncthreads();
    const U kmin = L.kmin, kmax = L.kmax;
    // The next segment's bounds, in flight while this segment is sorted.
    const bool more = job + G < count;
    const uint32_t nseg = rfl(nseg_v);
    const unsigned long long nlo_v = a.offsets[nseg];
    const uint32_t nhi_v = seg_hi_low(nseg);
    // Every key equal: the segment is sorted as it is.
    const bool moves = kmin != kmax;
    if (moves) {
      const uint64_t span = uint64_t(U(kmax - kmin));
      const int bits = 64 - __clzll((long long)span);
      const unsigned sh = bits > kMsdBits ? unsigned(bits - kMsdBits) : 0u;
      auto dig = [&](U x) { return uint32_t(uint64_t(U(x - kmin)) >> sh); };
      {
        const unsigned t = fresh_lane();
#pragma unroll
        for (int j = 0; j < IPT; ++j)
          if (uint32_t(j) * T + t < len) msd_bump(L.cnt, dig(k[j]));
      }
      __syncthreads();
      // Run starts (bins in order) and the longest run; the lane's counter
      // words are read again after the scan rather than held in registers.
      {
        const unsigned t = fresh_lane();
        uint32_t sum = 0, big = 0;
#pragma unroll
        for (int q = 0; q < PER / 2; ++q) {
          const uint32_t wv = L.cnt[t * (PER / 2) + q];
          sum += (wv & 0xFFFFu) + (wv >> 16);
          big = max(big, max(wv & 0xFFFFu, wv >> 16));
        }
        big = wave_max_dpp(big);
        if (lane == 0) atomicMax(&L.flags[2], big);
        uint32_t run = block_exclusive_add<T>(sum, L.wsum);
#pragma unroll
        for (int q = 0; q < PER / 2; ++q) {
          const uint32_t wv = L.cnt[t * (PER / 2) + q];
          const uint32_t s0 = run;
          run += wv & 0xFFFFu;
          L.cnt[t * (PER / 2) + q] = s0 | (run << 16);
          run += wv >> 16;
        }
      }
      __syncthreads();
      if (L.flags[2] <= uint32_t(kMsdRunMax)) {
        // Scatter: a key's slot is its bin's next free one (any order inside
        // a bin); the counters end as the bins' ends.
        {
          const unsigned t = fresh_lane();
#pragma unroll
          for (int j = 0; j < IPT; ++j)
            if (uint32_t(j) * T + t < len) L.keys[msd_bump(L.cnt, dig(k[j]))] = k[j];
        }
        __syncthreads();
        // Each slot's final place: its bin's start plus the keys of the bin
        // that are smaller, or equal and in earlier slots; then every key
        // moves at once.
        uint32_t to[(IPT + 1) / 2];
        {
          const unsigned t = fresh_lane();
#pragma unroll
          for (int j = 0; j < IPT; ++j) {
            const uint32_t p = uint32_t(j) * T + t;
            uint32_t r = p;
            if (p < len) {
              const U x = L.keys[p];
              const uint32_t d = dig(x);
              const uint32_t b = d ? msd_count(L.cnt, d - 1) : 0u, e = msd_count(L.cnt, d);
              r = b;
              for (uint32_t q = b; q < e; ++q) {
                const U y = L.keys[q];
                r += (y < x || (y == x && q < p)) ? 1u : 0u;
              }
            }
            if (j & 1)
              to[j >> 1] |= r << 16;
            else
              to[j >> 1] = r;
            __builtin_amdgcn_sched_barrier(0);  // one slot's rank at a time: bounded live ranges
          }
        }
        __syncthreads();
        {
          const unsigned t = fresh_lane();
#pragma unroll
          for (int j = 0; j < IPT; ++j) k[j] = L.keys[uint32_t(j) * T + t];
        }
        __syncthreads();
        {
          const unsigned t = fresh_lane();
#pragma unroll
          for (int j = 0; j < IPT; ++j)
            if (uint32_t(j) * T + t < len) L.keys[(to[j >> 1] >> (16 * (j & 1))) & 0xFFFFu] = k[j];
        }
      } else {
        // A run longer than kMsdRunMax keys: LSD byte passes over the bytes
        // in which the keys differ, from the wave-blocked layout; padding
        // slots hold the largest key and stay last.
        {
          const unsigned t = fresh_lane();
          U vary = 0;
#pragma unroll
          for (int j = 0; j < IPT; ++j) {
            const uint32_t p = uint32_t(j) * T + t;
            if (p < len) vary |= U(k[j] ^ kmin);
            L.keys[p] = p < len ? k[j] : U(~U(0));
          }
          const uint64_t x = wave_or<uint64_t>(uint64_t(vary));
          if (lane == 0) {
            atomicOr(&L.flags[0], uint32_t(x));
            if (sizeof(U) == 8) atomicOr(&L.flags[1], uint32_t(x >> 32));
          }
        }
        __syncthreads();
        const uint64_t varying = uint64_t(L.flags[0]) | (uint64_t(L.flags[1]) << 32);
        const uint32_t wb = w * 64 * IPT + lane;
        uint32_t none[IPT];
#pragma unroll
        for (int j = 0; j < IPT; ++j) k[j] = L.keys[wb + 64u * j];
        lds_byte_passes<U, T, IPT, false>(L, k, none, varying, a.atomic_rank);
      }
    }
    // Hand over: this segment is stored at the top of the next turn; the
    // next segment's keys (an empty window past the last) and the id of the
    // one after it are requested now.
    prev_lo = cur_lo;
    prev_len = moves ? len : 0u;
    {
      cur_len = more ? rfl(nhi_v - uint32_t(nlo_v)) : 0u;
      cur_lo = rfl64(nlo_v);
      const ElemWindow<U> nw(static_cast<U*>(a.keys) + cur_lo, cur_len);
      const unsigned t = fresh_lane();
#pragma unroll
      for (int j = 0; j < IPT; ++j) k[j] = nw.load((uint32_t(j) * T + t) * uint32_t(sizeof(U)));
    }
    nseg_v = list[min(job + 2 * G, count - 1)];
    job += G;
  }
}

// ---------------------------------------------------------------------------
// Three-way split of long keys-only segments around a pivot key (equal keys
// are equal bits, so no order among them is observable): keys below the
// pivot go to the front of the segment, keys above to the back, and the
// middle becomes copies of the pivot.  Split segment s is
// [lo[s], lo[s] + len[s]); its chunks of kSplitChunk keys are
// chunk_start[s] .. chunk_start[s + 1] - 1.

constexpr int kSplitThreads = 256;
constexpr int kSplitIpt = 16;
constexpr uint32_t kSplitChunk = uint32_t(kSplitThreads) * kSplitIpt;

struct SplitArgs {
  void* data;
  void* tmp;
  int mode;
  uint32_t nseg;
  const unsigned long long* lo;
  const unsigned long long* len;
  const uint32_t* chunk_start;
  unsigned long long* below;  // keys below the pivot, per segment
  unsigned long long* above;
  unsigned long long* pivot;  // radix key
};

// Pivot: the median of five keys spread over the segment; counters zeroed.
template <class U> __global__ void k_split_pivot(SplitArgs a) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.nseg) return;
  const U* d = static_cast<const U*>(a.data) + a.lo[s];
  const unsigned long long len = a.len[s];
  U v[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) v[k] = to_key<U>(d[len * (2 * k + 1) / 10], a.mode);
#pragma unroll
  for (int x = 0; x < 5; ++x)
#pragma unroll
    for (int y = 0; y + 1 < 5 - x; ++y)
      if (v[y + 1] < v[y]) {
        const U t = v[y];
        v[y] = v[y + 1];
        v[y + 1] = t;
      }
  a.pivot[s] = (unsigned long long)v[2];
  a.below[s] = 0;
  a.above[s] = 0;
}

__device__ __forceinline__ uint32_t split_segment(const SplitArgs& a, uint32_t chunk) {
  uint32_t lo = 0, hi = a.nseg - 1;  // largest s with chunk_start[s] <= chunk
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (a.chunk_start[mid] <= chunk)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

template <class U> __global__ __launch_bounds__(kSplitThreads) void k_split(SplitArgs a) {
  __shared__ uint32_t wtot[kSplitThreads / 64][2];
  __shared__ unsigned long long base[2];
  const unsigned tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const uint32_t s = split_segment(a, blockIdx.x);
  const unsigned long long slo = a.lo[s], slen = a.len[s];
  const unsigned long long lo = slo + (unsigned long long)(blockIdx.x - a.chunk_start[s]) * kSplitChunk;
  const unsigned long long hi = min(slo + slen, lo + kSplitChunk);
  const U p = U(a.pivot[s]);
  const U* __restrict__ d = static_cast<const U*>(a.data);
  U* __restrict__ tmp = static_cast<U*>(a.tmp);
  U x[kSplitIpt];
#pragma unroll
  for (int j = 0; j < kSplitIpt; ++j) x[j] = d[min(lo + (unsigned long long)j * kSplitThreads + tid, hi - 1)];
  // Ranks among this wave's keys of the same side, in key order.
  uint32_t rank[kSplitIpt];
  uint32_t nb = 0, na = 0;
#pragma unroll
  for (int j = 0; j < kSplitIpt; ++j) {
    const bool valid = lo + (unsigned long long)j * kSplitThreads + tid < hi;
    const U k = to_key<U>(x[j], a.mode);
    const bool b = valid && k < p, ab = valid && p < k;
    const uint64_t mb = __ballot(b), ma = __ballot(ab);
    rank[j] = b ? nb + lanes_below(mb) : (ab ? na + lanes_below(ma) : 0xFFFFFFFFu);
    if (ab) rank[j] |= 0x80000000u;
    nb += uint32_t(__popcll(mb));
    na += uint32_t(__popcll(ma));
  }
  if (lane == 0) {
    wtot[w][0] = nb;
    wtot[w][1] = na;
  }
  __syncthreads();
  uint32_t ob = 0, oa = 0, tb = 0, ta = 0;
#pragma unroll
  for (int x2 = 0; x2 < kSplitThreads / 64; ++x2) {
    ob += x2 < int(w) ? wtot[x2][0] : 0u;
    oa += x2 < int(w) ? wtot[x2][1] : 0u;
    tb += wtot[x2][0];
    ta += wtot[x2][1];
  }
  if (tid == 0) {
    base[0] = tb ? atomicAdd(&a.below[s], (unsigned long long)tb) : 0ull;
    base[1] = ta ? atomicAdd(&a.above[s], (unsigned long long)ta) : 0ull;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kSplitIpt; ++j) {
    const uint32_t r = rank[j];
    if (r == 0xFFFFFFFFu) continue;
    if (r & 0x80000000u)
      tmp[slo + slen - 1 - (base[1] + oa + (r & 0x7FFFFFFFu))] = x[j];
    else
      tmp[slo + base[0] + ob + r] = x[j];
  }
}

template <class U> __global__ __launch_bounds__(kSplitThreads) void k_split_fill(SplitArgs a) {
  const uint32_t s = split_segment(a, blockIdx.x);
  const unsigned long long slo = a.lo[s], slen = a.len[s];
  const unsigned long long lo = slo + (unsigned long long)(blockIdx.x - a.chunk_start[s]) * kSplitChunk;
  const unsigned long long hi = min(slo + slen, lo + kSplitChunk);
  const unsigned long long nb = a.below[s], na = a.above[s];
  const U pv = from_key<U>(U(a.pivot[s]), a.mode);
  U* __restrict__ d = static_cast<U*>(a.data);
  const U* __restrict__ tmp = static_cast<const U*>(a.tmp);
  for (unsigned long long i = lo + threadIdx.x; i < hi; i += kSplitThreads) {
    const unsigned long long o = i - slo;
    d[i] = (o < nb || o >= slen - na) ? tmp[i] : pv;
  }
}

// The device class's segment bounds and the class counts in one block
// (seg_long_bounds).
__global__ void k_seg_long_bounds(SegArgs a, unsigned long long* out, uint32_t max_pairs, const uint32_t* extra,
                                  uint32_t extra_words) {
  uint32_t* hdr = reinterpret_cast<uint32_t*>(out);
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < uint32_t(kSegClasses + 2)) hdr[g] = a.counts[g];
  if (g < extra_words) hdr[kSegClasses + 2 + g] = extra[g];
  const uint32_t nbig = min(a.counts[kSegDevice], max_pairs);
  const uint32_t* __restrict__ list = a.lists + size_t(kSegDevice) * a.n_segments;
  unsigned long long* pairs = out + kSegLongHeader / 8;
  for (uint32_t i = g; i < nbig; i += gridDim.x * blockDim.x) {
    const uint32_t seg = list[i];
    const unsigned long long lo = a.offsets[seg];
    pairs[2 * size_t(i)] = lo;
    pairs[2 * size_t(i) + 1] = seg_end(a, seg) - lo;
  }
}

template <class U> hipError_t run_split(const SplitJob& j, hipStream_t s) {
  SplitArgs a;
  a.data = j.data;
  a.tmp = j.tmp;
  a.mode = j.mode;
  a.nseg = j.nseg;
  a.lo = j.lo;
  a.len = j.len;
  a.chunk_start = j.chunk_start;
  a.below = j.below;
  a.above = j.above;
  a.pivot = j.pivot;
  hipLaunchKernelGGL((k_split_pivot<U>), dim3((j.nseg + 255) / 256), dim3(256), 0, s, a);
  hipLaunchKernelGGL((k_split<U>), dim3(j.chunks), dim3(kSplitThreads), 0, s, a);
  hipLaunchKernelGGL((k_split_fill<U>), dim3(j.chunks), dim3(kSplitThreads), 0, s, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Class table.  Class c takes lengths in (cap[c-1], cap[c]]; a cap of 0 marks
// a class this key width does not use.

struct ClassShape {
  int bitonic_w;  // > 0: wave network of this width
  int threads, ipt;
  bool bytes;  // byte digits (k_seg_radix8) instead of nibbles
  bool msd;    // keys only: one most-significant-digit pass and short sorts (k_seg_msd)
  int pair_bins;  // > 0: 4-byte keys with values, one MSD pass into 2^pair_bins bins (k_seg_pairs_msd)
  bool pipe;      // with msd: the pipelined form (k_seg_msd_pipe)
};

constexpr int kMaxVariants = 6;

constexpr ClassShape pairs_msd(int threads, int ipt, int bin_bits) {
  return ClassShape{0, threads, ipt, true, false, bin_bits, false};
}
constexpr ClassShape msd_pipe(int threads, int ipt) { return ClassShape{0, threads, ipt, true, true, 0, true}; }

// Class c, variant v: variant 0 is the default; the others are the
// alternatives an autotuner may pick (same capacity, another kernel shape).
// Only 4- and 8-byte keys carry alternatives.
template <class U, bool VALS> constexpr ClassShape class_shape(int c, int v = 0) {
  // 8-byte keys with values: 16384 keys would need 192 KB of LDS.
  constexpr bool wide = sizeof(U) == 8;
  constexpr bool alt = sizeof(U) >= 4;
  constexpr bool pairs4 = sizeof(U) == 4 && VALS;
  switch (c) {
    case 0: return v == 0 ? ClassShape{16, 0, 0, false} : ClassShape{};
    case 1: return v == 0 ? ClassShape{32, 0, 0, false} : ClassShape{};
    case 2: return v == 0 ? ClassShape{64, 0, 0, false} : ClassShape{};
    case 3:
      if (v == 0) return ClassShape{0, 64, 4, false};
      if (pairs4 && v == 1) return pairs_msd(64, 4, 8);
      if (pairs4 && v == 2) return pairs_msd(128, 2, 8);
      return ClassShape{};
    case 4:
      if (v == 0) return ClassShape{0, 64, 16, false};
      if (alt && v == 1) return ClassShape{0, 256, 4, true};
      if (pairs4 && v == 2) return pairs_msd(256, 4, 10);
      if (pairs4 && v == 3) return pairs_msd(128, 8, 10);
      return ClassShape{};
    case 5:
      if (v == 0) return ClassShape{0, 128, 16, false};
      if (alt && v == 1) return ClassShape{0, 256, 8, true};
      if (pairs4 && v == 2) return pairs_msd(256, 8, 11);
      if (pairs4 && v == 3) return pairs_msd(128, 16, 11);
      return ClassShape{};
    case 6:
      if (v == 0) return ClassShape{0, 256, 16, true};
      if (!alt) return ClassShape{};
      if (v == 1) return ClassShape{0, 256, 16, false};
      if (v == 2 && !VALS) return ClassShape{0, 256, 16, true, true};
      if (v == 2 && pairs4) return pairs_msd(256, 16, 12);
      if (v == 3 && !VALS) return msd_pipe(256, 16);
      return ClassShape{};
    case 7:
      if (v == 0) return ClassShape{0, 512, 16, true};
      if (!alt) return ClassShape{};
      if (v == 1) return ClassShape{0, 512, 16, false};
      if (v == 2) return ClassShape{0, 1024, 8, true};
      if (v == 3 && !VALS) return ClassShape{0, 512, 16, true, true};
      if (v == 3 && pairs4) return pairs_msd(512, 16, 13);
      if (v == 4 && pairs4) return pairs_msd(512, 16, 14);
      if (v == 4 && !VALS) return msd_pipe(512, 16);
      return ClassShape{};
    case 8:
      // 8-byte keys with values: 11264 pairs (132 KiB) are what one
      // workgroup's LDS holds.
      if (wide && VALS) return v == 0 ? ClassShape{0, 1024, 11, true} : ClassShape{};
      // 8-byte keys without values: the MSD shape only, with 17 keys per
      // lane (17408 keys in 136 KiB), which the other shapes' LDS cannot hold.
      if (wide) return v == 0 ? ClassShape{0, 1024, 17, true, true} : v == 1 ? msd_pipe(1024, 17) : ClassShape{};
      if (v == 0) return ClassShape{0, 1024, 16, true};
      if (!alt) return ClassShape{};
      if (v == 1) return ClassShape{0, 512, 32, false};
      if (v == 2 && !VALS) return ClassShape{0, 1024, 16, true, true};
      if (v == 2 && pairs4) return pairs_msd(1024, 16, 14);
      if (v == 3 && pairs4) return pairs_msd(1024, 16, 15);
      if (v == 3 && !VALS) return msd_pipe(1024, 16);
      return ClassShape{};
    default: return ClassShape{};
  }
}

template <class U, bool VALS> constexpr bool class_exists(int c, int v) {
  const ClassShape s = class_shape<U, VALS>(c, v);
  return s.bitonic_w > 0 || s.threads > 0;
}

template <class U, bool VALS> constexpr int class_variants(int c) {
  int n = 0;
  while (n < kMaxVariants && class_exists<U, VALS>(c, n)) ++n;
  return n;
}

template <class U, bool VALS> constexpr uint32_t class_cap(int c) {
  const ClassShape s = class_shape<U, VALS>(c);
  return s.bitonic_w ? uint32_t(s.bitonic_w) : uint32_t(s.threads * s.ipt);
}

template <class U, bool VALS, int C, int V>
void launch_shape(const SegArgs& a, int num_cus, hipStream_t s, KernelClock* clk) {
  constexpr ClassShape sh = class_shape<U, VALS>(C, V);
  if constexpr (sh.bitonic_w > 0) {
    if (clk) clk->begin(s, "k_seg_bitonic", 0);
    hipLaunchKernelGGL((k_seg_bitonic<U, sh.bitonic_w, VALS>), dim3(num_cus * 8), dim3(256), 0, s, a, C);
    if (clk) clk->end(s);
  } else if constexpr (sh.threads > 0 && sh.pair_bins > 0) {
    constexpr size_t lds = sizeof(SegPairMsdLds<sh.threads, sh.ipt, sh.pair_bins>);
    constexpr int by_lds = int((160 * 1024) / lds);
    constexpr int by_threads = 2048 / sh.threads;
    constexpr int per_cu = std::max(1, std::min(by_lds, by_threads));
    if (clk) clk->begin(s, "k_seg_pairs_msd", 0);
    hipLaunchKernelGGL((k_seg_pairs_msd<sh.threads, sh.ipt, sh.pair_bins>), dim3(num_cus * per_cu), dim3(sh.threads),
                       0, s, a, C);
    if (clk) clk->end(s);
  } else if constexpr (sh.threads > 0 && sh.msd && sh.pipe) {
    constexpr size_t lds = sizeof(SegMsdLds<U, sh.threads, sh.ipt>);
    constexpr int by_lds = int((160 * 1024) / lds);
    constexpr int by_threads = 2048 / sh.threads;
    constexpr int per_cu = std::max(1, std::min(by_lds, by_threads));
    if (clk) clk->begin(s, "k_seg_msd_pipe", 0);
    hipLaunchKernelGGL((k_seg_msd_pipe<U, sh.threads, sh.ipt>), dim3(num_cus * per_cu), dim3(sh.threads), 0, s, a,
                       C);
    if (clk) clk->end(s);
  } else if constexpr (sh.threads > 0 && sh.msd) {
    constexpr size_t lds = sizeof(SegMsdLds<U, sh.threads, sh.ipt>);
    constexpr int by_lds = int((160 * 1024) / lds);
    constexpr int by_threads = 2048 / sh.threads;
    constexpr int per_cu = std::max(1, std::min(by_lds, by_threads));
    if (clk) clk->begin(s, "k_seg_msd", 0);
    hipLaunchKernelGGL((k_seg_msd<U, sh.threads, sh.ipt>), dim3(num_cus * per_cu), dim3(sh.threads), 0, s, a, C);
    if (clk) clk->end(s);
  } else if constexpr (sh.threads > 0 && sh.bytes) {
    constexpr size_t lds = sizeof(SegByteLds<U, sh.threads, sh.ipt, VALS>);
    constexpr int by_lds = int((160 * 1024) / lds);
    constexpr int by_threads = 2048 / sh.threads;
    constexpr int per_cu = std::max(1, std::min(by_lds, by_threads));
    if (clk) clk->begin(s, "k_seg_radix8", 0);
    hipLaunchKernelGGL((k_seg_radix8<U, sh.threads, sh.ipt, VALS>), dim3(num_cus * per_cu), dim3(sh.threads), 0, s,
                       a, C);
    if (clk) clk->end(s);
  } else if constexpr (sh.threads > 0) {
    constexpr size_t lds = sizeof(SegRadixLds<U, sh.threads, sh.ipt, VALS>);
    constexpr int by_lds = int((160 * 1024) / lds);
    constexpr int by_threads = 2048 / sh.threads;
    constexpr int per_cu = std::max(1, std::min(by_lds, by_threads));
    if (clk) clk->begin(s, "k_seg_radix", 0);
    hipLaunchKernelGGL((k_seg_radix<U, sh.threads, sh.ipt, VALS>), dim3(num_cus * per_cu), dim3(sh.threads), 0, s,
                       a, C);
    if (clk) clk->end(s);
  }
}

template <class U, bool VALS, int C>
void launch_class(const SegArgs& a, int num_cus, hipStream_t s, KernelClock* clk, int v) {
  if (v == 1) {
    if constexpr (class_exists<U, VALS>(C, 1)) return launch_shape<U, VALS, C, 1>(a, num_cus, s, clk);
  } else if (v == 2) {
    if constexpr (class_exists<U, VALS>(C, 2)) return launch_shape<U, VALS, C, 2>(a, num_cus, s, clk);
  } else if (v == 3) {
    if constexpr (class_exists<U, VALS>(C, 3)) return launch_shape<U, VALS, C, 3>(a, num_cus, s, clk);
  } else if (v == 4) {
    if constexpr (class_exists<U, VALS>(C, 4)) return launch_shape<U, VALS, C, 4>(a, num_cus, s, clk);
  } else if (v == 5) {
    if constexpr (class_exists<U, VALS>(C, 5)) return launch_shape<U, VALS, C, 5>(a, num_cus, s, clk);
  }
  launch_shape<U, VALS, C, 0>(a, num_cus, s, clk);
}

template <class U, bool VALS> const char* variant_name(int c, int v) {
  static char buf[kSegClasses][kMaxVariants][64];
  if (c < 0 || c >= kSegDevice || v < 0 || v >= class_variants<U, VALS>(c)) return nullptr;
  ClassShape sh{};
  switch (c) {
    case 0: sh = class_shape<U, VALS>(0, v); break;
    case 1: sh = class_shape<U, VALS>(1, v); break;
    case 2: sh = class_shape<U, VALS>(2, v); break;
    case 3: sh = class_shape<U, VALS>(3, v); break;
    case 4: sh = class_shape<U, VALS>(4, v); break;
    case 5: sh = class_shape<U, VALS>(5, v); break;
    case 6: sh = class_shape<U, VALS>(6, v); break;
    case 7: sh = class_shape<U, VALS>(7, v); break;
    default: sh = class_shape<U, VALS>(8, v); break;
  }
  if (sh.bitonic_w)
    std::snprintf(buf[c][v], sizeof buf[c][v], "bitonic, %d lanes per segment", sh.bitonic_w);
  else if (sh.pair_bins)
    std::snprintf(buf[c][v], sizeof buf[c][v], "one %d-bit MSD pass, ranks in bins, %d lanes x %d keys",
                  sh.pair_bins, sh.threads, sh.ipt);
  else if (sh.msd)
    std::snprintf(buf[c][v], sizeof buf[c][v], "one %d-bit MSD pass%s, %d lanes x %d keys", kMsdBits,
                  sh.pipe ? ", pipelined" : "", sh.threads, sh.ipt);
  else
    std::snprintf(buf[c][v], sizeof buf[c][v], "%s digits, %d lanes x %d keys", sh.bytes ? "byte" : "4-bit",
                  sh.threads, sh.ipt);
  return buf[c][v];
}

template <class U, bool VALS> hipError_t run_segmented(const SegJob& j, hipStream_t s, KernelClock* clk) {
  SegArgs a;
  a.keys = j.keys;
  a.vals = VALS ? j.vals : nullptr;
  a.offsets = j.offsets;
  a.ends = j.ends;
  a.n = j.n;
  a.n_segments = j.n_segments;
  a.counts = j.counts;
  a.lists = j.lists;
  a.mode = j.mode;
  a.atomic_rank = j.atomic_rank ? 1 : 0;
  a.error = j.error;
  a.max_len = j.max_len;
  for (int c = 0; c < kSegClasses; ++c) a.caps[c] = c < kSegDevice ? class_cap<U, VALS>(c) : 0u;
  if (j.stage != 2) {
    hipLaunchKernelGGL(k_seg_clear, dim3(1), dim3(64), 0, s, j.counts);
    const size_t per_block = size_t(kClassifyThreads) * kClassifyItems;
    const unsigned cgrid = unsigned((j.n_segments + per_block - 1) / per_block);
    if (clk) clk->begin(s, "k_seg_classify", 0);
    hipLaunchKernelGGL(k_seg_classify, dim3(cgrid), dim3(kClassifyThreads), 0, s, a);
    if (clk) clk->end(s);
  }
  if (j.stage == 1) return hipGetLastError();
  // Longest segments first, so the short classes fill in behind them.
  if (j.class_mask >> 8 & 1u) launch_class<U, VALS, 8>(a, j.num_cus, s, clk, j.variant[8]);
  if (j.class_mask >> 7 & 1u) launch_class<U, VALS, 7>(a, j.num_cus, s, clk, j.variant[7]);
  if (j.class_mask >> 6 & 1u) launch_class<U, VALS, 6>(a, j.num_cus, s, clk, j.variant[6]);
  if (j.class_mask >> 5 & 1u) launch_class<U, VALS, 5>(a, j.num_cus, s, clk, j.variant[5]);
  if (j.class_mask >> 4 & 1u) launch_class<U, VALS, 4>(a, j.num_cus, s, clk, j.variant[4]);
  if (j.class_mask >> 3 & 1u) launch_class<U, VALS, 3>(a, j.num_cus, s, clk, j.variant[3]);
  if (j.class_mask >> 2 & 1u) launch_class<U, VALS, 2>(a, j.num_cus, s, clk, j.variant[2]);
  if (j.class_mask >> 1 & 1u) launch_class<U, VALS, 1>(a, j.num_cus, s, clk, j.variant[1]);
  if (j.class_mask >> 0 & 1u) launch_class<U, VALS, 0>(a, j.num_cus, s, clk, j.variant[0]);
  return hipGetLastError();
}

// One segment of n keys sorted inside one workgroup by the keys-only class
// kernel for its length: the plain sort's tier between the one-workgroup
// network (kSmallMax keys) and the device radix sort.  A one-lane-group
// kernel writes the class list of that single segment into `area`.
__global__ void k_one_segment(uint32_t* counts, uint32_t* lists, unsigned long long* offsets, int cls,
                              unsigned long long n) {
  const unsigned t = threadIdx.x;
  if (t < unsigned(kSegClasses + 2)) counts[t] = t == unsigned(cls) ? 1u : 0u;
  if (t < unsigned(kSegClasses)) lists[t] = 0;
  if (t == 0) {
    offsets[0] = 0;
    offsets[1] = n;
  }
}

template <class U>
hipError_t run_one_segment(int mode, void* keys, size_t n, void* area, const uint8_t* variant, bool atomic_rank,
                           uint32_t* error, hipStream_t s, KernelClock* clk) {
  SegArgs a = {};
  auto* counts = static_cast<uint32_t*>(area);
  auto* lists = counts + (kSegClasses + 2);
  auto* offsets = reinterpret_cast<unsigned long long*>(lists + kSegClasses + 2);  // 8-byte aligned
  a.keys = keys;
  a.vals = nullptr;
  a.offsets = offsets;
  a.n = n;
  a.n_segments = 1;
  a.counts = counts;
  a.lists = lists;
  a.mode = mode;
  a.atomic_rank = atomic_rank ? 1 : 0;
  a.error = error;
  int cls = -1;
  for (int c = 0; c < kSegClasses; ++c) {
    a.caps[c] = c < kSegDevice ? class_cap<U, false>(c) : 0u;
    if (cls < 0 && c < kSegDevice && n <= a.caps[c]) cls = c;
  }
  if (cls < 0 || n <= kSmallMax) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_one_segment, dim3(1), dim3(64), 0, s, counts, lists, offsets, cls, (unsigned long long)n);
  switch (cls) {  // one workgroup: the grid of a one-CU device
    case 6: launch_class<U, false, 6>(a, 1, s, clk, variant[6]); break;
    case 7: launch_class<U, false, 7>(a, 1, s, clk, variant[7]); break;
    case 8: launch_class<U, false, 8>(a, 1, s, clk, variant[8]); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <class U, bool VALS> uint32_t max_cap() {
  uint32_t m = 0;
  for (int c = 0; c < kSegDevice; ++c) m = std::max(m, class_cap<U, VALS>(c));
  return m;
}

}  // namespace

uint32_t seg_class_cap(int key_bytes, bool vals, int size_class) {
  if (size_class < 0 || size_class >= kSegDevice) return 0;
  switch (key_bytes) {
    case 1: return vals ? class_cap<uint8_t, true>(size_class) : class_cap<uint8_t, false>(size_class);
    case 2: return vals ? class_cap<uint16_t, true>(size_class) : class_cap<uint16_t, false>(size_class);
    case 4: return vals ? class_cap<uint32_t, true>(size_class) : class_cap<uint32_t, false>(size_class);
    case 8: return vals ? class_cap<uint64_t, true>(size_class) : class_cap<uint64_t, false>(size_class);
    default: return 0;
  }
}

uint32_t seg_lds_max_len(int key_bytes, bool vals) {
  switch (key_bytes) {
    case 1: return vals ? max_cap<uint8_t, true>() : max_cap<uint8_t, false>();
    case 2: return vals ? max_cap<uint16_t, true>() : max_cap<uint16_t, false>();
    case 4: return vals ? max_cap<uint32_t, true>() : max_cap<uint32_t, false>();
    case 8: return vals ? max_cap<uint64_t, true>() : max_cap<uint64_t, false>();
    default: return 0;
  }
}

int seg_variants(int key_bytes, bool vals, int size_class) {
  if (size_class < 0 || size_class >= kSegDevice) return 0;
  auto count = [&](auto u, auto v) {
    using U = decltype(u);
    constexpr bool V = decltype(v)::value;
    switch (size_class) {
      case 0: return class_variants<U, V>(0);
      case 1: return class_variants<U, V>(1);
      case 2: return class_variants<U, V>(2);
      case 3: return class_variants<U, V>(3);
      case 4: return class_variants<U, V>(4);
      case 5: return class_variants<U, V>(5);
      case 6: return class_variants<U, V>(6);
      case 7: return class_variants<U, V>(7);
      default: return class_variants<U, V>(8);
    }
  };
  using T_ = std::true_type;
  using F_ = std::false_type;
  switch (key_bytes) {
    case 1: return vals ? count(uint8_t(), T_()) : count(uint8_t(), F_());
    case 2: return vals ? count(uint16_t(), T_()) : count(uint16_t(), F_());
    case 4: return vals ? count(uint32_t(), T_()) : count(uint32_t(), F_());
    case 8: return vals ? count(uint64_t(), T_()) : count(uint64_t(), F_());
    default: return 0;
  }
}

const char* seg_variant_name(int key_bytes, bool vals, int size_class, int variant) {
  switch (key_bytes) {
    case 1: return vals ? variant_name<uint8_t, true>(size_class, variant) : variant_name<uint8_t, false>(size_class, variant);
    case 2: return vals ? variant_name<uint16_t, true>(size_class, variant) : variant_name<uint16_t, false>(size_class, variant);
    case 4: return vals ? variant_name<uint32_t, true>(size_class, variant) : variant_name<uint32_t, false>(size_class, variant);
    case 8: return vals ? variant_name<uint64_t, true>(size_class, variant) : variant_name<uint64_t, false>(size_class, variant);
    default: return nullptr;
  }
}

hipError_t lds_sort(int key_bytes, int mode, void* keys, size_t n, void* area, const uint8_t* variant,
                    bool atomic_rank, uint32_t* error, hipStream_t s, KernelClock* clk) {
  switch (key_bytes) {
    case 1: return run_one_segment<uint8_t>(mode, keys, n, area, variant, atomic_rank, error, s, clk);
    case 2: return run_one_segment<uint16_t>(mode, keys, n, area, variant, atomic_rank, error, s, clk);
    case 4: return run_one_segment<uint32_t>(mode, keys, n, area, variant, atomic_rank, error, s, clk);
    case 8: return run_one_segment<uint64_t>(mode, keys, n, area, variant, atomic_rank, error, s, clk);
    default: return hipErrorInvalidValue;
  }
}

uint32_t split_chunk() { return kSplitChunk; }

hipError_t seg_long_bounds(const SegJob& j, unsigned long long* out, uint32_t max_pairs, const uint32_t* extra,
                           uint32_t extra_words, hipStream_t s) {
  if (kSegLongHeader < (kSegClasses + 2 + extra_words) * 4) return hipErrorInvalidValue;
  SegArgs a = {};
  a.offsets = j.offsets;
  a.ends = j.ends;
  a.n = j.n;
  a.n_segments = j.n_segments;
  a.counts = j.counts;
  a.lists = j.lists;
  const unsigned grid = unsigned(std::min<size_t>(std::max<size_t>((size_t(max_pairs) + 255) / 256, 1), 1024));
  hipLaunchKernelGGL(k_seg_long_bounds, dim3(grid), dim3(256), 0, s, a, out, max_pairs, extra, extra_words);
  return hipGetLastError();
}

hipError_t split_segments(const SplitJob& j, hipStream_t s) {
  if (j.nseg == 0) return hipSuccess;
  switch (j.key_bytes) {
    case 1: return run_split<uint8_t>(j, s);
    case 2: return run_split<uint16_t>(j, s);
    case 4: return run_split<uint32_t>(j, s);
    case 8: return run_split<uint64_t>(j, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t segmented_sort(const SegJob& j, hipStream_t s, KernelClock* clk) {
  const bool v = j.vals != nullptr;
  switch (j.key_bytes) {
    case 1: return v ? run_segmented<uint8_t, true>(j, s, clk) : run_segmented<uint8_t, false>(j, s, clk);
    case 2: return v ? run_segmented<uint16_t, true>(j, s, clk) : run_segmented<uint16_t, false>(j, s, clk);
    case 4: return v ? run_segmented<uint32_t, true>(j, s, clk) : run_segmented<uint32_t, false>(j, s, clk);
    case 8: return v ? run_segmented<uint64_t, true>(j, s, clk) : run_segmented<uint64_t, false>(j, s, clk);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace ts
