// Device LSD radix sort for gfx950: one fused all-digit histogram pass, a
// one-block digit scan that also decides which digits are live, then one
// one-sweep scatter pass per live digit with decoupled look-back.
//
// Reference behaviour (turbosort 0.2.2): radix/mod.rs:205-271 (byte-digit
// LSD, live-pass skipping), histogram.rs:15-45 (all digits in one read),
// prefix_sum.rs:17-24 (exclusive scan), scatter.rs:26-40 (stable scatter by
// one digit).  The output is the unique ascending order of the radix keys.
//
// Code-generation notes for gfx950.  Loads are issued unconditionally (the
// last tile clamps its indices and masks the extra lanes afterwards) so that
// all of a lane's loads are in flight together; the key transform is applied
// in a separate loop under a uniform branch.  In-tile indices are 32-bit.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device.h"
#include "radix.h"

namespace ts {
namespace {

template <class U> __device__ __forceinline__ U first_lane(U x) {
  if constexpr (sizeof(U) == 8) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(x));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(uint64_t(x) >> 32));
    return U((uint64_t(hi) << 32) | lo);
  } else {
    return U(__builtin_amdgcn_readfirstlane(uint32_t(x)));
  }
}

// ------------------------------------------------------------------ init
// gate: the second sort after a divert (runs only when the finish set
// m->fallback, and keeps that flag).
__global__ void k_init(SortMeta* m, int P, int gate) {
  if (gate && !m->fallback) return;
  const unsigned g = blockIdx.x * blockDim.x + threadIdx.x, gs = gridDim.x * blockDim.x;
  for (unsigned i = g; i < unsigned(P) * kSegments * 256; i += gs) {
    (&m->pair[0][0][0])[i] = 0;
    (&m->pair_pos[0][0][0])[i] = 0;
  }
  if (blockIdx.x != 0) return;
  if (threadIdx.x < 8 * kSegments) (&m->seg_tickets[0][0])[threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    m->n_live = 0;
    m->error = 0;
    m->breaks_up = 0;
    m->breaks_down = 0;
    m->reverse = 0;
    m->skip = 0;
    m->finish_shift = 0;
    m->all_live = 0;
    if (!gate) {
      m->fallback = 0;
      m->diverted = 0;
    }
  }
}

// ------------------------------------------------------------- histogram
// All digit histograms in one read of the keys (histogram.rs:15-45), split by
// the segment each key falls in for that digit's pass: digit 0 by position
// segment (first_segment() keys each), digit p > 0 by the top three bits of
// digit p-1.  Each workgroup counts one contiguous chunk of one position
// segment into LDS and adds its counts to the global table once.  The
// forward key transform is fused into the load; the by-key path (REC) reads
// keys out of records and writes the transformed key array the scatter
// passes start from (unless the passes carry the records themselves:
// keys_out null).
//
// A digit (with its segment bits) on which all of a wave's keys of one step
// agree -- constant digits, sorted runs -- is counted with one LDS add
// instead of 64 colliding atomics.  The grid also clears the first look-back
// row.  Every workgroup's keys lie in one position segment, so its counts
// summed over value segments also give each digit's position-segment table
// (pair_pos), which a first executed pass on a digit above 0 uses (divert,
// or dead low digits).
#ifndef TS_HIST_THREADS
#define TS_HIST_THREADS 256
#endif
#ifndef TS_HIST_THREADS_WIDE
#define TS_HIST_THREADS_WIDE 1024
#endif
#ifndef TS_HIST_UNROLL_NARROW
#define TS_HIST_UNROLL_NARROW 16
#endif
#ifndef TS_HIST_UNROLL_WIDE
#define TS_HIST_UNROLL_WIDE 16
#endif
#ifndef TS_HIST_BLOCKS_PER_CU
#define TS_HIST_BLOCKS_PER_CU 4
#endif
#ifndef TS_HIST_BLOCKS_PER_CU_WIDE
#define TS_HIST_BLOCKS_PER_CU_WIDE 1
#endif
#ifndef TS_HIST_BLOCKS_PER_CU_LEARNED
#define TS_HIST_BLOCKS_PER_CU_LEARNED 4
#endif
// Fewest keys a histogram workgroup takes, down to a floor of workgroups
// per position segment.  Each workgroup ends by adding its nonzero (pass,
// segment, digit) counts to the global tables, so sorts of a few million
// keys spread over four workgroups per CU paid mostly for those adds
// (MI355X, u32: 2^22 keys 181 -> 143 us, 2^21 128 -> 108 us; 2^18 keys want
// the floor).  Sorts of 2^24 keys and more, and 8-byte keys, are unchanged.
#ifndef TS_HIST_MIN_KEYS
#define TS_HIST_MIN_KEYS 16384
#endif
#ifndef TS_HIST_MIN_BLOCKS_PER_SEG
#define TS_HIST_MIN_BLOCKS_PER_SEG 8
#endif
// Measured on MI355X at 2^28 keys: 16 keys per lane per step; 8-byte keys
// (eight digit tables per workgroup) prefer one 1024-lane workgroup per CU
// (0.55 ms against 0.65 ms for two of 256 lanes), 4-byte keys four of 256.
template <class U> struct HistUnroll {
  static constexpr int value = sizeof(U) >= 8 ? TS_HIST_UNROLL_WIDE : TS_HIST_UNROLL_NARROW;
  static constexpr int blocks_per_cu = sizeof(U) >= 8 ? TS_HIST_BLOCKS_PER_CU_WIDE : TS_HIST_BLOCKS_PER_CU;
  static constexpr int threads = sizeof(U) >= 8 ? TS_HIST_THREADS_WIDE : TS_HIST_THREADS;
};

// The digits of a pass come from the radix key itself, or (LEARNED) from the
// bucket (at most 24 bits) the learned model gives the key: three byte
// digits, of which the scan skips those no bucket id uses.
template <class U, bool LEARNED> struct DigitSource {
  using T = U;
  static constexpr int P = int(sizeof(U));
};
template <class U> struct DigitSource<U, true> {
  using T = uint32_t;
  static constexpr int P = 3;
};

// Copies the model's knots into LDS as knot_pair entries (every lane calls;
// a barrier follows).
__device__ __forceinline__ void load_knots(unsigned long long* lds, const LearnedModel* m) {
  for (unsigned i = threadIdx.x; i < unsigned(kKnots); i += blockDim.x) lds[i] = knot_pair(m->knots, int(i));
}

template <class U>
__device__ __forceinline__ uint32_t bucket_of(U k, const unsigned long long* knots, const LearnedCoord& m,
                                              uint32_t shift) {
  return learned_cdf_pairs(learned_coord((unsigned long long)k, m), knots) >> shift;
}

// Presorted scan (presorted.rs:15-65) on the device: flags whether some
// adjacent pair of radix keys decreases (breaks_up) and whether some pair
// increases (breaks_down).  Workgroup b checks the pairs (i-1, i) of its
// contiguous chunk, each lane 16 consecutive keys per step, and stops as
// soon as it has seen both kinds of break itself or, looking every 16 steps,
// finds both flags set: random input ends within a step, and only ordered
// (or almost ordered) input is read in full.
constexpr int kOrderKeys = 16;
#ifndef TS_ORDER_BLOCKS_PER_CU
#define TS_ORDER_BLOCKS_PER_CU 1
#endif
template <class U, bool REC>
__global__ __launch_bounds__(256) void k_order_check(const U* __restrict__ keys, const unsigned char* __restrict__ rec,
                                                     size_t stride, size_t key_off, size_t n, int mode,
                                                     SortMeta* m) {
  const size_t per = (n + gridDim.x - 1) / gridDim.x;
  const size_t b0 = max(size_t(1), size_t(blockIdx.x) * per), b1 = min(n, size_t(blockIdx.x + 1) * per);
  auto key = [&](size_t i) {
    const U x = REC ? *reinterpret_cast<const U*>(rec + i * stride + key_off) : keys[i];
    return to_key(x, mode);
  };
  bool up = false, down = false;
  uint32_t step = 0;
  for (size_t base = b0; base < b1; base += size_t(256) * kOrderKeys, ++step) {
    const size_t i0 = base + size_t(threadIdx.x) * kOrderKeys;
    U k[kOrderKeys + 1];
#pragma unroll
    for (int r = 0; r <= kOrderKeys; ++r) k[r] = key(min(i0 - 1 + size_t(r), n - 1));
#pragma unroll
    for (int r = 1; r <= kOrderKeys; ++r) {
      if (i0 - 1 + size_t(r) < b1) {
        up |= k[r - 1] > k[r];
        down |= k[r - 1] < k[r];
      }
    }
    up = __syncthreads_or(up);
    down = __syncthreads_or(down);
    if (up && down) break;
    if ((step & 15) == 15) {
      const bool both = __hip_atomic_load(&m->breaks_up, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &&
                        __hip_atomic_load(&m->breaks_down, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (__syncthreads_or(both)) return;  // the same answer in every lane
    }
  }
  if (threadIdx.x == 0) {
    if (up) atomicOr(&m->breaks_up, 1u);
    if (down) atomicOr(&m->breaks_down, 1u);
  }
}

template <class U, bool REC, bool LEARNED = false>
__global__ __launch_bounds__(HistUnroll<U>::threads) void k_hist(const U* __restrict__ keys, const unsigned char* __restrict__ rec,
                                                       size_t stride, size_t key_off, U* __restrict__ keys_out,
                                                       size_t n, int mode, size_t seg0, uint32_t blocks_per_seg,
                                                       size_t per_block, SortMeta* m, uint32_t* zero_ptr,
                                                       size_t zero_words, const LearnedModel* model,
                                                       int check_order, int gate, unsigned long long* fp) {
  using D = typename DigitSource<U, LEARNED>::T;
  constexpr int P = DigitSource<U, LEARNED>::P;
  constexpr int UN = HistUnroll<U>::value;
  constexpr int kHistThreads = HistUnroll<U>::threads;
  if (gate && !m->fallback) return;
  // Keys already in order (k_order_check; check_order as k_scan's presorted):
  // no pass runs, so no counts.
  if (!LEARNED && check_order && (!m->breaks_up || (check_order == 2 && !m->breaks_down))) return;
  __shared__ uint32_t h[P * kSegments * 256];
  __shared__ unsigned long long knots[LEARNED ? kKnots : 1];
  const unsigned tid = threadIdx.x, lane = tid & 63;
  LearnedCoord mc{};
  if constexpr (LEARNED) {
    mc = coord_of(*model);
    load_knots(knots, model);
  }
  for (unsigned i = tid; i < P * kSegments * 256; i += kHistThreads) h[i] = 0;
  {
    const size_t g = size_t(blockIdx.x) * kHistThreads + tid, gs = size_t(gridDim.x) * kHistThreads;
    for (size_t z = g; z < zero_words; z += gs) zero_ptr[z] = 0;
  }
  __syncthreads();
  const uint32_t s0 = blockIdx.x / blocks_per_seg;  // this block's position segment
  const size_t seg_lo = min(n, size_t(s0) * seg0), seg_hi = min(n, seg_lo + seg0);
  const size_t begin = min(seg_hi, seg_lo + size_t(blockIdx.x % blocks_per_seg) * per_block);
  const size_t end = min(seg_hi, begin + per_block);
  unsigned long long fp_sum = 0, fp_xor = 0;
  for (size_t base = begin; base < end; base += size_t(kHistThreads) * UN) {
    const bool full = base + size_t(kHistThreads) * UN <= end;
    U k[UN];
#pragma unroll
    for (int r = 0; r < UN; ++r) {
      size_t i = base + size_t(r) * kHistThreads + tid;
      if (!full) i = i < end ? i : end - 1;
      k[r] = REC ? *reinterpret_cast<const U*>(rec + i * stride + key_off) : keys[i];
    }
    if (LEARNED && fp) {
      // The input's multiset fingerprint for the learned sort's judge.
#pragma unroll
      for (int r = 0; r < UN; ++r) {
        if (full || base + size_t(r) * kHistThreads + tid < end) {
          const uint64_t hh = fingerprint_of(uint64_t(k[r]));
          fp_sum += hh;
          fp_xor ^= hh;
        }
      }
    }
    to_keys(k, mode);
    if (REC && keys_out) {
#pragma unroll
      for (int r = 0; r < UN; ++r) {
        const size_t i = base + size_t(r) * kHistThreads + tid;
        if (full || i < end) keys_out[i] = k[r];
      }
    }
    D dk[UN];
#pragma unroll
    for (int r = 0; r < UN; ++r) {
      if constexpr (LEARNED)
        dk[r] = bucket_of(k[r], knots, mc, mc.bucket_shift);
      else
        dk[r] = k[r];
    }
    // Bits in which the wave's valid digit sources differ from its first.
    const D k0 = first_lane(dk[0]);
    uint64_t diff = 0;
    uint32_t wave_count = 64u * UN;
#pragma unroll
    for (int r = 0; r < UN; ++r) {
      const size_t i = base + size_t(r) * kHistThreads + tid;
      if (full || i < end) diff |= uint64_t(dk[r] ^ k0);
    }
    if (!full) {
      wave_count = 0;
#pragma unroll
      for (int r = 0; r < UN; ++r)
        wave_count += uint32_t(__popcll(__ballot(base + size_t(r) * kHistThreads + tid < end)));
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
      // Digit p and, for p > 0, the segment bits (top three of digit p-1).
      const uint64_t care = p == 0 ? 0xFFull : (0x7FFull << (8 * p - 3));
      uint32_t* hp = h + p * kSegments * 256;
      if (__ballot((diff & care) != 0) == 0) {
        const uint32_t seg = p == 0 ? s0 : uint32_t(uint64_t(k0) >> (8 * p - 3)) & 7u;
        if (lane == 0) atomicAdd(&hp[seg * 256 + (uint32_t(uint64_t(k0) >> (8 * p)) & 0xFFu)], wave_count);
      } else if (p > 0 && __ballot((diff & (0xFFull << (8 * p))) != 0) == 0) {
        // The digit is the same in every key of the step but the segment
        // bits differ (a dead digit above a live one): count the segments in
        // 16-bit fields of two registers instead of 64 colliding LDS adds
        // onto at most eight counters (at most 64 * UN < 2^16 per field).
        uint64_t lo4 = 0, hi4 = 0;
#pragma unroll
        for (int r = 0; r < UN; ++r) {
          const size_t i = base + size_t(r) * kHistThreads + tid;
          if (full || i < end) {
            const uint32_t seg = uint32_t(uint64_t(dk[r]) >> (8 * p - 3)) & 7u;
            const uint64_t one = 1ull << (16 * (seg & 3u));
            if (seg < 4)
              lo4 += one;
            else
              hi4 += one;
          }
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
          lo4 += __shfl_xor(lo4, d, 64);
          hi4 += __shfl_xor(hi4, d, 64);
        }
        if (lane == 0) {
          const uint32_t dig = uint32_t(uint64_t(k0) >> (8 * p)) & 0xFFu;
#pragma unroll
          for (int sg = 0; sg < kSegments; ++sg) {
            const uint32_t c = uint32_t(((sg < 4 ? lo4 : hi4) >> (16 * (sg & 3))) & 0xFFFFu);
            if (c) atomicAdd(&hp[sg * 256 + dig], c);
          }
        }
      } else {
#pragma unroll
        for (int r = 0; r < UN; ++r) {
          const size_t i = base + size_t(r) * kHistThreads + tid;
          if (full || i < end) {
            const uint32_t seg = p == 0 ? s0 : uint32_t(uint64_t(dk[r]) >> (8 * p - 3)) & 7u;
            atomicAdd(&hp[seg * 256 + (uint32_t(uint64_t(dk[r]) >> (8 * p)) & 0xFFu)], 1u);
          }
        }
      }
    }
  }
  if (LEARNED && fp) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
      fp_sum += __shfl_xor(fp_sum, d, 64);
      fp_xor ^= __shfl_xor(fp_xor, d, 64);
    }
    if (lane == 0) {
      atomicAdd(&fp[0], fp_sum);
      atomicXor(&fp[1], fp_xor);
    }
  }
  __syncthreads();
  for (unsigned i = tid; i < P * kSegments * 256; i += kHistThreads) {
    const uint32_t c = h[i];
    if (c) atomicAdd(&m->pair[0][0][0] + i, (unsigned long long)c);
  }
  // Digit 0's table is already by position segment.
  for (unsigned i = tid + 256; i < P * 256; i += kHistThreads) {
    const unsigned p = i >> 8, x = i & 255u;
    uint32_t c = 0;
#pragma unroll
    for (int sg = 0; sg < kSegments; ++sg) c += h[(p * kSegments + sg) * 256 + x];
    if (c) atomicAdd(&m->pair_pos[p][s0][x], (unsigned long long)c);
  }
}

// ------------------------------------------------------------------ scan
// log2(x) in eighths of a bit, rounded down (the three bits after the leading
// one stand in for the fraction); x > 0.  The same estimate as the
// reference's divert decision (divert.rs:95-104).
__device__ __forceinline__ uint32_t log2_eighths(unsigned long long x) {
  const uint32_t e = 63u - uint32_t(__clzll((long long)x));
  const uint32_t frac = e >= 3 ? uint32_t(x >> (e - 3)) & 7u : uint32_t(x << (3 - e)) & 7u;
  return 8u * e + frac;
}

// Divert (divert.rs:19-54): walking down from the most significant live
// digit, each is credited with log2(n / its largest bucket) bits; once the
// credit reaches log2(n) + 2 bits, the live digits below are left unsorted,
// provided that skips at least three.
constexpr uint32_t kDivertSlackBits = 2;
constexpr uint32_t kDivertMinSkip = 3;

// Per digit p: its histogram (the segments' counts summed), the digit bases
// (prefix_sum.rs:17-24), its largest bucket and whether it is live (one bin
// holding every key makes it dead, as the reference decides from the first
// key's bin, radix/mod.rs:229-239).  Then which live digits run (all, or the
// top ones after a divert), and for each executed pass its segments, their
// tiles and each (segment, digit value)'s first output position.
//
// Segments of executed pass i on digit d: for i = 0 the position segments
// (the input order), counted in pair_pos (pair for d = 0); for i > 0 the
// top three bits of digit d-1, counted in pair.  The previous executed digit
// is d-1, or every digit between them is dead, so that digit d-1 is constant
// and its value segments collapse into one segment holding all keys in input
// order -- correct whichever digits are live.
// presorted: 0 ignores the order check; 1 skips every pass when the keys are
// already non-decreasing (finish_presorted, presorted.rs:15-27); 2 also
// reverses non-increasing keys in place instead of sorting them (keys only:
// equal keys are equal bits, presorted.rs:10-14).
// divert: 1 lets the scan skip low digits (keys-only sorts that finish the
// runs with k_finish).  gate: the second sort after a divert, which runs only
// when the finish set m->fallback.
__global__ __launch_bounds__(256) void k_scan(SortMeta* m, size_t n, int P, size_t seg0, uint32_t tile,
                                              int presorted, int divert, int gate) {
  if (gate && !m->fallback) return;
  __shared__ unsigned long long s_wave[4];
  __shared__ unsigned long long s_hist[8][256], s_base[8][256];
  __shared__ unsigned long long s_maxw[4];
  __shared__ unsigned long long s_maxb[8];
  __shared__ uint32_t s_tiles[kSegments];
  __shared__ uint32_t live[8];
  const unsigned tid = threadIdx.x;
  uint32_t n_live = 0;  // the same in every lane
  for (int p = 0; p < P; ++p) {
    unsigned long long tot = 0;
#pragma unroll
    for (int s = 0; s < kSegments; ++s) tot += m->pair[p][s][tid];
    const unsigned long long e = block256_exclusive_add(tot, s_wave);
    s_hist[p][tid] = tot;
    s_base[p][tid] = e;
    m->hist[p][tid] = tot;
    m->base[p][tid] = e;
    unsigned long long mx = tot;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) mx = max(mx, (unsigned long long)__shfl_xor(mx, d, 64));
    if ((tid & 63) == 0) s_maxw[tid >> 6] = mx;
    const int full = __syncthreads_or(tot == (unsigned long long)n);
    if (tid == 0) {
      const unsigned long long b = max(max(s_maxw[0], s_maxw[1]), max(s_maxw[2], s_maxw[3]));
      s_maxb[p] = b;
      m->maxb[p] = b;
    }
    if (!full) {
      if (tid == 0) live[n_live] = uint32_t(p);
      ++n_live;
    }
  }
  __syncthreads();
  if (presorted >= 1 && m->breaks_up == 0) {
    n_live = 0;
  } else if (presorted == 2 && m->breaks_down == 0) {
    n_live = 0;
    if (tid == 0) m->reverse = 1;
  }
  uint32_t skip = 0;
  if (divert && n_live > kDivertMinSkip) {
    const uint32_t log_len = log2_eighths(n), need = log_len + 8 * kDivertSlackBits;
    uint32_t have = 0;
    for (int i = int(n_live) - 1; i >= int(kDivertMinSkip); --i) {
      have += log_len - log2_eighths(s_maxb[live[i]]);
      if (have >= need) {
        skip = uint32_t(i);
        break;
      }
    }
  }
  const uint32_t n_exec = n_live - skip;
  if (tid == 0) {
    m->all_live = n_live;
    m->n_live = n_exec;
    m->skip = skip;
    m->finish_shift = skip ? 8u * live[skip] : 0u;
    if (!gate) m->diverted = skip;
    for (uint32_t i = 0; i < n_live; ++i) m->live[i] = live[i];
  }
  // Geometry of each executed pass, by execution order.
  for (uint32_t i = 0; i < n_exec; ++i) {
    const uint32_t d = live[skip + i];
    PassGeo* g = &m->geo[i];
    if (tid < kSegments) {
      unsigned long long st, ct;
      if (i == 0) {
        st = min((unsigned long long)n, (unsigned long long)tid * seg0);
        ct = min((unsigned long long)n, st + seg0) - st;
      } else {
        ct = 0;
        for (int x = 0; x < 32; ++x) ct += s_hist[d - 1][tid * 32 + x];
        st = s_base[d - 1][tid * 32];
      }
      g->start[tid] = st;
      g->count[tid] = ct;
      s_tiles[tid] = uint32_t((ct + tile - 1) / tile);
    }
    {
      const unsigned long long(*tab)[256] = (i == 0 && d > 0) ? m->pair_pos[d] : m->pair[d];
      unsigned long long run = s_base[d][tid];
#pragma unroll
      for (int s = 0; s < kSegments; ++s) {
        m->pbase[i][s][tid] = run;
        run += tab[s][tid];
      }
    }
    __syncthreads();
    if (tid == 0) {
      uint32_t slot = 0;
      for (int s = 0; s < kSegments; ++s) {
        g->slot[s] = slot;
        g->nt[s] = s_tiles[s];
        slot += s_tiles[s];
      }
      g->digit = d;
      g->total = slot;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------- one sweep
template <class S> struct Flags;
template <> struct Flags<uint32_t> {
  static constexpr int kBits = 30;
};
template <> struct Flags<uint64_t> {
  static constexpr int kBits = 62;
};

template <class U> struct SweepArgs {
  U* keys_a;
  U* keys_b;
  uint32_t* vals_a;
  uint32_t* vals_b;
  uint32_t* final_vals;
  size_t n;
  int mode_in;   // forward transform applied by the first pass
  int mode_out;  // inverse transform applied by the last pass
  const SortMeta* meta;
  const PassGeo* geo;  // meta->geo, read-only in the scatter
  uint32_t* seg_tickets;  // [pass][segment]
  uint32_t* error;  // kErrLookback is ORed in when a look-back wait exceeds its bound
  uint32_t num_tiles;
  const LearnedModel* model;  // learned partition: digits of the key's bucket
  int atomic_rank;            // rank with LDS atomics (rank_in_wave_atomic), else ballots
  int gate;                   // the second sort after a divert: run only when meta->fallback is set
  // Carried records (RW > 0): ping-pong halves (recs_a holds the input and
  // the result), and where each record's key lies: 32-bit word key_word,
  // shifted right by key_shift bits; rec_mode is the key's transform.
  void* recs_a;
  void* recs_b;
  uint32_t key_word, key_shift;
  int rec_mode;
};

// A carried record: RW 32-bit words.
template <int RW> struct RecOf {
  using T = uint32_t;
};
template <> struct RecOf<2> {
  using T = uint2;
};
template <> struct RecOf<4> {
  using T = uint4;
};

// The key stored in a carried record at 32-bit word `word`, bit `shift`
// (keys lie aligned to their width, so an 8-byte key starts at word 0 or 2).
// The selects are uniform: the word index is the same for every record.
template <class U, int RW>
__device__ __forceinline__ U rec_key(const typename RecOf<RW>::T& r, uint32_t word, uint32_t shift) {
  if constexpr (RW == 2) {
    if constexpr (sizeof(U) == 8)
      return U((uint64_t(r.y) << 32) | r.x);
    else
      return U((word ? r.y : r.x) >> shift);
  } else if constexpr (RW == 4) {
    if constexpr (sizeof(U) == 8) {
      const uint32_t lo = word ? r.z : r.x, hi = word ? r.w : r.y;
      return U((uint64_t(hi) << 32) | lo);
    } else {
      const uint32_t w = word == 0 ? r.x : word == 1 ? r.y : word == 2 ? r.z : r.w;
      return U(w >> shift);
    }
  } else {
    return U(0);
  }
}

// About two seconds of s_sleep(1) polling.
constexpr uint32_t kSpinLimit = 1u << 26;

// LDS of one scatter tile, declared once in the kernel.
template <class U, bool VALS, int IPT, bool LEARNED = false, bool SMALL = false, int RW = 0> struct SweepLds {
  static constexpr int NW = SweepShape<sizeof(U), LEARNED, VALS, SMALL, RW>::threads / 64;
  static constexpr int TILE = SweepShape<sizeof(U), LEARNED, VALS, SMALL, RW>::threads * IPT;
  uint32_t whist[NW][256];  // per-wave digit counts, then staging bases
  U keys[RW ? 1 : TILE];
  typename RecOf<RW>::T recs[RW ? TILE : 1];
  uint32_t vals[VALS ? TILE : 1];
  unsigned long long knots[LEARNED ? kKnots : 1];
  unsigned long long gofs[256];
  uint32_t wave[NW];
  uint32_t tcnt[256];
  uint32_t ticket;
};

// A persistent workgroup's claim of its next tile: (segment << 24) | round,
// or none.
constexpr uint32_t kNoClaim = 0xFFFFFFFFu;

// Everything a tile does after it knows where it is: load, publish its digit
// counts, rank, look back, stage, store.  FULL tiles (all but each segment's
// last) run without any per-key bounds test.  LOOP: the workgroup goes on to
// another tile afterwards (a persistent grid); it then clears the counters
// for that tile and claims the next round of segment sg0 (nt0 rounds) while
// its stores go out, and leaves the claim, or kNoClaim, in L.ticket for the
// barrier that ends the tile.
template <class U, bool VALS, class S, bool FULL, bool LEARNED, bool SMALL, int RW, bool LOOP = false>
__device__ __forceinline__ void run_tile(
    SweepLds<U, VALS, SweepShape<sizeof(U), LEARNED, VALS, SMALL, RW>::ipt, LEARNED, SMALL, RW>& L,
    const SweepArgs<U>& a, const LearnedCoord& mc, S* __restrict__ status, int pass, unsigned shift, bool first,
    bool last, size_t tile_base, uint32_t valid, size_t first_slot, size_t slot, bool first_round,
    unsigned long long digit_base, uint32_t sg0 = 0, uint32_t nt0 = 0) {
  using R = typename RecOf<RW>::T;
  constexpr int IPT = SweepShape<sizeof(U), LEARNED, VALS, SMALL, RW>::ipt;
  constexpr int THREADS = SweepShape<sizeof(U), LEARNED, VALS, SMALL, RW>::threads;
  constexpr int NW = THREADS / 64;
  constexpr int TILE = THREADS * IPT;
  constexpr int VB = Flags<S>::kBits;
  constexpr S kAgg = S(1) << VB, kPre = S(2) << VB, kMask = kAgg - 1;
  constexpr bool kEarlyClaim = LOOP;
  unsigned tid_ = threadIdx.x;
  // In a loop over tiles the lane index is taken afresh for each tile, so
  // that the lane's addresses are not hoisted out of the loop and kept live
  // (scratch spills) across it.
  if constexpr (LOOP) asm volatile("" : "+v"(tid_));
  const unsigned tid = tid_, lane = tid & 63, w = tid >> 6;
  const uint32_t lb = w * (64 * IPT) + lane;  // this lane's first key in the tile
  const U* __restrict__ src = (pass & 1) ? a.keys_b : a.keys_a;
  U* __restrict__ dst = (pass & 1) ? a.keys_a : a.keys_b;
  const uint32_t* __restrict__ vsrc = (pass & 1) ? a.vals_b : a.vals_a;
  uint32_t* __restrict__ vdst = (last && a.final_vals) ? a.final_vals : ((pass & 1) ? a.vals_a : a.vals_b);
  S* row = status + size_t(pass & 1) * a.num_tiles * 256;
  const bool dl = tid < 256;  // this lane also carries digit tid
  auto in_tile = [&](int j) { return FULL || lb + uint32_t(j) * 64 < valid; };

  U k[IPT];
  uint32_t v[VALS ? IPT : 1];
  R r[RW ? IPT : 1];
  {
    if constexpr (RW > 0) {
      const R* __restrict__ rs = static_cast<const R*>((pass & 1) ? a.recs_b : a.recs_a) + tile_base;
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const uint32_t li = lb + uint32_t(j) * 64;
        r[j] = rs[FULL || li < valid ? li : valid - 1];
      }
#pragma unroll
      for (int j = 0; j < IPT; ++j) k[j] = to_key(rec_key<U, RW>(r[j], a.key_word, a.key_shift), a.rec_mode);
    } else {
      const U* __restrict__ ts = src + tile_base;
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const uint32_t li = lb + uint32_t(j) * 64;
        k[j] = ts[FULL || li < valid ? li : valid - 1];
      }
    }
    if (VALS) {
      if (first) {
#pragma unroll
        for (int j = 0; j < IPT; ++j) v[j] = uint32_t(tile_base) + lb + uint32_t(j) * 64;
      } else {
        const uint32_t* __restrict__ tv = vsrc + tile_base;
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
          const uint32_t li = lb + uint32_t(j) * 64;
          v[j] = tv[FULL || li < valid ? li : valid - 1];
        }
      }
    }
    if (RW == 0 && first) to_keys(k, a.mode_in);
  }
  uint32_t bk[LEARNED ? IPT : 1];
  if constexpr (LEARNED) {
#pragma unroll
    for (int j = 0; j < IPT; ++j) bk[j] = bucket_of(k[j], L.knots, mc, mc.bucket_shift);
  }
  auto dig = [&](int j) {
    if constexpr (LEARNED)
      return digit_of(bk[j], shift);
    else
      return digit_of(k[j], shift);
  };

  S* my = row + slot * 256 + tid;
  uint32_t rk[(IPT + 1) / 2];
  if (a.atomic_rank) {
    // Ranking takes a few instructions per key: the counts are published
    // right after it.
    if constexpr (LEARNED)
      rank_in_wave_atomic<uint32_t, IPT, FULL>(bk, rk, shift, L.whist[w], lb, valid);
    else
      rank_in_wave_atomic<U, IPT, FULL>(k, rk, shift, L.whist[w], lb, valid);
  } else {
    // Publish the tile's digit counts before ranking, so that later tiles
    // rarely find this one unpublished.
#pragma unroll
    for (int j = 0; j < IPT; ++j)
      if (in_tile(j)) atomicAdd(&L.tcnt[dig(j)], 1u);
    __syncthreads();
    if (dl)
      __hip_atomic_store(my, S(L.tcnt[tid]) | (first_round ? kPre : kAgg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // Emit nothing, but make the keys opaque so that the digits computed above
    // are recomputed (one v_bfe each) instead of being held in IPT registers.
#pragma unroll
    for (int j = 0; j < IPT; ++j) opaque(k[j]);
    if constexpr (LEARNED)
      rank_in_wave<uint32_t, IPT, FULL>(bk, rk, shift, L.whist[w], lb, valid);
    else
      rank_in_wave<U, IPT, FULL>(k, rk, shift, L.whist[w], lb, valid);
  }
#pragma unroll
  for (int j = 0; j < IPT; ++j) opaque(k[j]);  // recompute digits when staging
  __syncthreads();

  // Per digit: exclusive offsets of the waves, and the tile's count.
  uint32_t cnt = 0;
  uint32_t wexcl[NW];
#pragma unroll
  for (int ww = 0; ww < NW; ++ww) {
    wexcl[ww] = cnt;
    if (dl) cnt += L.whist[ww][tid];
  }
  if (a.atomic_rank && dl)
    __hip_atomic_store(my, S(cnt) | (first_round ? kPre : kAgg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  // L.wave is written again only by the next tile, after several barriers.
  const uint32_t excl_in_tile = block256_exclusive_add_once(cnt, L.wave);
  if (dl) {
#pragma unroll
    for (int ww = 0; ww < NW; ++ww) L.whist[ww][tid] = excl_in_tile + wexcl[ww];
  }
  __syncthreads();

  // Stage the tile in digit order.  It does not depend on the look-back, so it
  // comes first and the keys and ranks die before the wait.
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    if (in_tile(j)) {
      const uint32_t d = dig(j);
      const uint32_t pos = L.whist[w][d] + ((rk[j >> 1] >> (16 * (j & 1))) & 0xFFFFu);
      if constexpr (RW > 0)
        L.recs[pos] = r[j];
      else
        L.keys[pos] = k[j];
      if (VALS) L.vals[pos] = v[j];
    }
    // Keep the scheduler from hoisting every iteration's LDS read ahead of
    // the writes, which costs more registers than it hides latency.
    if ((j & 7) == 7) __builtin_amdgcn_sched_barrier(0);
  }

  // Decoupled look-back over the segment's earlier tiles for this lane's
  // digit, reading a window of predecessors per step and consuming them
  // newest first until one holds an inclusive prefix.  Every tile with a
  // smaller ticket is running and publishes without waiting on later tiles,
  // so the walk ends; the spin bound only turns a broken invariant into a
  // flagged wrong result instead of a hung GPU.
  unsigned long long before_tiles = 0;
  if (dl && !first_round) {
    constexpr int kWindow = TS_LOOKBACK_WINDOW;
    int64_t j = int64_t(slot) - 1;
    const int64_t stop = int64_t(first_slot);
    uint32_t spins = 0;
    while (true) {
      S st[kWindow];
#pragma unroll
      for (int q = 0; q < kWindow; ++q)
        st[q] = j - q >= stop ? __hip_atomic_load(row + size_t(j - q) * 256 + tid, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT)
                              : S(0);
      int used = 0;
      bool done = false;
#pragma unroll
      for (int q = 0; q < kWindow; ++q) {
        const S fl = st[q] >> VB;
        if (fl == 0) break;  // not published yet: wait for it
        before_tiles += (unsigned long long)(st[q] & kMask);
        ++used;
        if (fl == 2 || j - q == stop) {
          done = true;
          break;
        }
      }
      if (done) break;
      j -= used;
      if (used == 0) {
        if (++spins > kSpinLimit) {
          atomicOr(a.error, kErrLookback);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __hip_atomic_store(my, S(before_tiles + cnt) | kPre, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (dl) L.gofs[tid] = digit_base + before_tiles - excl_in_tile;
  __syncthreads();
  if constexpr (LOOP) {
    // The next tile's counters: this tile reads them no more (staging is
    // done), so they are cleared while its stores go out.
    for (unsigned i = tid; i < NW * 256; i += THREADS) (&L.whist[0][0])[i] = 0;
    if (tid < 256) L.tcnt[tid] = 0;
  }

  // Write each digit's run of the tile contiguously.  The output transform
  // and whether keys are written are uniform, so each case is its own loop.
  auto store = [&](auto transform, bool keys_out) {
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint32_t s = uint32_t(j) * THREADS + tid;
      if (FULL || s < valid) {
        const U kk = L.keys[s];
        // Learned digits are evaluated again rather than staged: the pass
        // is far from VALU-bound, and the LDS goes to a larger tile.
        const uint32_t d = LEARNED ? digit_of(bucket_of(kk, L.knots, mc, mc.bucket_shift), shift)
                                   : digit_of(kk, shift);
        const size_t o = size_t(L.gofs[d]) + s;
        if (keys_out) dst[o] = transform(kk);
        if (VALS) vdst[o] = L.vals[s];
      }
      if ((j & 7) == 7) __builtin_amdgcn_sched_barrier(0);
    }
  };
  // The next tile's claim, in flight with the stores: claimed after this
  // tile's look-back, so a tile starts right after its claim and successors
  // rarely wait for it.  The address goes through a lane register so that the
  // compiler keeps the atomic one lane's (its wave-combined form reads the
  // result back at once, waiting for the atomic before the stores).
  uint32_t claimed = 0;
  if (kEarlyClaim && tid == 0 && nt0 != 0) {
    uint32_t z = tid;
    asm volatile("" : "+v"(z));
    claimed = atomicAdd(&a.seg_tickets[pass * kSegments + sg0 + z], 1u);
  }
  if constexpr (RW > 0) {
    // Records leave as they came; the key is read out of each again.
    R* __restrict__ rdst = static_cast<R*>((pass & 1) ? a.recs_a : a.recs_b);
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint32_t s = uint32_t(j) * THREADS + tid;
      if (FULL || s < valid) {
        const R rr = L.recs[s];
        const U kk = to_key(rec_key<U, RW>(rr, a.key_word, a.key_shift), a.rec_mode);
        const size_t o = size_t(L.gofs[digit_of(kk, shift)]) + s;
        rdst[o] = rr;
        if (VALS) vdst[o] = L.vals[s];
      }
      if ((j & 7) == 7) __builtin_amdgcn_sched_barrier(0);
    }
    if (kEarlyClaim && tid == 0) L.ticket = (nt0 != 0 && claimed < nt0) ? ((sg0 << 24) | claimed) : kNoClaim;
    (void)TILE;
    return;
  }
  const bool write_keys = !(VALS && last && a.final_vals);
  const int out_mode = last ? a.mode_out : int(kIdentity);
  if (!write_keys)
    store([](U x) { return x; }, false);
  else if (out_mode == kSigned)
    store([](U x) { return from_key(x, kSigned); }, true);
  else if (out_mode == kFloat)
    store([](U x) { return from_key(x, kFloat); }, true);
  else
    store([](U x) { return x; }, true);
  if (kEarlyClaim && tid == 0) L.ticket = (nt0 != 0 && claimed < nt0) ? ((sg0 << 24) | claimed) : kNoClaim;
  (void)TILE;
}

// One stable scatter pass by the byte digit geo[pass].digit (scatter.rs:26-40).
//
// A workgroup claims a round of the segment numbered like its XCD through a
// per-(pass, segment) counter, moving on to the following segments once that
// one is exhausted: a segment's consecutive tiles, whose runs of each digit
// abut in the output, are then written through one L2.  Rounds of a segment
// are claimed in order, so a tile's predecessor was claimed earlier by a
// running workgroup.  Each wave ranks its IPT x 64 keys stably; the tile's
// digit counts are published; the exclusive prefix over the segment's
// earlier tiles is gathered by decoupled look-back, several predecessors per
// step; and the tile is staged in LDS in digit order so that each digit's
// keys leave as one coalesced run.  A key's position is (the first position
// of its digit in its segment, pbase) + (earlier tiles of its segment) +
// (its rank in the tile).
//
// Pass 0 of a by-key sort generates the values (positions 0..n-1); the last
// pass applies the inverse key transform, and in a by-key sort writes only the
// order to final_vals.  RW > 0: the by-key sort's 8- or 16-byte records are
// staged and written instead of keys, each pass reading the key out of them,
// so that the records end in order without a gather.  PERSIST: a grid
// smaller than the tile count whose workgroups claim tiles until none is left:
// the gated second sort, which then costs a few microseconds when it does not
// run, and the keys-only, record-carrying and learned passes of large
// sorts, where a workgroup claims its next tile while its stores go out
// instead of ending and a new one starting (MI355X, 2^28 u32 keys: 2.36 ->
// 2.27 ms per sort; 10^8 16-byte records by key 3.56 -> 3.53 ms; 10^9
// learned keys about 3 %; below about 8 tiles per CU the end of the pass
// costs more than that saves).
template <class U, bool VALS, class S, bool LEARNED = false, bool SMALL = false, bool PERSIST = false, int RW = 0>
__global__ __launch_bounds__((SweepShape<sizeof(U), LEARNED, VALS, SMALL, RW>::threads)) void k_sweep(
    SweepArgs<U> a, S* __restrict__ status, int pass) {
  constexpr int IPT = SweepShape<sizeof(U), LEARNED, VALS, SMALL, RW>::ipt;
  constexpr int THREADS = SweepShape<sizeof(U), LEARNED, VALS, SMALL, RW>::threads;
  constexpr int NW = THREADS / 64;
  constexpr int TILE = THREADS * IPT;
  __shared__ SweepLds<U, VALS, IPT, LEARNED, SMALL, RW> L;

  const SortMeta* __restrict__ meta = a.meta;
  if (a.gate && !meta->fallback) return;
  const uint32_t n_live = meta->n_live;
  if (uint32_t(pass) >= n_live) return;
  const unsigned tid = threadIdx.x;
  // The pass's geometry, from wave-uniform addresses (scalar loads).
  const PassGeo* __restrict__ g = a.geo + pass;
  const unsigned digit = g->digit;
  const uint32_t total = g->total;
  uint32_t nt[kSegments];
#pragma unroll
  for (int s = 0; s < kSegments; ++s) nt[s] = g->nt[s];
  const unsigned shift = 8 * digit;
  const bool first = pass == 0;
  const bool last = uint32_t(pass) + 1 == n_live;
  uint32_t xcc;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc));

  LearnedCoord mc{};
  if constexpr (LEARNED) {
    mc = coord_of(*a.model);
    load_knots(L.knots, a.model);
  }
  // Rows of the next pass's look-back words: a claimed tile clears its
  // slot's, and the rows past this pass's tiles are cleared here.
  S* next_row = status + size_t((pass + 1) & 1) * a.num_tiles * 256;
  if (tid < 256) {
    if (PERSIST) {
      for (uint32_t b = total + blockIdx.x; b < a.num_tiles; b += gridDim.x) next_row[size_t(b) * 256 + tid] = 0;
    } else if (blockIdx.x >= total && blockIdx.x < a.num_tiles) {
      next_row[size_t(blockIdx.x) * 256 + tid] = 0;
    }
  }
  for (unsigned i = tid; i < NW * 256; i += THREADS) (&L.whist[0][0])[i] = 0;
  if (tid < 256) L.tcnt[tid] = 0;
  bool have = false;  // (uniform) L.ticket holds the claim made during the last tile
  bool first_claim = true;
  while (true) {
    if (!have) {
      if (tid == 0) {
        uint32_t got = kSegments, r = 0;
        // Past a workgroup's first claim its own segment is exhausted: read
        // every counter at once rather than try each one in turn.
        uint32_t seen[kSegments];
#pragma unroll
        for (int q = 0; q < kSegments; ++q)
          seen[q] = first_claim ? 0u
                                : __hip_atomic_load(&a.seg_tickets[pass * kSegments + q], __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
        for (uint32_t i = 0; i < uint32_t(kSegments); ++i) {
          const uint32_t sg = (xcc + i) & uint32_t(kSegments - 1);
          if (nt[sg] == 0 || seen[sg] >= nt[sg]) continue;
          r = atomicAdd(&a.seg_tickets[pass * kSegments + sg], 1u);
          if (r < nt[sg]) {
            got = sg;
            break;
          }
        }
        L.ticket = (got << 24) | r;
      }
      first_claim = false;
      __syncthreads();
    }
    const uint32_t claim = __builtin_amdgcn_readfirstlane(L.ticket);
    const uint32_t seg = claim >> 24;
    if (seg >= uint32_t(kSegments)) return;
    const uint32_t round = claim & 0xFFFFFFu;
    // The next claim is tried on this tile's segment first: the workgroup's
    // own one, or the one it went on to once that was exhausted.
    const uint32_t seg_tiles = g->nt[seg];
    if (tid < 256) next_row[(size_t(g->slot[seg]) + round) * 256 + tid] = 0;

    const size_t seg_lo = g->start[seg];
    const size_t seg_hi = seg_lo + g->count[seg];
    const size_t tile_base = seg_lo + size_t(round) * TILE;
    const uint32_t valid = seg_hi - tile_base >= size_t(TILE) ? uint32_t(TILE) : uint32_t(seg_hi - tile_base);
    const size_t first_slot = g->slot[seg];
    const size_t slot = first_slot + round;
    // This lane's digit: the output position of the first key of that digit
    // in this segment.
    const unsigned long long digit_base = tid < 256 ? meta->pbase[pass][seg][tid] : 0ull;
    if (valid == uint32_t(TILE))
      run_tile<U, VALS, S, true, LEARNED, SMALL, RW, PERSIST>(L, a, mc, status, pass, shift, first, last, tile_base,
                                                                valid, first_slot, slot, round == 0, digit_base, seg,
                                                                seg_tiles);
    else
      run_tile<U, VALS, S, false, LEARNED, SMALL, RW, PERSIST>(L, a, mc, status, pass, shift, first, last, tile_base,
                                                                 valid, first_slot, slot, round == 0, digit_base, seg,
                                                                 seg_tiles);
    if (!PERSIST) return;
    // The tile's LDS is reused by the next tile, whose claim lane 0 has left
    // in L.ticket.
    __syncthreads();
    have = __builtin_amdgcn_readfirstlane(L.ticket) != kNoClaim;
  }
}

// Non-increasing keys found by the scan: reversing them sorts them.  Each
// lane swaps 8 keys at each end per step (the grid's lanes striped over the
// steps, so that every load and store is coalesced), loads before stores.
template <class U> __device__ __forceinline__ void reverse_keys(U* __restrict__ a, size_t n) {
  constexpr int R = 8;
  const size_t half = n / 2;
  const size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x, gs = size_t(gridDim.x) * blockDim.x;
  for (size_t i0 = g; i0 < half; i0 += gs * R) {
    U lo[R], hi[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const size_t i = min(i0 + size_t(r) * gs, half - 1);
      lo[r] = a[i];
      hi[r] = a[n - 1 - i];
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const size_t i = i0 + size_t(r) * gs;
      if (i < half) {
        a[i] = hi[r];
        a[n - 1 - i] = lo[r];
      }
    }
  }
}

// The last step of a keys-only sort: an odd number of live passes leaves the
// keys in the scratch half, and keys the scan found non-increasing are
// reversed.  After a divert the finish has already written the output.
template <class U>
__global__ void k_copy_back(const SortMeta* m, const U* __restrict__ src, U* __restrict__ dst, size_t n) {
  if (m->reverse) return reverse_keys(dst, n);
  if (m->skip || (m->n_live & 1u) == 0) return;
  const size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x, gs = size_t(gridDim.x) * blockDim.x;
  for (size_t i = g; i < n; i += gs) dst[i] = src[i];
}

// A record-carrying sort with an odd number of live passes ends in the
// scratch half.
template <class R>
__global__ void k_copy_back_records(const SortMeta* m, const R* __restrict__ src, R* __restrict__ dst, size_t n) {
  if ((m->n_live & 1u) == 0) return;
  const size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x, gs = size_t(gridDim.x) * blockDim.x;
  for (size_t i = g; i < n; i += gs) dst[i] = src[i];
}

// With no live digit every key is equal, or the keys were already in order:
// the stable order is the identity.
__global__ void k_iota_if_dead(const SortMeta* m, uint32_t* out, size_t n) {
  if (m->n_live != 0) return;
  const size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x, gs = size_t(gridDim.x) * blockDim.x;
  for (size_t i = g; i < n; i += gs) out[i] = uint32_t(i);
}

// ---------------------------------------------------------------- finish
// After a divert the keys are in order of their bits from finish_shift up
// (the prefix), and each run of keys sharing a prefix only needs sorting
// within itself (finish_runs, divert.rs:60-90).  Two kernels do it.
//
// k_finish_local: each workgroup takes tiles of kFinishTile positions (no
// overlap) and runs kLocalRounds rounds of odd-even transposition on the
// tile in LDS, comparing whole keys.  Neighbouring keys of different runs
// are already in order, so only keys of one run ever swap, and every run of
// at most kLocalRounds keys inside the tile comes out sorted.  The tile then
// records (one word per tile, in the look-back status area, free after the
// scatter passes) whether an out-of-order pair is left inside it (bit 0),
// and whether a run crosses its first position, so that the tile before it
// and it must be looked at again (bit 1).  It reads the position before its
// tile only for that run's prefix.
//
// k_finish: the tiles marked above (and the tile before a bit-1 tile) are
// finished again with windows of kFinishHalo keys on either side.  Every run
// of at most kFinishHalo keys is sorted by the tile that holds its first
// key: each key's rank is the number of keys of its run that are smaller, or
// equal and earlier (the order among equal keys is not observable).  A
// longer run is left as it is when it is already in order (a run of one
// repeated key, say); otherwise the finish sets m->fallback and the gated
// second sort sorts every digit.  Every tile classifies every run the same
// way (short runs lie inside the window of each tile they touch).  Run
// starts are kept as one bit per window position (wave ballots).
//
// The workgroups of one XCD (dispatch deals workgroups round-robin over the
// 8 XCDs) take consecutive tiles of one eighth of the array, so that
// neighbouring tiles are read through one L2.  The data is in the first
// buffer after an even number of executed passes: both kernels then work in
// place and write only keys that move; a k_finish tile never writes a run it
// does not own, and what it reads of other runs is only their prefixes,
// which moves inside a run do not change.  After an odd number the local
// finish moves every key from the second buffer into the first.
constexpr int kFinishThreads = 256;
constexpr int kFinishTile = kFinishThreads * 16;
constexpr int kFinishPer = kFinishTile / kFinishThreads;
constexpr int kLocalRounds = 4;
constexpr int kFinishHalo = 256;
constexpr int kFinishWin = kFinishTile + 2 * kFinishHalo;
constexpr int kFinishSteps = kFinishWin / kFinishThreads;  // window positions per lane
constexpr int kFinishWords = kFinishWin / 32 + 1;  // run-start bits, and the bit just past the window
constexpr int kFinishReach = kFinishHalo / 32 + 1;  // words a short run's bound can lie away
static_assert(kFinishWin % kFinishThreads == 0 && kFinishThreads % 64 == 0, "whole steps");
static_assert(kFinishHalo < kFinishTile && kFinishHalo % 32 == 0, "a short run touches at most two tiles");

// Bytes of per-tile marks the local finish writes into the status area.
inline size_t finish_mark_bytes(size_t n) { return ((n + kFinishTile - 1) / kFinishTile) * sizeof(uint32_t); }

// LDS slot of tile position q: one pad word per 16 keys, so that each lane's
// run of 16 consecutive keys starts 17 words after the previous lane's and
// the lanes' 8-byte reads spread over the banks.
__device__ __forceinline__ int finish_slot(int q) { return q + (q >> 4); }

// Four workgroups per CU (the grid is four per CU): at most 128 VGPRs.
template <class U>
__global__ __launch_bounds__(kFinishThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_finish_local(
    SortMeta* m, U* __restrict__ keys_a, const U* __restrict__ keys_b, size_t n, int mode,
    uint32_t* __restrict__ marks) {
  if (m->skip == 0) return;
  const uint32_t shift = m->finish_shift;
  const bool in_place = (m->n_live & 1u) == 0;
  const U* src = in_place ? keys_a : keys_b;
  U* dst = keys_a;
  constexpr int NW = kFinishThreads / 64;
  __shared__ U w[kFinishTile + kFinishTile / kFinishPer];
  // Each wave's first and last key, one copy per exchange (two odd rounds,
  // then the final order check, which needs only the first), so that no
  // exchange waits for the reads of the one before.
  __shared__ U edge_first[3][NW], edge_last[2][NW];
  const size_t tiles = (n + kFinishTile - 1) / kFinishTile;
  const uint32_t xcd = blockIdx.x & 7u, per_xcd = gridDim.x >> 3;  // gridDim.x is a multiple of 8
  const size_t stretch = (tiles + 7) / 8;
  const size_t t_end = min(tiles, size_t(xcd + 1) * stretch);
  // The key after this lane's last one (nf: the next lane's first), and the
  // key before its first (pl: the previous lane's last), across waves
  // through LDS; past the tile's ends, keys that never move a key (the
  // largest and the smallest key).
  unsigned tid = 0, lane = 0, wave = 0;
  auto exchange = [&](U first, U last, int x, U& nf, U& pl) {
    const U f = __shfl_down(first, 1, 64), l = __shfl_up(last, 1, 64);
    if (lane == 0) edge_first[x][wave] = first;
    if (lane == 63 && x < 2) edge_last[x][wave] = last;
    __syncthreads();
    nf = lane != 63 ? f : (wave + 1 < unsigned(NW) ? edge_first[x][wave + 1] : U(~U(0)));
    pl = lane != 0 ? l : (wave > 0 && x < 2 ? edge_last[x][wave - 1] : U(0));
  };
  for (size_t tile = size_t(xcd) * stretch + blockIdx.x / 8; tile < t_end; tile += per_xcd) {
    // The lane index is taken afresh for each tile, so that the lane's many
    // LDS and global addresses are not hoisted out of the loop and spilled.
    tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    lane = tid & 63;
    wave = tid >> 6;
    const size_t base = tile * kFinishTile;
    const int valid = int(min(n - base, size_t(kFinishTile)));
    const U* __restrict__ b = src + base;  // uniform: one 32-bit offset per load
    U reg[kFinishPer];
#pragma unroll
    for (int r = 0; r < kFinishPer; ++r) reg[r] = b[min(r * kFinishThreads + int(tid), valid - 1)];
    to_keys(reg, mode);
    __syncthreads();  // the previous tile's keys are no longer read
    // Positions past the tile's last key hold the largest key: they never
    // move a key of the tile and are never written.
#pragma unroll
    for (int r = 0; r < kFinishPer; ++r) {
      const int q = r * kFinishThreads + int(tid);
      w[finish_slot(q)] = q < valid ? reg[r] : U(~U(0));
    }
    __syncthreads();
    // Lane t holds positions 16t .. 16t + 15 and runs the odd-even rounds on
    // them in registers: even rounds pair (2i, 2i + 1) inside the lane; odd
    // rounds pair (2i + 1, 2i + 2) inside it and its last key with the next
    // lane's first.
    U c[kFinishPer];
#pragma unroll
    for (int i = 0; i < kFinishPer; ++i) c[i] = w[finish_slot(int(tid) * kFinishPer + i)];
    auto order = [&](int i) {
      const U lo = c[i] < c[i + 1] ? c[i] : c[i + 1], hi = c[i] < c[i + 1] ? c[i + 1] : c[i];
      c[i] = lo;
      c[i + 1] = hi;
    };
#pragma unroll
    for (int round = 0; round < kLocalRounds; ++round) {
      if ((round & 1) == 0) {
#pragma unroll
        for (int i = 0; i + 1 < kFinishPer; i += 2) order(i);
      } else {
        U nf, pl;
        exchange(c[0], c[kFinishPer - 1], round >> 1, nf, pl);
#pragma unroll
        for (int i = 1; i + 1 < kFinishPer; i += 2) order(i);
        c[kFinishPer - 1] = nf < c[kFinishPer - 1] ? nf : c[kFinishPer - 1];
        c[0] = pl > c[0] ? pl : c[0];
      }
    }
    static_assert(kLocalRounds == 4, "two odd rounds: exchanges 0 and 1");
    U nf, pl;
    exchange(c[0], c[kFinishPer - 1], 2, nf, pl);
    bool left = c[kFinishPer - 1] > nf;
#pragma unroll
    for (int i = 0; i + 1 < kFinishPer; ++i) left |= c[i] > c[i + 1];
    left = __syncthreads_or(left);
    if (tid == 0) {
      // A run crossing into this tile: compare prefixes only (the tile
      // before may be moving keys inside that run right now).
      const bool cross = base > 0 && (uint64_t(to_key(src[base - 1], mode)) >> shift) == (uint64_t(c[0]) >> shift);
      marks[tile] = (left ? 1u : 0u) | (cross ? 2u : 0u);
    }
    // Each lane writes its own positions: the keys that differ from the
    // tile's keys as read (still in LDS), or every key when the input was in
    // the second buffer.
    U* __restrict__ d = dst + base + size_t(tid) * kFinishPer;
    const KeyMasks<U> km(mode);
#pragma unroll
    for (int i = 0; i < kFinishPer; ++i) {
      const int q = int(tid) * kFinishPer + i;
      if (q < valid && (!in_place || c[i] != w[finish_slot(q)])) d[i] = km.from(c[i]);
    }
  }
}

template <class U>
__global__ __launch_bounds__(kFinishThreads) void k_finish(SortMeta* m, U* __restrict__ keys_a, size_t n, int mode,
                                                           const uint32_t* __restrict__ marks) {
  if (m->skip == 0) return;
  const uint32_t shift = m->finish_shift;
  // After the local finish the keys are in the first buffer.
  U* src = keys_a;
  U* dst = keys_a;
  __shared__ U w[kFinishWin];
  __shared__ uint32_t starts[kFinishWords];
  const unsigned tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  auto pre = [&](U k) { return uint64_t(k) >> shift; };
  bool unsorted = false;
  const size_t tiles = (n + kFinishTile - 1) / kFinishTile;
  const uint32_t xcd = blockIdx.x & 7u, per_xcd = gridDim.x >> 3;  // gridDim.x is a multiple of 8
  const size_t stretch = (tiles + 7) / 8;
  const size_t t_end = min(tiles, size_t(xcd + 1) * stretch);
  for (size_t tile = size_t(xcd) * stretch + blockIdx.x / 8; tile < t_end; tile += per_xcd) {
    // Only marked tiles, and a tile whose successor's first run crosses
    // back into it.
    if (!(marks[tile] & 3u) && !(tile + 1 < tiles && (marks[tile + 1] & 2u))) continue;  // uniform
    const size_t base = tile * kFinishTile;
    const size_t lo = base >= size_t(kFinishHalo) ? base - kFinishHalo : 0;
    const size_t hi = min(n, base + kFinishTile + kFinishHalo);
    const int wn = int(hi - lo);
    const int own0 = int(base - lo), own1 = int(min(n, base + kFinishTile) - lo);  // this tile: [own0, own1)
    U reg[kFinishSteps];
    {
      const U* __restrict__ b = src + lo;
#pragma unroll
      for (int r = 0; r < kFinishSteps; ++r) reg[r] = b[min(r * kFinishThreads + int(tid), wn - 1)];
    }
    to_keys(reg, mode);
    __syncthreads();  // the previous tile's window is no longer read
#pragma unroll
    for (int r = 0; r < kFinishSteps; ++r) {
      const int q = r * kFinishThreads + int(tid);
      if (q < wn) w[q] = reg[r];
    }
    if (tid == 0) starts[kFinishWords - 1] = (wn == kFinishWin && hi == n) ? 1u : 0u;
    __syncthreads();
    // Bit q: a run starts at window position q (its prefix differs from
    // q-1's; position 0 only at the array's start), or q = wn is the array's
    // end.  Before the window nothing is known.
#pragma unroll 3
    for (int r = 0; r < kFinishSteps; ++r) {
      const int q = r * kFinishThreads + int(tid);
      bool bt;
      if (q < wn)
        bt = q == 0 ? lo == 0 : pre(w[q]) != pre(w[q - 1]);
      else
        bt = q == wn && hi == n;
      const uint64_t bal = __ballot(bt);
      if (lane == 0) {
        const int wd = (r * kFinishThreads + int(wave) * 64) >> 5;
        starts[wd] = uint32_t(bal);
        starts[wd + 1] = uint32_t(bal >> 32);
      }
    }
    __syncthreads();
    // Each wave takes 64 consecutive positions per step; their run-start
    // bits are one 64-bit word every lane reads (a broadcast).
#pragma unroll 1
    for (int r = 0; r < kFinishSteps; ++r) {
      const int q0 = r * kFinishThreads + int(wave) * 64, q = q0 + int(lane);
      if (q0 + 64 <= own0 || q0 >= wn) continue;  // the whole wave is outside: uniform
      const int wd = q0 >> 5;
      const uint64_t B = uint64_t(starts[wd]) | uint64_t(starts[wd + 1]) << 32;
      const uint64_t upto = B & ((2ull << lane) - 1ull);  // bits 0 .. lane (all for lane 63)
      const uint64_t later = B & ~((2ull << lane) - 1ull);
      if (q < own0 || q >= wn) continue;
      const bool is_start = (B >> lane) & 1u;
      // The run's start: the last start bit at or before q.
      int rs = -1;
      if (upto) {
        rs = q0 + 63 - int(__clzll((long long)upto));
      } else {
        int wi = wd;
        uint32_t bitsw = 0;
        for (int k = 0; !bitsw && wi > 0 && k < kFinishReach; ++k) bitsw = starts[--wi];
        if (bitsw) rs = wi * 32 + 31 - __clz(bitsw);
      }
      // Its end: the first start bit after q.
      int re = -1;
      if (rs >= 0 && q - rs < kFinishHalo) {
        if (later) {
          re = q0 + __ffsll((unsigned long long)later) - 1;
        } else {
          int wi = wd + 1;
          uint32_t bitsw = 0;
          for (int k = 0; !bitsw && wi < kFinishWords - 1 && k < kFinishReach; ++k) bitsw = starts[++wi];
          if (bitsw) re = wi * 32 + __ffs(bitsw) - 1;
        }
      }
      const bool short_run = rs >= 0 && re >= 0 && re - rs <= kFinishHalo;
      if (short_run) {
        if (rs < own0 || rs >= own1 || re - rs == 1) continue;  // another tile's run, or one key
        const U k = w[q];
        int rank = 0;
        for (int j = rs; j < re; ++j) {
          const U x = w[j];
          rank += (x < k) || (x == k && j < q);
        }
        const int o = rs + rank;
        if (o != q) dst[lo + size_t(o)] = from_key(k, mode);
      } else if (q < own1 && !is_start && w[q - 1] > w[q]) {
        // A long run out of order (q is not its start, so q - 1 lies in it).
        unsorted = true;
      }
    }
  }
  if (__syncthreads_or(unsorted) && tid == 0) atomicOr(&m->fallback, 1u);
}

// Length of the first pass's position segments (a whole number of tiles).
size_t first_segment(size_t n, int key_bytes, bool learned, bool vals, int rec_words) {
  const size_t tile = sweep_tile_for(n, key_bytes, learned, vals, rec_words);
  const size_t per = (n + kSegments - 1) / kSegments;
  return ((per + tile - 1) / tile) * tile;
}

// Keys-only sorts of 8-byte keys that may divert launch their passes from
// this one on as persistent grids: after a divert they do not run, and a
// persistent grid ends in microseconds where a grid of one workgroup per
// tile takes about 11 us at 2^28 keys.
#ifndef TS_SWEEP_PERSIST_FROM
#define TS_SWEEP_PERSIST_FROM 4
#endif
// Keys-only, key-value, record-carrying and learned passes (one workgroup
// per CU) run as persistent grids from this many tiles per CU on (MI355X:
// u32 2^26 keys, 8 tiles per CU, 0.711 -> 0.686 ms, 2^27 1.266 -> 1.213 ms;
// 2^25, 4 tiles per CU, 0.415 -> 0.418 ms with a threshold of 4; u64 2^27
// 2.700 -> 2.648 ms).
#ifndef TS_SWEEP_PERSIST_TILES_PER_CU
#define TS_SWEEP_PERSIST_TILES_PER_CU 8
#endif

// gate: the second sort after a divert (persistent launches for keys-only
// 8-byte keys, run only when the finish found an out-of-order long run).
// One pass of a record-carrying sort (RW words per record).
template <class U, class S, bool VALS, int RW>
void launch_rec_sweep(uint32_t tiles, unsigned cus, hipStream_t s, const SweepArgs<U>& a, void* status, int p) {
  if (size_t(tiles) >= size_t(cus) * TS_SWEEP_PERSIST_TILES_PER_CU)
    hipLaunchKernelGGL((k_sweep<U, VALS, S, false, false, true, RW>), dim3(cus),
                       dim3(SweepShape<sizeof(U), false, VALS, false, RW>::threads), 0, s, a, static_cast<S*>(status), p);
  else
    hipLaunchKernelGGL((k_sweep<U, VALS, S, false, false, false, RW>), dim3(tiles),
                       dim3(SweepShape<sizeof(U), false, VALS, false, RW>::threads), 0, s, a, static_cast<S*>(status), p);
}

template <class U, class S, bool LEARNED>
hipError_t run_sweeps(const RadixJob& j, hipStream_t s, KernelClock* clk, bool gate, bool divert = false) {
  const bool vals = j.final_vals != nullptr;
  const int rw = LEARNED ? 0 : j.rec_words;
  const uint32_t tiles = sweep_tiles(j.n, sizeof(U), LEARNED, vals, rw);
  SweepArgs<U> a;
  a.keys_a = static_cast<U*>(j.keys_a);
  a.keys_b = static_cast<U*>(j.keys_b);
  a.vals_a = j.vals_a;
  a.vals_b = j.vals_b;
  a.final_vals = j.final_vals;
  a.n = j.n;
  a.mode_in = j.records ? int(kIdentity) : j.mode;
  a.mode_out = a.mode_in;
  a.meta = j.meta;
  a.geo = j.meta->geo;
  a.seg_tickets = &j.meta->seg_tickets[0][0];
  a.error = j.error ? j.error : &j.meta->error;
  a.num_tiles = tiles;
  a.model = j.model;
  a.atomic_rank = j.atomic_rank ? 1 : 0;
  a.gate = gate ? 1 : 0;
  a.recs_a = const_cast<unsigned char*>(j.records);
  a.recs_b = j.recs_b;
  a.key_word = uint32_t(j.key_offset / 4);
  a.key_shift = uint32_t(8 * (j.key_offset % 4));
  a.rec_mode = j.mode;
  const double bytes = double(j.n) * (2.0 * (rw ? 4.0 * rw : double(sizeof(U))) + (vals ? 8.0 : 0.0));
  const unsigned cus = unsigned(std::max(1, j.num_cus));
  for (int p = 0; p < DigitSource<U, LEARNED>::P; ++p) {
    if (clk) clk->begin(s, gate ? "k_sweep_fallback" : rw ? "k_sweep_records" : "k_sweep", bytes);
    if (rw) {
      if (rw == 2 && vals)
        launch_rec_sweep<U, S, true, 2>(tiles, cus, s, a, j.status, p);
      else if (rw == 2)
        launch_rec_sweep<U, S, false, 2>(tiles, cus, s, a, j.status, p);
      else if (vals)
        launch_rec_sweep<U, S, true, 4>(tiles, cus, s, a, j.status, p);
      else
        launch_rec_sweep<U, S, false, 4>(tiles, cus, s, a, j.status, p);
    } else if constexpr (LEARNED) {
      if (size_t(tiles) >= size_t(cus) * TS_SWEEP_PERSIST_TILES_PER_CU)
        hipLaunchKernelGGL((k_sweep<U, false, S, true, false, true>), dim3(cus),
                           dim3(SweepShape<sizeof(U), true>::threads), 0, s, a, static_cast<S*>(j.status), p);
      else
        hipLaunchKernelGGL((k_sweep<U, false, S, true>), dim3(tiles), dim3(SweepShape<sizeof(U), true>::threads), 0, s,
                           a, static_cast<S*>(j.status), p);
    } else if (vals) {
      if (sizeof(U) >= 4 && size_t(tiles) >= size_t(cus) * TS_SWEEP_PERSIST_TILES_PER_CU) {
        if constexpr (sizeof(U) >= 4)
          hipLaunchKernelGGL((k_sweep<U, true, S, false, false, true>), dim3(cus),
                             dim3(SweepShape<sizeof(U), false, true>::threads), 0, s, a, static_cast<S*>(j.status), p);
      } else {
        hipLaunchKernelGGL((k_sweep<U, true, S>), dim3(tiles), dim3(SweepShape<sizeof(U), false, true>::threads), 0, s,
                           a, static_cast<S*>(j.status), p);
      }
    } else if (gate || (divert && p >= TS_SWEEP_PERSIST_FROM)) {
      if constexpr (sizeof(U) == 8) {
        const unsigned grid = unsigned(std::min<size_t>(tiles, size_t(j.num_cus)));
        hipLaunchKernelGGL((k_sweep<U, false, S, false, false, true>), dim3(grid), dim3(SweepShape<sizeof(U)>::threads),
                           0, s, a, static_cast<S*>(j.status), p);
      } else {
        return hipErrorInvalidValue;
      }
    } else if (small_sweep(j.n, false, false)) {
      hipLaunchKernelGGL((k_sweep<U, false, S, false, true>), dim3(tiles),
                         dim3(SweepShape<sizeof(U), false, false, true>::threads), 0, s, a, static_cast<S*>(j.status), p);
    } else if (sizeof(U) >= 4 && size_t(tiles) >= size_t(cus) * TS_SWEEP_PERSIST_TILES_PER_CU) {
      if constexpr (sizeof(U) >= 4)
        hipLaunchKernelGGL((k_sweep<U, false, S, false, false, true>), dim3(cus), dim3(SweepShape<sizeof(U)>::threads),
                           0, s, a, static_cast<S*>(j.status), p);
    } else {
      hipLaunchKernelGGL((k_sweep<U, false, S>), dim3(tiles), dim3(SweepShape<sizeof(U)>::threads), 0, s, a,
                         static_cast<S*>(j.status), p);
    }
    if (clk) clk->end(s);
  }
  return hipGetLastError();
}

template <class U, bool LEARNED = false> hipError_t run_radix(const RadixJob& j, hipStream_t s, KernelClock* clk) {
  constexpr int P = DigitSource<U, LEARNED>::P;
  constexpr int KB = int(sizeof(U));
  const bool vals = j.final_vals != nullptr;
  const int rw = j.rec_words;
  if (rw && (LEARNED || !j.records || !j.recs_b || j.stride != size_t(4 * rw) ||
             reinterpret_cast<uintptr_t>(j.records) % j.stride != 0 ||
             reinterpret_cast<uintptr_t>(j.recs_b) % j.stride != 0 || j.key_offset + KB > j.stride ||
             j.key_offset % KB != 0))
    return hipErrorInvalidValue;
  // The look-back rows this sort's launches index must lie inside the
  // caller's status area.
  if (!j.status || status_bytes_used(j.n, KB, LEARNED, vals, rw) > j.status_cap) return hipErrorInvalidValue;
  const bool divert = !LEARNED && !vals && !j.records && KB == 8 && j.divert && !small_sweep(j.n, false, false);
  hipLaunchKernelGGL(k_init, dim3(16), dim3(256), 0, s, j.meta, P, 0);
  // Histogram: blocks_per_seg workgroups per position segment, each over a
  // contiguous chunk of it.
  const size_t seg0 = first_segment(j.n, KB, LEARNED, vals, rw);
  // Histogram workgroups per CU.  The learned histogram holds three digit
  // tables (and the knots), not eight, so more fit a CU;
  // TS_HIST_BLOCKS_PER_CU_LEARNED counts 256-lane workgroups.
  constexpr int kHistThreads = HistUnroll<U>::threads;
  const int hist_per_cu = LEARNED ? std::max(1, TS_HIST_BLOCKS_PER_CU_LEARNED * 256 / kHistThreads)
                                  : HistUnroll<U>::blocks_per_cu;
  const size_t step = size_t(kHistThreads) * HistUnroll<U>::value;
  const size_t min_keys = std::max(step, size_t(TS_HIST_MIN_KEYS));
  const size_t by_keys = std::max<size_t>(TS_HIST_MIN_BLOCKS_PER_SEG, (seg0 + min_keys - 1) / min_keys);
  const uint32_t blocks_per_seg =
      uint32_t(std::max<size_t>(1, std::min<size_t>(size_t(j.num_cus) * hist_per_cu / kSegments, by_keys)));
  size_t per_block = (seg0 + blocks_per_seg - 1) / blocks_per_seg;
  per_block = std::max(step, ((per_block + step - 1) / step) * step);
  const uint32_t tiles = sweep_tiles(j.n, KB, LEARNED, vals, rw);
  const size_t zero_words = size_t(tiles) * 256 * (wide_status(j.n) ? 2 : 1);
  const dim3 hgrid(blocks_per_seg * kSegments);
  const uint32_t tile = uint32_t(sweep_tile_for(j.n, KB, LEARNED, vals, rw));
  // The order check does not apply to bucket digits, and only keys without
  // values may be reversed (a stable order of non-increasing keys with ties
  // is no reversal).
  const bool vals_or_rec = j.records != nullptr || vals;
  const int check_order = LEARNED ? 0 : vals_or_rec ? 1 : 2;
  if constexpr (!LEARNED) {
    const unsigned og = unsigned(std::max(1, j.num_cus * TS_ORDER_BLOCKS_PER_CU));
    if (j.records)
      hipLaunchKernelGGL((k_order_check<U, true>), dim3(og), dim3(256), 0, s, nullptr, j.records, j.stride,
                         j.key_offset, j.n, j.mode, j.meta);
    else
      hipLaunchKernelGGL((k_order_check<U, false>), dim3(og), dim3(256), 0, s, static_cast<const U*>(j.keys_a),
                         nullptr, size_t(0), size_t(0), j.n, j.mode, j.meta);
  }
  auto hist = [&](int gate) {
    const int co = gate ? 0 : check_order;
    if (clk) clk->begin(s, gate ? "k_hist_fallback" : "k_hist", double(j.n) * (j.records ? double(j.stride) + KB : KB));
    if constexpr (LEARNED)
      hipLaunchKernelGGL((k_hist<U, false, true>), hgrid, dim3(kHistThreads), 0, s, static_cast<const U*>(j.keys_a),
                         nullptr, size_t(0), size_t(0), nullptr, j.n, j.mode, seg0, blocks_per_seg, per_block, j.meta,
                         static_cast<uint32_t*>(j.status), zero_words, j.model, co, gate, j.fingerprint);
    else if (j.records)
      hipLaunchKernelGGL((k_hist<U, true>), hgrid, dim3(kHistThreads), 0, s, nullptr, j.records, j.stride,
                         j.key_offset, rw ? nullptr : static_cast<U*>(j.keys_a), j.n, j.mode, seg0, blocks_per_seg,
                         per_block, j.meta,
                         static_cast<uint32_t*>(j.status), zero_words, nullptr, co, gate, nullptr);
    else
      hipLaunchKernelGGL((k_hist<U, false>), hgrid, dim3(kHistThreads), 0, s, static_cast<const U*>(j.keys_a),
                         nullptr, size_t(0), size_t(0), nullptr, j.n, j.mode, seg0, blocks_per_seg, per_block, j.meta,
                         static_cast<uint32_t*>(j.status), zero_words, nullptr, co, gate, nullptr);
    if (clk) clk->end(s);
  };
  hist(0);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(256), 0, s, j.meta, j.n, P, seg0, tile, check_order, divert ? 1 : 0, 0);
  hipError_t e = wide_status(j.n) ? run_sweeps<U, uint64_t, LEARNED>(j, s, clk, false, divert)
                                  : run_sweeps<U, uint32_t, LEARNED>(j, s, clk, false, divert);
  if (e != hipSuccess) return e;
  if (divert) {
    if constexpr (!LEARNED && KB == 8) {
      // Finish the runs, then the second sort that runs only if a long run
      // was out of order.  The per-tile marks live in the status area.
      if (status_bytes_used(j.n, KB, false, false) < finish_mark_bytes(j.n)) return hipErrorInvalidValue;
      uint32_t* marks = static_cast<uint32_t*>(j.status);
      const unsigned fgrid = std::max(8u, (unsigned(j.num_cus) * 4u) & ~7u);  // a multiple of 8 (XCDs)
      if (clk) clk->begin(s, "k_finish_local", double(j.n) * KB);
      hipLaunchKernelGGL((k_finish_local<U>), dim3(fgrid), dim3(kFinishThreads), 0, s, j.meta,
                         static_cast<U*>(j.keys_a), static_cast<const U*>(j.keys_b), j.n, j.mode, marks);
      if (clk) clk->end(s);
      if (clk) clk->begin(s, "k_finish", 0.0);
      hipLaunchKernelGGL((k_finish<U>), dim3(fgrid), dim3(kFinishThreads), 0, s, j.meta, static_cast<U*>(j.keys_a),
                         j.n, j.mode, static_cast<const uint32_t*>(marks));
      if (clk) clk->end(s);
      hipLaunchKernelGGL(k_init, dim3(16), dim3(256), 0, s, j.meta, P, 1);
      hist(1);
      hipLaunchKernelGGL(k_scan, dim3(1), dim3(256), 0, s, j.meta, j.n, P, seg0, tile, 0, 0, 1);
      e = wide_status(j.n) ? run_sweeps<U, uint64_t, LEARNED>(j, s, clk, true)
                           : run_sweeps<U, uint32_t, LEARNED>(j, s, clk, true);
      if (e != hipSuccess) return e;
    }
  }
  const unsigned cb = unsigned(std::min<size_t>((j.n + 255) / 256, size_t(j.num_cus) * 8));
  if (rw == 2)
    hipLaunchKernelGGL(k_copy_back_records<uint2>, dim3(cb ? cb : 1), dim3(256), 0, s, j.meta,
                       static_cast<const uint2*>(j.recs_b), reinterpret_cast<uint2*>(const_cast<unsigned char*>(j.records)),
                       j.n);
  else if (rw == 4)
    hipLaunchKernelGGL(k_copy_back_records<uint4>, dim3(cb ? cb : 1), dim3(256), 0, s, j.meta,
                       static_cast<const uint4*>(j.recs_b), reinterpret_cast<uint4*>(const_cast<unsigned char*>(j.records)),
                       j.n);
  if (j.final_vals) {
    hipLaunchKernelGGL(k_iota_if_dead, dim3(cb ? cb : 1), dim3(256), 0, s, j.meta, j.final_vals, j.n);
  } else if (!rw) {
    hipLaunchKernelGGL((k_copy_back<U>), dim3(cb ? cb : 1), dim3(256), 0, s, j.meta, static_cast<const U*>(j.keys_b),
                       static_cast<U*>(j.keys_a), j.n);
  }
  return hipGetLastError();
}

// Ranks eight digit patterns (uniform bytes, two values, sixteen values, one
// value, one digit per lane, pairs of lanes, 90 % one digit, runs of eight)
// both with LDS atomics and with ballots and adds the number of keys whose
// ranks differ to *bad.  Block b uses pattern b % 8.
__global__ __launch_bounds__(512) void k_rank_check(uint32_t* bad) {
  constexpr int IPT = 16;
  __shared__ uint32_t ha[8][256], hb[8][256];
  const unsigned tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  for (unsigned i = tid; i < 8 * 256; i += 512) {
    (&ha[0][0])[i] = 0;
    (&hb[0][0])[i] = 0;
  }
  __syncthreads();
  const unsigned pattern = blockIdx.x & 7u;
  uint32_t k[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    uint32_t x = (blockIdx.x * 512u + tid) * 0x9E3779B1u + uint32_t(j) * 0x85EBCA77u;
    x ^= x >> 15;
    x *= 0x2C1B3C6Du;
    x ^= x >> 12;
    uint32_t d;
    switch (pattern) {
      case 0: d = x; break;
      case 1: d = (x & 1u) ? 7u : 200u; break;
      case 2: d = (x & 15u) << 4; break;
      case 3: d = 42u; break;
      case 4: d = lane; break;
      case 5: d = lane >> 1; break;
      case 6: d = (x % 10u) ? 3u : (x >> 8); break;
      default: d = ((uint32_t(j) * 64u + lane) >> 3) * 37u; break;
    }
    k[j] = d & 255u;
  }
  uint32_t ra[IPT / 2], rb[IPT / 2];
  rank_in_wave_atomic<uint32_t, IPT, true>(k, ra, 0u, ha[w], 0u, 0u);
  rank_in_wave<uint32_t, IPT, true>(k, rb, 0u, hb[w], 0u, 0u);
  uint32_t diff = 0;
#pragma unroll
  for (int q = 0; q < IPT / 2; ++q) diff += (ra[q] & 0xFFFFu) != (rb[q] & 0xFFFFu);
#pragma unroll
  for (int q = 0; q < IPT / 2; ++q) diff += (ra[q] >> 16) != (rb[q] >> 16);
  if (diff) atomicAdd(bad, diff);
}

}  // namespace

// ---------------------------------------------------------- tile geometry
namespace {
size_t tile_of(int key_bytes, bool learned, bool vals, bool small, int rec_words) {
  if (rec_words == 2)
    return size_t(vals ? SweepShape<4, false, true, false, 2>::tile : SweepShape<4, false, false, false, 2>::tile);
  if (rec_words == 4)
    return size_t(vals ? SweepShape<4, false, true, false, 4>::tile : SweepShape<4, false, false, false, 4>::tile);
  if (small) return size_t(key_bytes < 8 ? SweepShape<4, false, false, true>::tile : SweepShape<8, false, false, true>::tile);
  if (key_bytes < 8)
    return size_t(learned ? SweepShape<4, true>::tile : vals ? SweepShape<4, false, true>::tile : SweepShape<4>::tile);
  return size_t(learned ? SweepShape<8, true>::tile : vals ? SweepShape<8, false, true>::tile : SweepShape<8>::tile);
}
size_t status_row_bytes(size_t n) { return size_t(256) * (wide_status(n) ? 8 : 4); }
}  // namespace

bool small_sweep(size_t n, bool learned, bool vals) { return !learned && !vals && n <= TS_SWEEP_SMALL_N; }

size_t sweep_tile_for(size_t n, int key_bytes, bool learned, bool vals, int rec_words) {
  return tile_of(key_bytes, learned, vals, rec_words == 0 && small_sweep(n, learned, vals), rec_words);
}

uint32_t sweep_tiles(size_t n, int key_bytes, bool learned, bool vals, int rec_words) {
  const size_t tile = sweep_tile_for(n, key_bytes, learned, vals, rec_words);
  return uint32_t((n + tile - 1) / tile + kSegments);
}

bool wide_status(size_t n) { return n >= (size_t(1) << 30); }

size_t status_bytes_used(size_t n, int key_bytes, bool learned, bool vals, int rec_words) {
  return size_t(2) * sweep_tiles(n, key_bytes, learned, vals, rec_words) * status_row_bytes(n);
}

size_t status_bytes(size_t n, int key_bytes) {
  // Per kind the row count grows with n, except that keys-only sorts of at
  // most TS_SWEEP_SMALL_N keys switch to smaller tiles: the longest such sort
  // bounds them.  Rows are as wide as n's.
  size_t rows = std::max({sweep_tiles(n, key_bytes, false, false), sweep_tiles(n, key_bytes, false, true),
                          sweep_tiles(n, key_bytes, true, false),
                          sweep_tiles(std::min<size_t>(n, TS_SWEEP_SMALL_N), key_bytes, false, false)});
  return size_t(2) * rows * status_row_bytes(n);
}

hipError_t rank_order_check(uint32_t* bad, hipStream_t s) {
  hipLaunchKernelGGL(k_rank_check, dim3(64), dim3(512), 0, s, bad);
  return hipGetLastError();
}

hipError_t radix_sort(const RadixJob& job, hipStream_t s, KernelClock* clk) {
  if (job.model) {
    if (job.records || job.final_vals) return hipErrorInvalidValue;
    switch (job.key_bytes) {
      case 4: return run_radix<uint32_t, true>(job, s, clk);
      case 8: return run_radix<uint64_t, true>(job, s, clk);
      default: return hipErrorInvalidValue;
    }
  }
  switch (job.key_bytes) {
    case 1: return run_radix<uint8_t>(job, s, clk);
    case 2: return run_radix<uint16_t>(job, s, clk);
    case 4: return run_radix<uint32_t>(job, s, clk);
    case 8: return run_radix<uint64_t>(job, s, clk);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace ts
