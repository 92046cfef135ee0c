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
__global__ void k_init(SortMeta* m, int P) {
  for (unsigned i = threadIdx.x; i < unsigned(P) * kSegments * 256; i += blockDim.x) (&m->pair[0][0][0])[i] = 0;
  if (threadIdx.x < 8) m->tickets[threadIdx.x] = 0;
  if (threadIdx.x < 8 * kSegments) (&m->seg_tickets[0][0])[threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    m->n_live = 0;
    m->error = 0;
    m->breaks_up = 0;
    m->breaks_down = 0;
    m->reverse = 0;
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
// passes start from.
//
// A digit (with its segment bits) on which all of a wave's keys of one step
// agree -- constant digits, sorted runs -- is counted with one LDS add
// instead of 64 colliding atomics.  The grid also clears the first look-back
// row.
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
__device__ __forceinline__ uint32_t bucket_of(U k, const unsigned long long* knots, const LearnedModel& m,
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
                                                       int check_order) {
  using D = typename DigitSource<U, LEARNED>::T;
  constexpr int P = DigitSource<U, LEARNED>::P;
  constexpr int UN = HistUnroll<U>::value;
  constexpr int kHistThreads = HistUnroll<U>::threads;
  // Keys already in order (k_order_check; check_order as k_scan's presorted):
  // no pass runs, so no counts.
  if (!LEARNED && check_order && (!m->breaks_up || (check_order == 2 && !m->breaks_down))) return;
  __shared__ uint32_t h[P * kSegments * 256];
  __shared__ unsigned long long knots[LEARNED ? kKnots : 1];
  const unsigned tid = threadIdx.x, lane = tid & 63;
  if (LEARNED) load_knots(knots, model);
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
  for (size_t base = begin; base < end; base += size_t(kHistThreads) * UN) {
    const bool full = base + size_t(kHistThreads) * UN <= end;
    U k[UN];
#pragma unroll
    for (int r = 0; r < UN; ++r) {
      size_t i = base + size_t(r) * kHistThreads + tid;
      if (!full) i = i < end ? i : end - 1;
      k[r] = REC ? *reinterpret_cast<const U*>(rec + i * stride + key_off) : keys[i];
    }
    to_keys(k, mode);
    if (REC) {
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
        dk[r] = bucket_of(k[r], knots, *model, model->bucket_shift);
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
  __syncthreads();
  for (unsigned i = tid; i < P * kSegments * 256; i += kHistThreads) {
    const uint32_t c = h[i];
    if (c) atomicAdd(&m->pair[0][0][0] + i, (unsigned long long)c);
  }
}

// ------------------------------------------------------------------ scan
// Per digit p: its histogram (the segments' counts summed), the digit bases
// (prefix_sum.rs:17-24), each segment's offset inside each digit's run, the
// segment bounds and tile counts of p's pass, and whether p is live.  A digit
// is dead when one bin holds every key, as the reference decides from the
// first key's bin (radix/mod.rs:229-239).  When digit p-1 is dead its value
// segments collapse into one segment holding all keys in input order, which
// is the plain single-chain pass, so the segments stay correct whichever
// digits are live.
// presorted: 0 ignores the order check; 1 skips every pass when the keys are
// already non-decreasing (finish_presorted, presorted.rs:15-27); 2 also
// reverses non-increasing keys in place instead of sorting them (keys only:
// equal keys are equal bits, presorted.rs:10-14).
__global__ __launch_bounds__(256) void k_scan(SortMeta* m, size_t n, int P, size_t seg0, uint32_t tile,
                                              int presorted) {
  __shared__ unsigned long long s_wave[4];
  __shared__ unsigned long long s_prev[256], s_prev_base[256];
  __shared__ uint32_t s_tiles[kSegments];
  const unsigned tid = threadIdx.x;
  uint32_t n_live = 0;
  for (int p = 0; p < P; ++p) {
    unsigned long long tot = 0;
#pragma unroll
    for (int s = 0; s < kSegments; ++s) {
      const unsigned long long c = m->pair[p][s][tid];
      m->segpre[p][s][tid] = tot;
      tot += c;
    }
    const unsigned long long e = block256_exclusive_add(tot, s_wave);
    m->base[p][tid] = e;
    const int full = __syncthreads_or(tot == (unsigned long long)n);
    if (!full) {
      if (tid == 0) m->live[n_live] = uint32_t(p);
      ++n_live;
    }
    if (tid < kSegments) {
      unsigned long long st, ct;
      if (p == 0) {
        st = min((unsigned long long)n, (unsigned long long)tid * seg0);
        ct = min((unsigned long long)n, st + seg0) - st;
      } else {
        ct = 0;
        for (int x = 0; x < 32; ++x) ct += s_prev[tid * 32 + x];
        st = s_prev_base[tid * 32];
      }
      m->seg_start[p][tid] = st;
      m->seg_count[p][tid] = ct;
      s_tiles[tid] = uint32_t((ct + tile - 1) / tile);
    }
    __syncthreads();
    if (tid == 0) {
      uint32_t slot = 0;
      for (int s = 0; s < kSegments; ++s) {
        m->seg_slot[p][s] = slot;
        m->seg_tiles[p][s] = s_tiles[s];
        slot += s_tiles[s];
      }
      m->total_tiles[p] = slot;
    }
    s_prev[tid] = tot;
    s_prev_base[tid] = e;
  }
  __syncthreads();
  if (presorted >= 1 && m->breaks_up == 0) {
    n_live = 0;
  } else if (presorted == 2 && m->breaks_down == 0) {
    n_live = 0;
    if (tid == 0) m->reverse = 1;
  }
  if (tid == 0) m->n_live = n_live;
  // Geometry of each executed pass, by execution order.
  for (uint32_t i = 0; i < n_live; ++i) {
    const uint32_t p = m->live[i];
    PassGeo* g = &m->geo[i];
    if (tid == 0) {
      g->digit = p;
      g->total = m->total_tiles[p];
    }
    if (tid < kSegments) {
      g->nt[tid] = m->seg_tiles[p][tid];
      g->slot[tid] = m->seg_slot[p][tid];
      g->start[tid] = m->seg_start[p][tid];
      g->count[tid] = m->seg_count[p][tid];
    }
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
  uint32_t* tickets;
  uint32_t* seg_tickets;  // [pass][segment]
  uint32_t* error;  // set when a look-back wait exceeds its bound
  uint32_t num_tiles;
  const LearnedModel* model;  // learned partition: digits of the key's bucket
  int atomic_rank;            // rank with LDS atomics (rank_in_wave_atomic), else ballots
};

#ifndef TS_SWEEP_XCD_AFFINE
#define TS_SWEEP_XCD_AFFINE 1
#endif

// About two seconds of s_sleep(1) polling.
constexpr uint32_t kSpinLimit = 1u << 26;

// One stable scatter pass by the byte digit meta->live[pass] (scatter.rs:26-40).
//
// Tiles are claimed through a ticket in an order that interleaves the pass's
// segments (round t of every segment before round t+1), so every tile a tile
// waits on has already started, and each segment's look-back chain advances
// at 1/kSegments of the tile rate.  Each wave ranks its IPT x 64 keys stably
// (rank_in_wave); the tile's digit counts are published first from an LDS
// histogram; the exclusive prefix over the segment's earlier tiles is
// gathered by decoupled look-back, several predecessors per step; and the
// tile is staged in LDS in digit order so that each digit's keys leave as one
// coalesced run.  A key's position is digit base + the count of its digit in
// earlier segments + earlier tiles of its segment + its rank in the tile.
//
// Pass 0 of a by-key sort generates the values (positions 0..n-1); the last
// pass applies the inverse key transform, and in a by-key sort writes only the
// order to final_vals.
// LDS of one scatter tile, declared once in the kernel.
template <class U, bool VALS, int IPT, bool LEARNED = false, bool SMALL = false> struct SweepLds {
  static constexpr int NW = SweepShape<sizeof(U), LEARNED, VALS, SMALL>::threads / 64;
  static constexpr int TILE = SweepShape<sizeof(U), LEARNED, VALS, SMALL>::threads * IPT;
  uint32_t whist[NW][256];  // per-wave digit counts, then staging bases
  U keys[TILE];
  uint32_t vals[VALS ? TILE : 1];
  unsigned long long knots[LEARNED ? kKnots : 1];
  unsigned long long gofs[256];
  uint32_t wave[NW];
  uint32_t tcnt[256];
  uint32_t ticket;
};

// Everything a tile does after it knows where it is: load, publish its digit
// counts, rank, look back, stage, store.  FULL tiles (all but each segment's
// last) run without any per-key bounds test.
template <class U, bool VALS, class S, bool FULL, bool LEARNED, bool SMALL>
__device__ __forceinline__ void run_tile(SweepLds<U, VALS, SweepShape<sizeof(U), LEARNED, VALS, SMALL>::ipt, LEARNED, SMALL>& L,
                                           const SweepArgs<U>& a,
                                           S* __restrict__ status, int pass, unsigned shift, bool first, bool last,
                                           size_t tile_base, uint32_t valid, size_t first_slot, size_t slot,
                                           bool first_round, unsigned long long digit_base) {
  constexpr int IPT = SweepShape<sizeof(U), LEARNED, VALS, SMALL>::ipt;
  constexpr int THREADS = SweepShape<sizeof(U), LEARNED, VALS, SMALL>::threads;
  constexpr int NW = THREADS / 64;
  constexpr int TILE = THREADS * IPT;
  constexpr int VB = Flags<S>::kBits;
  constexpr S kAgg = S(1) << VB, kPre = S(2) << VB, kMask = kAgg - 1;
  const unsigned tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
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
  {
    const U* __restrict__ ts = src + tile_base;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint32_t li = lb + uint32_t(j) * 64;
      k[j] = ts[FULL || li < valid ? li : valid - 1];
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
    if (first) to_keys(k, a.mode_in);
  }
  uint32_t bk[LEARNED ? IPT : 1];
  if constexpr (LEARNED) {
#pragma unroll
    for (int j = 0; j < IPT; ++j) bk[j] = bucket_of(k[j], L.knots, *a.model, a.model->bucket_shift);
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
  const uint32_t excl_in_tile = block256_exclusive_add(cnt, L.wave);
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
          atomicOr(a.error, 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __hip_atomic_store(my, S(before_tiles + cnt) | kPre, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (dl) L.gofs[tid] = digit_base + before_tiles - excl_in_tile;
  __syncthreads();

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
        const uint32_t d = LEARNED ? digit_of(bucket_of(kk, L.knots, *a.model, a.model->bucket_shift), shift)
                                   : digit_of(kk, shift);
        const size_t o = size_t(L.gofs[d]) + s;
        if (keys_out) dst[o] = transform(kk);
        if (VALS) vdst[o] = L.vals[s];
      }
      if ((j & 7) == 7) __builtin_amdgcn_sched_barrier(0);
    }
  };
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
  (void)TILE;
}

// One stable scatter pass by the byte digit geo[pass].digit (scatter.rs:26-40).
//
// Tiles are claimed through a ticket in an order that interleaves the pass's
// segments (round t of every segment before round t+1), so every tile a tile
// waits on has already started, and each segment's look-back chain advances
// at 1/kSegments of the tile rate.  Each wave ranks its IPT x 64 keys stably
// (rank_in_wave); the tile's digit counts are published first from an LDS
// histogram; the exclusive prefix over the segment's earlier tiles is
// gathered by decoupled look-back, several predecessors per step; and the
// tile is staged in LDS in digit order so that each digit's keys leave as one
// coalesced run.  A key's position is digit base + the count of its digit in
// earlier segments + earlier tiles of its segment + its rank in the tile.
//
// Pass 0 of a by-key sort generates the values (positions 0..n-1); the last
// pass applies the inverse key transform, and in a by-key sort writes only the
// order to final_vals.
template <class U, bool VALS, class S, bool LEARNED = false, bool SMALL = false>
__global__ __launch_bounds__((SweepShape<sizeof(U), LEARNED, VALS, SMALL>::threads)) void k_sweep(
    SweepArgs<U> a, S* __restrict__ status, int pass) {
  constexpr int IPT = SweepShape<sizeof(U), LEARNED, VALS, SMALL>::ipt;
  constexpr int THREADS = SweepShape<sizeof(U), LEARNED, VALS, SMALL>::threads;
  constexpr int NW = THREADS / 64;
  constexpr int TILE = THREADS * IPT;
  __shared__ SweepLds<U, VALS, IPT, LEARNED, SMALL> L;

  const SortMeta* __restrict__ meta = a.meta;
  const uint32_t n_live = meta->n_live;
  if (uint32_t(pass) >= n_live) return;
  const unsigned tid = threadIdx.x;
  if (!TS_SWEEP_XCD_AFFINE && tid == 0) L.ticket = atomicAdd(&a.tickets[pass], 1u);
  // The pass's geometry, from wave-uniform addresses (scalar loads) issued
  // while the ticket is in flight.
  const PassGeo* __restrict__ g = a.geo + pass;
  const unsigned digit = g->digit;
  const uint32_t total = g->total;
  uint32_t nt[kSegments];
#pragma unroll
  for (int s = 0; s < kSegments; ++s) nt[s] = g->nt[s];
  const unsigned shift = 8 * digit;
  const bool first = pass == 0;
  const bool last = uint32_t(pass) + 1 == n_live;

  for (unsigned i = tid; i < NW * 256; i += THREADS) (&L.whist[0][0])[i] = 0;
  if (tid < 256) L.tcnt[tid] = 0;
  if (LEARNED) load_knots(L.knots, a.model);
  uint32_t seg, round;
  S* next_row = status + size_t((pass + 1) & 1) * a.num_tiles * 256;
  if (TS_SWEEP_XCD_AFFINE) {
    // Claim a round of the segment numbered like this workgroup's XCD, so
    // that a segment's consecutive tiles, whose digit runs abut, write
    // through one L2; once it is exhausted, of the next segments in turn.
    // Within a segment rounds are claimed in order, so a tile's predecessor
    // was claimed first by a running workgroup.
    if (tid == 0) {
      uint32_t x;
      asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(x));
      uint32_t got = kSegments, r = 0;
      for (uint32_t i = 0; i < uint32_t(kSegments); ++i) {
        const uint32_t sg = (x + i) & uint32_t(kSegments - 1);
        if (nt[sg] == 0) continue;
        r = atomicAdd(&a.seg_tickets[pass * kSegments + sg], 1u);
        if (r < nt[sg]) {
          got = sg;
          break;
        }
      }
      L.ticket = (got << 24) | r;
    }
    __syncthreads();
    const uint32_t claim = __builtin_amdgcn_readfirstlane(L.ticket);
    // Rows of the next pass's look-back words: a claimed tile clears its
    // slot's, and workgroup b clears row b when b is no tile's slot.
    if (tid < 256 && blockIdx.x >= total && blockIdx.x < a.num_tiles) next_row[size_t(blockIdx.x) * 256 + tid] = 0;
    seg = claim >> 24;
    if (seg >= uint32_t(kSegments)) return;
    round = claim & 0xFFFFFFu;
    if (tid < 256) next_row[(size_t(g->slot[seg]) + round) * 256 + tid] = 0;
  } else {
    __syncthreads();
    const uint32_t ticket = __builtin_amdgcn_readfirstlane(L.ticket);
    if (tid < 256) next_row[size_t(ticket) * 256 + tid] = 0;  // the next pass's look-back row
    if (ticket >= total) return;
    ticket_to_tile(ticket, nt, &seg, &round);  // uniform: stays in scalar registers
    seg = __builtin_amdgcn_readfirstlane(seg);
    round = __builtin_amdgcn_readfirstlane(round);
  }

  const size_t seg_lo = g->start[seg];
  const size_t seg_hi = seg_lo + g->count[seg];
  const size_t tile_base = seg_lo + size_t(round) * TILE;
  const uint32_t valid = seg_hi - tile_base >= size_t(TILE) ? uint32_t(TILE) : uint32_t(seg_hi - tile_base);
  const size_t first_slot = g->slot[seg];
  const size_t slot = first_slot + round;
  // This lane's digit: its base and its count in earlier segments.
  const unsigned long long digit_base =
      tid < 256 ? meta->base[digit][tid] + meta->segpre[digit][seg][tid] : 0ull;
  if (valid == uint32_t(TILE))
    run_tile<U, VALS, S, true, LEARNED, SMALL>(L, a, status, pass, shift, first, last, tile_base, valid, first_slot,
                                                 slot, round == 0, digit_base);
  else
    run_tile<U, VALS, S, false, LEARNED, SMALL>(L, a, status, pass, shift, first, last, tile_base, valid, first_slot,
                                                  slot, round == 0, digit_base);
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
// reversed.
template <class U>
__global__ void k_copy_back(const SortMeta* m, const U* __restrict__ src, U* __restrict__ dst, size_t n) {
  if (m->reverse) return reverse_keys(dst, n);
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

template <class U, class S, bool LEARNED> hipError_t run_sweeps(const RadixJob& j, hipStream_t s, KernelClock* clk) {
  const bool vals = j.final_vals != nullptr;
  const uint32_t tiles = sweep_tiles(j.n, sizeof(U), LEARNED, vals);
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
  a.tickets = j.meta->tickets;
  a.seg_tickets = &j.meta->seg_tickets[0][0];
  a.error = &j.meta->error;
  a.num_tiles = tiles;
  a.model = j.model;
  a.atomic_rank = j.atomic_rank ? 1 : 0;
  const double bytes = double(j.n) * (2.0 * sizeof(U) + (vals ? 8.0 : 0.0));
  for (int p = 0; p < DigitSource<U, LEARNED>::P; ++p) {
    if (clk) clk->begin(s, "k_sweep", bytes);
    if constexpr (LEARNED)
      hipLaunchKernelGGL((k_sweep<U, false, S, true>), dim3(tiles), dim3(SweepShape<sizeof(U), true>::threads), 0, s, a,
                         static_cast<S*>(j.status), p);
    else if (vals)
      hipLaunchKernelGGL((k_sweep<U, true, S>), dim3(tiles), dim3(SweepShape<sizeof(U), false, true>::threads), 0, s, a,
                         static_cast<S*>(j.status), p);
    else if (small_sweep(j.n, false, false))
      hipLaunchKernelGGL((k_sweep<U, false, S, false, true>), dim3(tiles),
                         dim3(SweepShape<sizeof(U), false, false, true>::threads), 0, s, a, static_cast<S*>(j.status), p);
    else
      hipLaunchKernelGGL((k_sweep<U, false, S>), dim3(tiles), dim3(SweepShape<sizeof(U)>::threads), 0, s, a, static_cast<S*>(j.status), p);
    if (clk) clk->end(s);
  }
  return hipGetLastError();
}

template <class U, bool LEARNED = false> hipError_t run_radix(const RadixJob& j, hipStream_t s, KernelClock* clk) {
  constexpr int P = DigitSource<U, LEARNED>::P;
  constexpr int KB = int(sizeof(U));
  hipLaunchKernelGGL(k_init, dim3(1), dim3(256), 0, s, j.meta, P);
  // Histogram: blocks_per_seg workgroups per position segment, each over a
  // contiguous chunk of it.
  const bool vals = j.final_vals != nullptr;
  const size_t seg0 = first_segment(j.n, KB, LEARNED, vals);
  // Histogram workgroups per CU.  The learned histogram holds three digit
  // tables (and the knots), not eight, so more fit a CU;
  // TS_HIST_BLOCKS_PER_CU_LEARNED counts 256-lane workgroups.
  constexpr int kHistThreads = HistUnroll<U>::threads;
  const int hist_per_cu = LEARNED ? std::max(1, TS_HIST_BLOCKS_PER_CU_LEARNED * 256 / kHistThreads)
                                  : HistUnroll<U>::blocks_per_cu;
  const uint32_t blocks_per_seg = uint32_t(std::max(1, j.num_cus * hist_per_cu / kSegments));
  const size_t step = size_t(kHistThreads) * HistUnroll<U>::value;
  size_t per_block = (seg0 + blocks_per_seg - 1) / blocks_per_seg;
  per_block = std::max(step, ((per_block + step - 1) / step) * step);
  const uint32_t tiles = sweep_tiles(j.n, KB, LEARNED, vals);
  const size_t zero_words = size_t(tiles) * 256 * (wide_status(j.n) ? 2 : 1);
  const dim3 hgrid(blocks_per_seg * kSegments);
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
  if (clk) clk->begin(s, "k_hist", double(j.n) * (j.records ? double(j.stride) + KB : KB));
  if constexpr (LEARNED)
    hipLaunchKernelGGL((k_hist<U, false, true>), hgrid, dim3(kHistThreads), 0, s, static_cast<const U*>(j.keys_a),
                       nullptr, size_t(0), size_t(0), nullptr, j.n, j.mode, seg0, blocks_per_seg, per_block, j.meta,
                       static_cast<uint32_t*>(j.status), zero_words, j.model, check_order);
  else if (j.records)
    hipLaunchKernelGGL((k_hist<U, true>), hgrid, dim3(kHistThreads), 0, s, nullptr, j.records, j.stride,
                       j.key_offset, static_cast<U*>(j.keys_a), j.n, j.mode, seg0, blocks_per_seg, per_block, j.meta,
                       static_cast<uint32_t*>(j.status), zero_words, nullptr, check_order);
  else
    hipLaunchKernelGGL((k_hist<U, false>), hgrid, dim3(kHistThreads), 0, s, static_cast<const U*>(j.keys_a), nullptr,
                       size_t(0), size_t(0), nullptr, j.n, j.mode, seg0, blocks_per_seg, per_block, j.meta,
                       static_cast<uint32_t*>(j.status), zero_words, nullptr, check_order);
  if (clk) clk->end(s);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(256), 0, s, j.meta, j.n, P, seg0,
                     uint32_t(sweep_tile(KB, LEARNED, vals, small_sweep(j.n, LEARNED, vals))),
                     check_order);
  hipError_t e = wide_status(j.n) ? run_sweeps<U, uint64_t, LEARNED>(j, s, clk)
                                  : run_sweeps<U, uint32_t, LEARNED>(j, s, clk);
  if (e != hipSuccess) return e;
  const unsigned cb = unsigned(std::min<size_t>((j.n + 255) / 256, size_t(j.num_cus) * 8));
  if (j.final_vals) {
    hipLaunchKernelGGL(k_iota_if_dead, dim3(cb ? cb : 1), dim3(256), 0, s, j.meta, j.final_vals, j.n);
  } else {
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
