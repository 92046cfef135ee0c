// Device LSD radix sort for gfx950: one fused all-digit histogram pass, a
// one-block digit scan that also decides which digits are live, then one
// one-sweep scatter pass per live digit with decoupled look-back.
//
// Reference behaviour (turbosort 0.2.2): radix/mod.rs:205-271 (byte-digit
// LSD, live-pass skipping), histogram.rs:15-45 (all digits in one read),
// prefix_sum.rs:17-24 (exclusive scan), scatter.rs:26-40 (stable scatter by
// one digit).  The output is the unique ascending order of the radix keys.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "radix.h"

namespace ts {
namespace {

template <int KB> struct UOf;
template <> struct UOf<1> { using T = uint8_t; };
template <> struct UOf<2> { using T = uint16_t; };
template <> struct UOf<4> { using T = uint32_t; };
template <> struct UOf<8> { using T = uint64_t; };

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

template <class T> __device__ __forceinline__ T wave_inclusive_add(T v) {
  const unsigned lane = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T o = __shfl_up(v, d, 64);
    if (lane >= unsigned(d)) v += o;
  }
  return v;
}

// Exclusive scan of one value per lane over a 256-lane workgroup.
template <class T> __device__ __forceinline__ T block256_exclusive_add(T v, T* s_wave) {
  const unsigned tid = threadIdx.x, w = tid >> 6;
  T inc = wave_inclusive_add(v);
  if ((tid & 63) == 63) s_wave[w] = inc;
  __syncthreads();
  T off = 0;
  for (unsigned i = 0; i < w; ++i) off += s_wave[i];
  __syncthreads();
  return off + inc - v;
}

// ------------------------------------------------------------------ init
__global__ void k_init(SortMeta* m) {
  for (unsigned i = threadIdx.x; i < 8 * 256; i += blockDim.x) (&m->hist[0][0])[i] = 0;
  if (threadIdx.x < 8) m->tickets[threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    m->n_live = 0;
    m->error = 0;
  }
}

// ------------------------------------------------------------- histogram
// All P digit histograms in one read of the keys (histogram.rs:15-45).  Each
// workgroup counts a contiguous chunk into LDS and adds its counts to the
// global histogram once.  The forward key transform is fused into the load;
// the by-key path reads keys out of records and writes the transformed key
// array the scatter passes start from.  The grid also clears the first
// look-back row.
constexpr int kHistThreads = 256;
constexpr int kHistUnroll = 8;

template <class U>
__global__ __launch_bounds__(kHistThreads) void k_hist(const U* __restrict__ keys, const unsigned char* __restrict__ rec,
                                                       size_t stride, size_t key_off, U* __restrict__ keys_out,
                                                       size_t n, int mode, size_t per_block, SortMeta* m,
                                                       uint32_t* zero_ptr, size_t zero_words) {
  constexpr int P = int(sizeof(U));
  __shared__ uint32_t h[P * 256];
  const unsigned tid = threadIdx.x;
  for (unsigned i = tid; i < P * 256; i += kHistThreads) h[i] = 0;
  {
    const size_t g = size_t(blockIdx.x) * kHistThreads + tid, gs = size_t(gridDim.x) * kHistThreads;
    for (size_t z = g; z < zero_words; z += gs) zero_ptr[z] = 0;
  }
  __syncthreads();
  const size_t begin = size_t(blockIdx.x) * per_block;
  const size_t end = begin + per_block < n ? begin + per_block : n;
  for (size_t base = begin; base < end; base += size_t(kHistThreads) * kHistUnroll) {
    U k[kHistUnroll];
#pragma unroll
    for (int r = 0; r < kHistUnroll; ++r) {
      const size_t i = base + size_t(r) * kHistThreads + tid;
      U x = 0;
      if (i < end) x = rec ? *reinterpret_cast<const U*>(rec + i * stride + key_off) : keys[i];
      k[r] = to_key(x, mode);
    }
#pragma unroll
    for (int r = 0; r < kHistUnroll; ++r) {
      const size_t i = base + size_t(r) * kHistThreads + tid;
      if (i < end) {
        if (keys_out) keys_out[i] = k[r];
#pragma unroll
        for (int p = 0; p < P; ++p) atomicAdd(&h[p * 256 + unsigned((uint64_t(k[r]) >> (8 * p)) & 0xFF)], 1u);
      }
    }
  }
  __syncthreads();
  for (unsigned i = tid; i < P * 256; i += kHistThreads) {
    const uint32_t c = h[i];
    if (c) atomicAdd(&m->hist[0][0] + i, (unsigned long long)c);
  }
}

// ------------------------------------------------------------------ scan
// Digit bases (prefix_sum.rs:17-24) and live digits: a digit is dead when one
// bin holds every key, as the reference decides from the first key's bin
// (radix/mod.rs:229-239).
__global__ __launch_bounds__(256) void k_scan(SortMeta* m, size_t n, int P) {
  __shared__ unsigned long long s_wave[4];
  const unsigned tid = threadIdx.x;
  uint32_t n_live = 0;
  for (int p = 0; p < P; ++p) {
    const unsigned long long c = m->hist[p][tid];
    const unsigned long long e = block256_exclusive_add(c, s_wave);
    m->base[p][tid] = e;
    const int full = __syncthreads_or(c == (unsigned long long)n);
    if (!full) {
      if (tid == 0) m->live[n_live] = uint32_t(p);
      ++n_live;
    }
  }
  if (tid == 0) m->n_live = n_live;
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
  uint32_t* tickets;
  uint32_t* error;  // set when a look-back wait exceeds its bound
  uint32_t num_tiles;
};

// About two seconds of s_sleep(1) polling.
constexpr uint32_t kSpinLimit = 1u << 26;

// One stable scatter pass by the byte digit meta->live[pass] (scatter.rs:26-40).
//
// Tiles are claimed in launch order through a ticket, so every tile a tile
// waits on has already started.  Each wave ranks its 16x64 keys stably with
// eight ballots per key (lanes agreeing on all digit bits), counting into a
// wave-private LDS histogram; a cross-wave scan then gives tile-local ranks.
// The tile's digit counts are published, the exclusive prefix over earlier
// tiles is gathered by decoupled look-back, and the tile is staged in LDS in
// digit order so that each digit's keys leave as one coalesced run.
//
// Pass 0 of a by-key sort generates the values (positions 0..n-1); the last
// pass applies the inverse key transform, and in a by-key sort writes only the
// order to final_vals.
template <class U, bool VALS, class S>
__global__ __launch_bounds__(kSweepThreads) void k_sweep(SweepArgs<U> a, S* __restrict__ status, int pass) {
  constexpr int IPT = SweepShape<sizeof(U)>::ipt;
  constexpr int NW = kSweepThreads / 64;
  constexpr int TILE = kSweepThreads * IPT;
  constexpr int VB = Flags<S>::kBits;
  constexpr S kAgg = S(1) << VB, kPre = S(2) << VB, kMask = kAgg - 1;

  __shared__ uint32_t s_whist[NW][256];
  __shared__ U s_keys[TILE];
  __shared__ uint32_t s_vals[VALS ? TILE : 1];
  __shared__ uint32_t s_excl[256];
  __shared__ unsigned long long s_gofs[256];
  __shared__ uint32_t s_wave[NW];
  __shared__ uint32_t s_tile;

  const uint32_t n_live = a.meta->n_live;
  if (uint32_t(pass) >= n_live) return;
  const unsigned digit = a.meta->live[pass];
  const unsigned shift = 8 * digit;
  const bool first = pass == 0;
  const bool last = uint32_t(pass) + 1 == n_live;
  const U* __restrict__ src = (pass & 1) ? a.keys_b : a.keys_a;
  U* __restrict__ dst = (pass & 1) ? a.keys_a : a.keys_b;
  const uint32_t* __restrict__ vsrc = (pass & 1) ? a.vals_b : a.vals_a;
  uint32_t* __restrict__ vdst = (last && a.final_vals) ? a.final_vals : ((pass & 1) ? a.vals_a : a.vals_b);
  const bool write_keys = !(VALS && last && a.final_vals);

  const unsigned tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  if (tid == 0) s_tile = atomicAdd(&a.tickets[pass], 1u);
  for (unsigned i = tid; i < NW * 256; i += kSweepThreads) (&s_whist[0][0])[i] = 0;
  __syncthreads();
  const uint32_t tile = s_tile;
  S* row = status + size_t(pass & 1) * a.num_tiles * 256;
  S* next_row = status + size_t((pass + 1) & 1) * a.num_tiles * 256;
  next_row[size_t(tile) * 256 + tid] = 0;  // the next pass's look-back row

  const size_t n = a.n;
  const size_t tile_base = size_t(tile) * TILE;
  const size_t wbase = tile_base + size_t(w) * 64 * IPT;

  U k[IPT];
  uint32_t v[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const size_t i = wbase + size_t(j) * 64 + lane;
    k[j] = i < n ? (first ? to_key(src[i], a.mode_in) : src[i]) : U(~U(0));
    if (VALS) v[j] = first ? uint32_t(i) : (i < n ? vsrc[i] : 0u);
  }

  // Stable rank of each key among the wave's keys with the same digit.
  uint32_t rk[IPT];
  const uint64_t lt_mask = (1ull << lane) - 1;
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const bool ok = wbase + size_t(j) * 64 + lane < n;
    const unsigned d = unsigned(uint64_t(k[j]) >> shift) & 0xFF;
    uint64_t same = __ballot(ok);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1;
      const uint64_t vote = __ballot(bit);
      same &= bit ? vote : ~vote;
    }
    const uint64_t below = same & lt_mask;
    const uint32_t before = s_whist[w][d];
    if (ok && below == 0) s_whist[w][d] = before + uint32_t(__popcll(same));
    rk[j] = before + uint32_t(__popcll(below));
  }
  __syncthreads();

  // Per digit: exclusive offsets of the waves, and the tile's count.
  uint32_t cnt = 0;
  for (int ww = 0; ww < NW; ++ww) {
    const uint32_t c = s_whist[ww][tid];
    s_whist[ww][tid] = cnt;
    cnt += c;
  }
  S* my = row + size_t(tile) * 256 + tid;
  __hip_atomic_store(my, S(cnt) | (tile == 0 ? kPre : kAgg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint32_t excl_in_tile = block256_exclusive_add(cnt, s_wave);
  s_excl[tid] = excl_in_tile;

  // Decoupled look-back over earlier tiles for this lane's digit.
  // Every tile with a smaller ticket is running and publishes without waiting
  // on later tiles, so the walk ends; the spin bound only turns a broken
  // invariant into a flagged wrong result instead of a hung GPU.
  unsigned long long before_tiles = 0;
  if (tile > 0) {
    int64_t j = int64_t(tile) - 1;
    uint32_t spins = 0;
    while (true) {
      const S st = __hip_atomic_load(row + size_t(j) * 256 + tid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const S fl = st >> VB;
      if (fl == 0) {
        if (++spins > kSpinLimit) {
          atomicOr(a.error, 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        continue;
      }
      before_tiles += (unsigned long long)(st & kMask);
      if (fl == 2 || j == 0) break;
      --j;
    }
    __hip_atomic_store(my, S(before_tiles + cnt) | kPre, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  s_gofs[tid] = a.meta->base[digit][tid] + before_tiles - excl_in_tile;
  __syncthreads();

  // Stage the tile in digit order.
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const bool ok = wbase + size_t(j) * 64 + lane < n;
    if (ok) {
      const unsigned d = unsigned(uint64_t(k[j]) >> shift) & 0xFF;
      const uint32_t pos = s_excl[d] + s_whist[w][d] + rk[j];
      s_keys[pos] = k[j];
      if (VALS) s_vals[pos] = v[j];
    }
  }
  __syncthreads();

  // Write each digit's run of the tile contiguously.
  const uint32_t valid = n - tile_base < size_t(TILE) ? uint32_t(n - tile_base) : uint32_t(TILE);
  const int out_mode = last ? a.mode_out : int(kIdentity);
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t s = uint32_t(j) * kSweepThreads + tid;
    if (s < valid) {
      const U kk = s_keys[s];
      const unsigned d = unsigned(uint64_t(kk) >> shift) & 0xFF;
      const size_t o = size_t(s_gofs[d]) + s;
      if (write_keys) dst[o] = from_key(kk, out_mode);
      if (VALS) vdst[o] = s_vals[s];
    }
  }
}

// An odd number of live passes leaves the keys in the scratch half.
template <class U>
__global__ void k_copy_back(const SortMeta* m, const U* __restrict__ src, U* __restrict__ dst, size_t n) {
  if ((m->n_live & 1u) == 0) return;
  const size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x, gs = size_t(gridDim.x) * blockDim.x;
  for (size_t i = g; i < n; i += gs) dst[i] = src[i];
}

// With no live digit every key is equal: the stable order is the identity.
__global__ void k_iota_if_dead(const SortMeta* m, uint32_t* out, size_t n) {
  if (m->n_live != 0) return;
  const size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x, gs = size_t(gridDim.x) * blockDim.x;
  for (size_t i = g; i < n; i += gs) out[i] = uint32_t(i);
}

template <class U, class S> hipError_t run_sweeps(const RadixJob& j, hipStream_t s, KernelClock* clk) {
  const uint32_t tiles = sweep_tiles(j.n, sizeof(U));
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
  a.tickets = j.meta->tickets;
  a.error = &j.meta->error;
  a.num_tiles = tiles;
  const bool vals = j.final_vals != nullptr;
  const double bytes = double(j.n) * (2.0 * sizeof(U) + (vals ? 8.0 : 0.0));
  for (int p = 0; p < int(sizeof(U)); ++p) {
    if (clk) clk->begin(s, "k_sweep", bytes);
    if (vals)
      hipLaunchKernelGGL((k_sweep<U, true, S>), dim3(tiles), dim3(kSweepThreads), 0, s, a, static_cast<S*>(j.status), p);
    else
      hipLaunchKernelGGL((k_sweep<U, false, S>), dim3(tiles), dim3(kSweepThreads), 0, s, a, static_cast<S*>(j.status), p);
    if (clk) clk->end(s);
  }
  return hipGetLastError();
}

template <class U> hipError_t run_radix(const RadixJob& j, hipStream_t s, KernelClock* clk) {
  constexpr int P = int(sizeof(U));
  hipLaunchKernelGGL(k_init, dim3(1), dim3(256), 0, s, j.meta);
  // Histogram: a few workgroups per CU, each over a contiguous chunk.
  const size_t step = size_t(kHistThreads) * kHistUnroll;
  const size_t want_blocks = size_t(j.num_cus) * 4;
  size_t per_block = (j.n + want_blocks - 1) / want_blocks;
  per_block = ((per_block + step - 1) / step) * step;
  if (per_block == 0) per_block = step;
  const size_t blocks = (j.n + per_block - 1) / per_block;
  const uint32_t tiles = sweep_tiles(j.n, P);
  const size_t zero_words = size_t(tiles) * 256 * (wide_status(j.n) ? 2 : 1);
  if (clk) clk->begin(s, "k_hist", double(j.n) * (j.records ? double(j.stride) + P : P));
  hipLaunchKernelGGL((k_hist<U>), dim3(blocks ? blocks : 1), dim3(kHistThreads), 0, s, static_cast<const U*>(j.keys_a),
                     j.records, j.stride, j.key_offset, j.records ? static_cast<U*>(j.keys_a) : nullptr, j.n, j.mode,
                     per_block, j.meta, static_cast<uint32_t*>(j.status), zero_words);
  if (clk) clk->end(s);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(256), 0, s, j.meta, j.n, P);
  hipError_t e = wide_status(j.n) ? run_sweeps<U, uint64_t>(j, s, clk) : run_sweeps<U, uint32_t>(j, s, clk);
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

}  // namespace

hipError_t radix_sort(const RadixJob& job, hipStream_t s, KernelClock* clk) {
  switch (job.key_bytes) {
    case 1: return run_radix<uint8_t>(job, s, clk);
    case 2: return run_radix<uint16_t>(job, s, clk);
    case 4: return run_radix<uint32_t>(job, s, clk);
    case 8: return run_radix<uint64_t>(job, s, clk);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace ts
