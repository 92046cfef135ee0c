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

// A 64-bit finaliser (the SplitMix64 mixer): the synthetic key streams and the
// learned sample's strata.
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// One key's share of the multiset fingerprint: its bits (zero-extended) plus
// a salt, through a one-multiply xor-shift mixer (cheap enough for the
// passes that take it on the side).  A fingerprint is the sum and the xor of
// these over the keys, which every permutation keeps.
__host__ __device__ __forceinline__ uint64_t fingerprint_of(uint64_t bits) {
  uint64_t x = bits + 0x632BE59BD9B4E019ull;
  x ^= x >> 32;
  x *= 0xD6E8FEB86659FD93ull;
  return x ^ (x >> 32);
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

// Device bookkeeping of one sort, decided on the device so that a sort needs
// no host round trip.
//
// Each scatter pass splits its input into kSegments contiguous segments and
// runs one decoupled look-back chain per segment (each chain is claimed by
// the workgroups of one XCD first, so it advances at a fraction of the tile
// rate).  The first executed pass reads the input in its original order and
// uses position segments; a later pass reads keys ordered by the previous
// digit and uses value segments, the top three bits of the previous digit.
// The segmented digit tables are filled in the single histogram read, and
// give the digit bases, each segment's first output position for each digit
// value (pbase) and the segment bounds.
constexpr int kSegments = 8;

// The geometry of one executed pass (the pass-th live digit), written by the
// scan and read by that pass's scatter kernel in its first instructions.
struct PassGeo {
  uint32_t digit;
  uint32_t total;  // tiles of the pass
  uint32_t nt[kSegments];
  uint32_t slot[kSegments];
  uint32_t pad[2];
  unsigned long long start[kSegments];
  unsigned long long count[kSegments];
};

// Sticky error bits that kernels OR into a context's device error word; the
// host reads and clears them (ts_ctx_check, and the calls that wait anyway).
constexpr uint32_t kErrLookback = 1u;    // a look-back wait hit its bound: the result is wrong
constexpr uint32_t kErrSegHint = 2u;     // a segment was longer than the caller's max_segment_len
constexpr uint32_t kErrSegOffsets = 4u;  // segment offsets decreasing or past n

struct SortMeta {
  PassGeo geo[8];
  // pair[p][s][x]: keys with digit p = x in value segment s of digit p's
  // pass (the top three bits of digit p-1; for p = 0 the position segment).
  unsigned long long pair[8][kSegments][256];
  // pair_pos[p][s][x]: keys with digit p = x in position segment s, for a
  // first executed pass on a digit above 0.
  unsigned long long pair_pos[8][kSegments][256];
  unsigned long long hist[8][256];  // per digit: all keys
  unsigned long long base[8][256];  // per digit: exclusive prefix of hist
  // Per executed pass, segment and digit value: the output position of the
  // first such key (digit base + the digit's count in earlier segments).
  unsigned long long pbase[8][kSegments][256];
  unsigned long long maxb[8];  // per digit: its largest bucket
  uint32_t seg_tickets[8][kSegments];  // per pass and segment: rounds claimed (XCD-affine claims)
  uint32_t n_live;  // executed passes
  uint32_t live[8];
  uint32_t error;  // non-zero: a look-back wait hit its bound (never expected)
  // Presorted input (presorted.rs:15-65): set by the histogram read when some
  // adjacent pair of radix keys decreases / increases; reverse: the scan
  // found the keys non-increasing and a reverse finishes the sort.
  uint32_t breaks_up, breaks_down, reverse;
  // Divert (divert.rs:38-90): skip = live digits left unsorted below the
  // executed ones (0: none); finish_shift = the lowest sorted bit; fallback:
  // the finish found a run longer than it sorts that is out of order, and
  // the gated second sort runs every digit.
  uint32_t skip, finish_shift, fallback;
  uint32_t all_live;  // live digits before the divert decision
  uint32_t diverted;  // the first sort's skip, kept when the second sort runs
  uint32_t pad[3];
  unsigned long long check[4];  // checksum scratch: sum, xor, unsorted pairs
};

// ---------------------------------------------------------------------------
// Learned partition model: a monotone piecewise-linear CDF over a key
// coordinate, evaluated in integer arithmetic so that the bucket of a key
// never decreases as the key grows.
//
// The coordinate of a radix key k is d = max(k - kmin, 0) (kmin: the sample's
// smallest key), then either log2_fixed(d) -- for integer keys spread over
// many orders of magnitude -- or d >> shift (linear; float keys, whose bit
// patterns already grow with the logarithm of the value).  Both are
// non-decreasing in k.  log2_fixed(d) = e * 2^23 + (the 23 bits below d's
// leading one), e its position.  The sample's coordinate range [0, xhi] is
// stretched onto u in [0, kKnots * 2^19); knot i covers u >> 19 == i, and
// F(u) = knots[i] + (knots[i+1] - knots[i]) * (u mod 2^19) / 2^19 (rounded
// down) never exceeds knots[i+1], so F is non-decreasing across knots too.
// A key's bucket is F >> bucket_shift.
constexpr int kKnots = 1024;
constexpr int kKnotFracBits = 19;  // u bits below the knot index

struct LearnedModel {
  uint32_t knots[kKnots + 1];  // F at each knot, 32-bit fixed point, non-decreasing
  uint32_t bucket_shift;       // 32 - log2(number of buckets)
  uint32_t linear;             // 1: coordinate d >> shift; 0: log2_fixed(d)
  uint32_t shift;
  uint32_t cmul, cshift;    // u = x * cmul >> cshift stretches [0, xhi] onto [0, 2^29)
  unsigned long long kmin;  // smallest sample key
  unsigned long long xhi;   // coordinate of the largest sample key (below 2^32)
};

// The coordinate parameters of a LearnedModel, copied into registers once by
// the kernels that evaluate the model for every key (read through the
// model's pointer, they would be reloaded -- and waited for -- per key).
struct LearnedCoord {
  unsigned long long kmin, xhi;
  uint32_t linear, shift, cmul, cshift, bucket_shift;
};

__host__ __device__ __forceinline__ LearnedCoord coord_of(const LearnedModel& m) {
  LearnedCoord c;
  c.kmin = m.kmin;
  c.xhi = m.xhi;
  c.linear = m.linear;
  c.shift = m.shift;
  c.cmul = m.cmul;
  c.cshift = m.cshift;
  c.bucket_shift = m.bucket_shift;
  return c;
}

#if defined(__HIP_DEVICE_COMPILE__)
#define TS_CLZ64(x) __clzll((long long)(x))
#else
#define TS_CLZ64(x) __builtin_clzll(x)
#endif

// e * 2^23 + (the 23 bits below k's leading one, zero-filled), e the
// leading one's position; below 2^29.
__host__ __device__ __forceinline__ uint32_t log2_fixed(unsigned long long k) {
  if (k == 0) return 0;
  const int lz = TS_CLZ64(k);
  return uint32_t(63 - lz) << 23 | (uint32_t((k << lz) >> 40) & 0x7FFFFFu);
}

template <class M>
__host__ __device__ __forceinline__ unsigned long long learned_x(unsigned long long key, const M& m) {
  const unsigned long long d = key > m.kmin ? key - m.kmin : 0ull;
  return m.linear ? d >> m.shift : log2_fixed(d);
}

// cmul and cshift for a coordinate range [0, xhi], xhi < 2^32:
// cmul = floor(2^(61 - t) / (xhi + 1)) and cshift = 32 - t, with the
// smallest t >= 0 that keeps cmul below 2^32 for every xhi of its bit length,
// so that x * cmul >> cshift < 2^29 for every x <= xhi (a 32 x 32-bit
// product).
__host__ __device__ __forceinline__ void learned_scale(unsigned long long xhi, uint32_t* cmul, uint32_t* cshift) {
  const int len = 64 - TS_CLZ64(xhi + 1);  // 2^(len-1) <= xhi + 1 < 2^len
  const int t = len >= 31 ? 0 : 31 - len;
  *cmul = uint32_t((1ull << (61 - t)) / (xhi + 1));
  *cshift = uint32_t(32 - t);
}

// The model coordinate of a radix key: u in [0, kKnots << kKnotFracBits).
// x * cmul >> cshift, computed as the high word of (x << (32 - cshift)) *
// cmul: the same value (x << (32 - cshift) stays below 2^32), one multiply.
template <class M>
__host__ __device__ __forceinline__ uint32_t learned_coord(unsigned long long key, const M& m) {
  unsigned long long x = learned_x(key, m);
  x = x > m.xhi ? m.xhi : x;
  const uint32_t xs = uint32_t(x) << (32 - m.cshift);
#if defined(__HIP_DEVICE_COMPILE__)
  return __umulhi(xs, m.cmul);
#else
  return uint32_t((unsigned long long)xs * m.cmul >> 32);
#endif
}

// The knots as evaluated on the device: entry i packs F at knot i (low word)
// and its rise to knot i + 1 (high word), so one read serves a key.
__host__ __device__ __forceinline__ unsigned long long knot_pair(const uint32_t* knots, int i) {
  return (unsigned long long)knots[i] | (unsigned long long)(knots[i + 1] - knots[i]) << 32;
}

// learned_cdf over knot_pair entries (the same value).
__host__ __device__ __forceinline__ uint32_t learned_cdf_pairs(uint32_t u, const unsigned long long* pairs) {
  const uint32_t i = u >> kKnotFracBits, f = u & ((1u << kKnotFracBits) - 1);
  const unsigned long long p = pairs[i];
  return uint32_t(p) + uint32_t((p >> 32) * f >> kKnotFracBits);
}

__host__ __device__ __forceinline__ uint32_t learned_cdf(uint32_t u, const uint32_t* knots) {
  const uint32_t i = u >> kKnotFracBits, f = u & ((1u << kKnotFracBits) - 1);
  const uint32_t y0 = knots[i], y1 = knots[i + 1];
  return y0 + uint32_t((unsigned long long)(y1 - y0) * f >> kKnotFracBits);
}

}  // namespace ts
