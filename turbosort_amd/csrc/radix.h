// Host-side entry points of the device radix sort (radix.hip) and the other
// kernel files, used by the C ABI in api.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "keys.h"

namespace ts {

// Records the duration of each kernel of a sort with HIP events on the sort's
// stream, when enabled on the context.
struct KernelClock {
  static constexpr int kMax = 256;
  bool on = false;
  int count = 0;
  const char* name[kMax];
  double bytes[kMax];
  hipEvent_t ev[2 * kMax];
  bool made = false;
  void begin(hipStream_t s, const char* nm, double b) {
    if (!on || count >= kMax) return;
    name[count] = nm;
    bytes[count] = b;
    (void)hipEventRecord(ev[2 * count], s);
  }
  void end(hipStream_t s) {
    if (!on || count >= kMax) return;
    (void)hipEventRecord(ev[2 * count + 1], s);
    ++count;
  }
};

// Tile geometry of the one-sweep scatter: `threads` lanes per workgroup (a
// multiple of 256; the first 256 lanes also carry one digit each through the
// per-digit steps) and `ipt` keys per lane, per key width (4-byte and
// narrower / 8-byte) and kind: keys only, keys with values (both staged in
// LDS), keys ordered by learned buckets (whose model knots also take LDS).  Measured on MI355X at 2^28 keys with LDS-atomic ranking:
// the passes are bound by memory latency and the scattered write runs, and
// larger tiles (longer runs per digit, fewer look-back steps) win down to one
// workgroup per CU.
#ifndef TS_SWEEP_KEYS_NARROW
#define TS_SWEEP_KEYS_NARROW 1024, 32
#endif
#ifndef TS_SWEEP_KEYS_WIDE
#define TS_SWEEP_KEYS_WIDE 512, 36
#endif
#ifndef TS_SWEEP_VALS_NARROW
#define TS_SWEEP_VALS_NARROW 512, 36
#endif
#ifndef TS_SWEEP_VALS_WIDE
#define TS_SWEEP_VALS_WIDE 512, 24
#endif
#ifndef TS_SWEEP_LEARNED_NARROW
#define TS_SWEEP_LEARNED_NARROW 512, 32
#endif
#ifndef TS_SWEEP_LEARNED_WIDE
#define TS_SWEEP_LEARNED_WIDE 1024, 16
#endif
// Keys-only sorts of at most TS_SWEEP_SMALL_N keys use 8192-key tiles, so
// that their tiles cover more CUs (measured on MI355X, u32: 100 against
// 115 us at 2^20 keys, 129 against 134 us at 2^21; larger sorts are faster
// with the large tiles, and 4096-key tiles are slower from 2^20 on).
#ifndef TS_SWEEP_KEYS_SMALL
#define TS_SWEEP_KEYS_SMALL 512, 16
#endif
#ifndef TS_SWEEP_SMALL_N
#define TS_SWEEP_SMALL_N (size_t(1) << 21)
#endif
// By-key sorts of 8- and 16-byte records carry the records themselves through
// the passes (each pass reads the key out of the staged record), with or
// without the u32 order beside them; LDS holds a tile of records (and order
// values), no separate keys.
#ifndef TS_SWEEP_REC8_VALS
#define TS_SWEEP_REC8_VALS 512, 24
#endif
#ifndef TS_SWEEP_REC8
#define TS_SWEEP_REC8 512, 36
#endif
#ifndef TS_SWEEP_REC16_VALS
#define TS_SWEEP_REC16_VALS 512, 14
#endif
#ifndef TS_SWEEP_REC16
#define TS_SWEEP_REC16 512, 16
#endif
// Predecessor tiles one look-back step reads at once.
#ifndef TS_LOOKBACK_WINDOW
#define TS_LOOKBACK_WINDOW 4
#endif
template <int THREADS, int IPT> struct ShapeOf {
  static constexpr int threads = THREADS;
  static constexpr int ipt = IPT;
};
template <int RW, bool VALS>
using RecShape = std::conditional_t<RW == 4, std::conditional_t<VALS, ShapeOf<TS_SWEEP_REC16_VALS>, ShapeOf<TS_SWEEP_REC16>>,
                                    std::conditional_t<VALS, ShapeOf<TS_SWEEP_REC8_VALS>, ShapeOf<TS_SWEEP_REC8>>>;
// RW: 32-bit words of the records a by-key sort carries (0: keys move, the
// records are gathered afterwards; 2 or 4).
template <int KB, bool LEARNED = false, bool VALS = false, bool SMALL = false, int RW = 0>
struct SweepShape
    : std::conditional_t<
          (RW > 0), RecShape<RW, VALS>,
          std::conditional_t<
              SMALL, ShapeOf<TS_SWEEP_KEYS_SMALL>,
              std::conditional_t<
                  (KB >= 8),
                  std::conditional_t<LEARNED, ShapeOf<TS_SWEEP_LEARNED_WIDE>,
                                     std::conditional_t<VALS, ShapeOf<TS_SWEEP_VALS_WIDE>, ShapeOf<TS_SWEEP_KEYS_WIDE>>>,
                  std::conditional_t<LEARNED, ShapeOf<TS_SWEEP_LEARNED_NARROW>,
                                     std::conditional_t<VALS, ShapeOf<TS_SWEEP_VALS_NARROW>,
                                                        ShapeOf<TS_SWEEP_KEYS_NARROW>>>>>> {
  static_assert(SweepShape::threads % 256 == 0, "the first 256 lanes carry one digit each");
  static_assert(!SMALL || (!LEARNED && !VALS && RW == 0), "small tiles are for keys-only sorts");
  static_assert(RW == 0 || RW == 2 || RW == 4, "records of 8 or 16 bytes are carried");
  static_assert(RW == 0 || !LEARNED, "learned digits come from keys");
  static constexpr int tile = SweepShape::threads * SweepShape::ipt;
};

// Tile geometry on the host side.  These are defined in radix.hip, the one
// translation unit that instantiates the scatter kernels, so that workspace
// sizing and launch shapes always come from the same tile constants.
//
// The kind of a sort picks its tile shape: keys only, keys with u32 order
// values (by-key), learned bucket digits, or 8- / 16-byte records carried
// with the order (by-key) or without it (by-key sorts beyond u32::MAX).
enum SweepKind : int {
  kSweepKeys = 0,
  kSweepVals = 1,
  kSweepLearned = 2,
  kSweepRec8Vals = 3,
  kSweepRec16Vals = 4,
  kSweepRec8 = 5,
  kSweepRec16 = 6,
};
// Whether a sort of n keys runs the small-tile shape.
bool small_sweep(size_t n, bool learned, bool vals);
// Keys per tile of the shape a sort of n keys of this kind runs (rec_words:
// the 32-bit words of carried records, 0 when keys move).
size_t sweep_tile_for(size_t n, int key_bytes, bool learned, bool vals, int rec_words = 0);
// Upper bound on the tiles of one pass of a sort of n keys: every segment may
// end in a partial tile.
uint32_t sweep_tiles(size_t n, int key_bytes, bool learned, bool vals, int rec_words = 0);
// Look-back words are 32-bit (2 flag bits, 30 value bits) while every count
// fits in 30 bits, else 64-bit.
bool wide_status(size_t n);
// Bytes of look-back status that one radix sort of n keys of the given kind
// uses (two alternating rows).
size_t status_bytes_used(size_t n, int key_bytes, bool learned, bool vals, int rec_words = 0);
// Bytes of look-back status to reserve so that every radix sort of AT MOST n
// keys of this width, of the keys, values or learned kinds, fits: shorter
// sorts may run the small tiles, which need more rows per key than the large
// ones.  (Record-carrying sorts reserve status_bytes_used of their own kind.)
size_t status_bytes(size_t n, int key_bytes);

// One LSD sort on the device.  Keys start in keys_a (plain sort) or are read
// from records (by-key sort; keys_a then receives the transformed keys).
struct RadixJob {
  int key_bytes = 4;
  int mode = kIdentity;  // transform of the source keys
  size_t n = 0;
  const unsigned char* records = nullptr;  // by-key: record base
  size_t stride = 0, key_offset = 0;
  void* keys_a = nullptr;
  void* keys_b = nullptr;
  uint32_t* vals_a = nullptr;  // by-key: order ping-pong
  uint32_t* vals_b = nullptr;
  uint32_t* final_vals = nullptr;  // by-key: the last pass writes order here
  // By-key sorts of 8- or 16-byte records (stride 4 * rec_words, aligned to
  // it): the records move through the passes, ping-ponging between
  // `records` (where they end, in stable key order) and recs_b, and every
  // pass reads the key at key_offset of each record; keys_a / keys_b are
  // unused.  With final_vals the order is written too.
  int rec_words = 0;
  void* recs_b = nullptr;
  SortMeta* meta = nullptr;
  void* status = nullptr;
  size_t status_cap = 0;  // bytes at `status`; a sort needing more is refused
  // Sticky error bits (device word; kErr* below), kept across sorts until the
  // host reads and clears them.  Null: meta->error, cleared by each sort.
  uint32_t* error = nullptr;
  int num_cus = 256;
  // Keys-only sorts of 8-byte keys: sort only the top live digits that carry
  // about log2(n) + 2 bits, then finish the runs that share them (the
  // reference's divert, radix/mod.rs:240-242,268-270, divert.rs:38-90).
  bool divert = false;
  // Learned partition: the passes sort by the bytes of each key's model
  // bucket (up to 24 bits) instead of by the key (keys only; 4- and 8-byte
  // keys).
  const LearnedModel* model = nullptr;
  // Rank keys within a wave by LDS atomic adds instead of ballots; only where
  // rank_order_check found the adds applied in lane order.
  bool atomic_rank = false;
  // Learned partition: the histogram read also adds the keys' multiset
  // fingerprint (sum, xor of fingerprint_of) into fingerprint[0..1], which
  // the caller has zeroed (null: not taken).
  unsigned long long* fingerprint = nullptr;
};

hipError_t radix_sort(const RadixJob& job, hipStream_t s, KernelClock* clk);

// Adds to *bad (device memory, zeroed by the caller) the number of keys whose
// LDS-atomic rank differs from their ballot rank over a set of digit patterns.
hipError_t rank_order_check(uint32_t* bad, hipStream_t s);

// Small-input tier: one workgroup sorts n <= kSmallMax keys in LDS with a
// bitonic network.  With order != nullptr it sorts (key, position) pairs and
// writes the stable order instead of moving keys (by-key path).
constexpr size_t kSmallMax = 4096;
hipError_t small_sort(int key_bytes, int mode, void* keys, size_t n, const unsigned char* records, size_t stride,
                      size_t key_offset, uint32_t* order, hipStream_t s, KernelClock* clk);

// out[i] = src[order[i]] for n elements of elem_bytes bytes.
hipError_t gather(const uint32_t* order, const void* src, void* dst, size_t n, size_t elem_bytes, hipStream_t s,
                  KernelClock* clk);
// The same with the positions taken from words 2-3 of n 16-byte records
// (pack_key_pos records, sorted by key).
hipError_t gather_packed(const void* packed, const void* src, void* dst, size_t n, size_t elem_bytes, hipStream_t s,
                         KernelClock* clk);
// out[i] = (the key_bytes-wide key at keys + i * stride + key_offset,
// zero-extended to 64 bits; i as 64 bits): 16 bytes per element.
hipError_t pack_key_pos(int key_bytes, const void* keys, size_t stride, size_t key_offset, void* out, size_t n,
                        hipStream_t s);

// Batched segmented sort (segmented.hip).  Segments of up to
// seg_lds_max_len keys are sorted inside one workgroup each; longer ones are
// listed in class kSegDevice (count in counts[kSegDevice], ids from
// lists + kSegDevice * n_segments) for the caller to sort with the device
// radix sort.  counts[kSegClasses] becomes non-zero when an offset pair is
// decreasing or past n.
constexpr int kSegClasses = 10;
constexpr int kSegDevice = kSegClasses - 1;
struct SegJob {
  int key_bytes = 4;
  int mode = kIdentity;
  void* keys = nullptr;
  uint32_t* vals = nullptr;  // null: keys only
  size_t n = 0;
  const unsigned long long* offsets = nullptr;  // n_segments + 1 entries (segment starts)
  const unsigned long long* ends = nullptr;     // optional: segment s is [offsets[s], ends[s])
  size_t n_segments = 0;
  uint32_t* counts = nullptr;  // kSegClasses + 2 words
  uint32_t* lists = nullptr;   // kSegClasses * n_segments words
  int num_cus = 256;
  uint8_t variant[kSegClasses] = {};  // kernel shape per size class (0: default)
  bool atomic_rank = false;           // as RadixJob::atomic_rank
  // Sticky error word (as RadixJob::error): the classify ORs in
  // kErrSegOffsets for invalid offsets and kErrSegHint for a segment longer
  // than max_len (0: no bound given).
  uint32_t* error = nullptr;
  size_t max_len = 0;
  // 0: classify and sort; 1: classify only (counts and lists filled);
  // 2: sort the classes of an earlier stage-1 call.
  int stage = 0;
  // Size classes to sort (bit c: class c); a caller that knows some classes
  // are empty may leave them out and save their launches.
  uint32_t class_mask = ~0u;
};
uint32_t seg_lds_max_len(int key_bytes, bool vals);
// The keys-only sort of kSmallMax < n <= seg_lds_max_len(key_bytes, false)
// keys inside one workgroup (segmented.hip): the keys-only class kernel for
// that length (variant[c] picks its shape) on a one-segment class list that
// a one-wave kernel writes into `area` (kOneSegmentBytes of device memory).
constexpr size_t kOneSegmentBytes = 256;
hipError_t lds_sort(int key_bytes, int mode, void* keys, size_t n, void* area, const uint8_t* variant,
                    bool atomic_rank, uint32_t* error, hipStream_t s, KernelClock* clk);
// Largest segment of size class c (0: a class this width does not use).
uint32_t seg_class_cap(int key_bytes, bool vals, int size_class);
// Kernel shapes a size class can run (variant 0 is the default), and a name
// for each.
int seg_variants(int key_bytes, bool vals, int size_class);
const char* seg_variant_name(int key_bytes, bool vals, int size_class, int variant);

// Three-way split of long keys-only segments (segmented.hip): segment s is
// [lo[s], lo[s] + len[s]), cut into chunks of split_chunk() keys numbered
// chunk_start[s] .. chunk_start[s + 1] - 1 (chunks in total).  Keys below a
// pivot (a median of five) move to the front, keys above it to the back, the
// middle is filled with the pivot; below[s] and above[s] receive the counts.
struct SplitJob {
  int key_bytes = 8;
  int mode = kIdentity;
  void* data = nullptr;
  void* tmp = nullptr;  // as long as data
  uint32_t nseg = 0;
  uint32_t chunks = 0;
  const unsigned long long* lo = nullptr;
  const unsigned long long* len = nullptr;
  const uint32_t* chunk_start = nullptr;  // nseg entries
  unsigned long long* below = nullptr;
  unsigned long long* above = nullptr;
  unsigned long long* pivot = nullptr;
};
uint32_t split_chunk();
hipError_t split_segments(const SplitJob& job, hipStream_t s);
hipError_t segmented_sort(const SegJob& job, hipStream_t s, KernelClock* clk);

// After a classify (stage 0 or 1), one block for the host to read back: a
// kSegLongHeader-byte header holding counts[0 .. kSegClasses + 1], then
// extra_words words copied from `extra`; then (start, length) of the first
// min(counts[kSegDevice], max_pairs) segments of the device class.
constexpr size_t kSegLongHeader = 128;
hipError_t seg_long_bounds(const SegJob& job, unsigned long long* out, uint32_t max_pairs, const uint32_t* extra,
                           uint32_t extra_words, hipStream_t s);

// Learned partition (learned.hip): sample, fit, then group the keys of `data`
// by model bucket in place and write the bucket bounds.  Every buffer is
// caller-provided device memory.
struct LearnedJob {
  int key_bytes = 8;
  int mode = kIdentity;
  void* data = nullptr;
  void* scratch = nullptr;  // n keys
  size_t n = 0;
  size_t sample = 0;
  unsigned long long* sample_a = nullptr;  // sample keys, two halves
  unsigned long long* sample_b = nullptr;
  uint32_t points = 0;  // training points
  float* u = nullptr;
  float* t = nullptr;
  float* theta = nullptr;  // kKnots
  float* grad = nullptr;   // kKnots
  float* loss = nullptr;   // 3: at the start, after fitting, of uniform knots
  unsigned long long* knot_first = nullptr;  // kKnots + 1
  bool train = true;
  int iters = 300;
  float lr = 0.05f;
  LearnedModel* model = nullptr;
  uint32_t bucket_shift = 16;
  uint32_t buckets = 65536;
  unsigned long long* offsets = nullptr;  // buckets + 1
  unsigned long long* stats = nullptr;    // [0]: keys in the largest bucket
  SortMeta* meta = nullptr;
  void* status = nullptr;  // look-back status for the sample's sort and the partition
  size_t status_cap = 0;   // bytes at `status`
  uint32_t* error = nullptr;  // as RadixJob::error
  int num_cus = 256;
  bool atomic_rank = false;  // as RadixJob::atomic_rank
  unsigned long long* fingerprint = nullptr;  // as RadixJob::fingerprint, for the partition's histogram
};
hipError_t learned_partition(const LearnedJob& job, hipStream_t s, KernelClock* clk);

// One fit of the piecewise-linear CDF (see k_pwl_fit): theta in/out, the
// gradient at the initial theta, the loss before and after `iters` steps.
hipError_t pwl_fit(const float* u, const float* t, uint32_t T, float* theta, float* grad, float* loss, int iters,
                   float lr, hipStream_t s);

hipError_t fill_splitmix64(void* out, size_t n, uint64_t seed, int kind, hipStream_t s);
// Order-independent fingerprint of n keys into out[0..2] (device memory):
// sum and xor of mix64(bits + salt), and the adjacent pairs out of order.
hipError_t checksum(int key_bytes, int mode, const void* data, size_t n, unsigned long long* out, hipStream_t s,
                    KernelClock* clk = nullptr);
// Zeroes the three words of a fingerprint slot (a kernel, not a memset).
hipError_t fingerprint_zero(unsigned long long* out, hipStream_t s);

}  // namespace ts
