// C ABI (include/tsort.h): argument checks with the reference's panic
// messages, workspace management and tier dispatch.
//
// Tiers (dispatch.rs:31-67): n <= 1 is a no-op; up to kSmallMax keys one
// workgroup sorts in LDS (the reference's tiny/small network tiers); up to
// one workgroup's LDS capacity (16384 keys of 4 bytes, 17408 of 8) keys-only
// sorts take the segmented sort's kernel for that length on one segment;
// above that the device LSD radix sort runs.  The reference's presorted scan and u8
// counting sort are output-invisible shortcuts (presorted.rs:10-14): the
// radix sort produces the same unique output for those inputs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/tsort.h"
#include "radix.h"

struct ts_ctx {
  int device = 0;
  int num_cus = 256;
  char* ws = nullptr;
  size_t ws_bytes = 0;
  std::string err;
  ts::KernelClock clk;
  // Segmented sort: kernel shape per (key width 1/2/4/8, values, size class).
  uint8_t seg_variant[4][2][ts::kSegClasses] = {};
  // Whether this device applies a wave's LDS atomic adds to one address in
  // lane order (rank_order_check), and whether the sorts use that ranking.
  bool atomic_rank_ok = false;
  bool atomic_rank = false;
  // Three-way split rounds for long keys-only segments before the radix sort.
  int split_rounds = 12;
  // ts_sort_host: device copy of the caller's host slice (grow-only).
  char* hdev = nullptr;
  size_t hdev_bytes = 0;
  // ts_learned_sort: a second stream (default priority) on which the
  // host-driven rounds for overflowing buckets run beside the bucket sort.
  hipStream_t aux = nullptr;
  hipEvent_t aux_done = nullptr;
  hipEvent_t back_ready = nullptr;  // the learned sort's readback has landed
  // Pinned host memory for readbacks and the split rounds' segment lists
  // (grow-only), so those copies are direct transfers.
  char* pin = nullptr;
  size_t pin_bytes = 0;
  // Device block of 256 bytes: the sticky error word (kErr* bits, read and
  // cleared by ts_ctx_check and by the calls that wait for the device
  // anyway) at byte 0, then two fingerprints of three words each (input and
  // output of a judged sort) at kFpIn and kFpOut.
  uint32_t* dflags = nullptr;
  // Sorts of 8-byte keys without values divert from this many keys on.
  size_t divert_min = size_t(1) << 23;
  // Bookkeeping of the last radix sort (the workspace, or the static
  // scratch), for ts_ctx_last_sort_info.
  const ts::SortMeta* last_meta = nullptr;
};

namespace {

constexpr size_t kFpIn = 64, kFpOut = 128;  // byte offsets in ts_ctx::dflags
unsigned long long* fp_slot(const ts_ctx* c, size_t off) {
  return reinterpret_cast<unsigned long long*>(reinterpret_cast<char*>(c->dflags) + off);
}

constexpr size_t kAlign = 256;
inline size_t up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Runs a call with the context's device current and restores the caller's.
struct DeviceScope {
  int prev = -1;
  explicit DeviceScope(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) hipSetDevice(dev);
  }
  ~DeviceScope() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) hipSetDevice(prev);
  }
};

ts_status fail(ts_ctx* c, ts_status st, const std::string& msg) {
  if (c) c->err = msg;
  return st;
}

ts_status hip_fail(ts_ctx* c, hipError_t e, const char* where) {
  return fail(c, TS_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

ts_status ensure(ts_ctx* c, size_t bytes) {
  if (bytes <= c->ws_bytes) return TS_OK;
  if (c->ws) {
    hipError_t e = hipFree(c->ws);  // waits for work still using the old workspace
    if (c->last_meta && reinterpret_cast<const char*>(c->last_meta) >= c->ws &&
        reinterpret_cast<const char*>(c->last_meta) < c->ws + c->ws_bytes)
      c->last_meta = nullptr;
    c->ws = nullptr;
    c->ws_bytes = 0;
    if (e != hipSuccess) return hip_fail(c, e, "hipFree");
  }
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&c->ws), bytes);
  if (e != hipSuccess) {
    c->ws = nullptr;
    return hip_fail(c, e, "hipMalloc(workspace)");
  }
  c->ws_bytes = bytes;
  return TS_OK;
}

// Grows the pinned host area; copies from an earlier call may still read
// the old one, so the device is drained before it is freed.
ts_status ensure_pinned(ts_ctx* c, size_t bytes) {
  if (bytes <= c->pin_bytes) return TS_OK;
  if (c->pin) {
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipHostFree(c->pin);
    c->pin = nullptr;
    c->pin_bytes = 0;
    if (e != hipSuccess) return hip_fail(c, e, "hipHostFree");
  }
  hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&c->pin), bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    c->pin = nullptr;
    return hip_fail(c, e, "hipHostMalloc(staging)");
  }
  c->pin_bytes = bytes;
  return TS_OK;
}

bool valid_dtype(int dt) { return dt >= TS_U8 && dt <= TS_F64; }

// What a set of sticky error bits means, for ts_ctx_last_error.
std::string flag_text(uint32_t f) {
  std::string m = "device check failed:";
  if (f & ts::kErrLookback) m += " a look-back wait hit its bound (the sort's output is wrong);";
  if (f & ts::kErrSegHint) m += " a segment was longer than max_segment_len (any beyond one workgroup's capacity was left unsorted);";
  if (f & ts::kErrSegOffsets) m += " segment offsets were decreasing or past n (those segments were skipped);";
  return m;
}

ts_status judge_verdict(ts_ctx* c, const unsigned long long* got, const uint64_t* expect) {
  if (got[2] != 0)
    return fail(c, TS_ERR_VERIFY_FAILED,
                "output not sorted: " + std::to_string(got[2]) + " adjacent pairs out of order");
  if (expect && (got[0] != expect[0] || got[1] != expect[1]))
    return fail(c, TS_ERR_VERIFY_FAILED, "output keys differ from the input's (multiset fingerprint)");
  return TS_OK;
}

// Fingerprints data into fp_slot(c, kFpOut), waits, and compares it with
// `expect` (sum, xor; null: sortedness only).
ts_status verify_output(ts_ctx* c, int kb, int mode, const void* data, size_t n, const uint64_t* expect,
                        hipStream_t s) {
  unsigned long long* out = fp_slot(c, kFpOut);
  hipError_t e = ts::checksum(kb, mode, data, n, out, s, &c->clk);
  unsigned long long h[3] = {0, 0, 0};
  if (e == hipSuccess) e = hipMemcpyAsync(h, out, sizeof h, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(c, e, "verify");
  return judge_verdict(c, h, expect);
}

// Waits for stream s, then reads and clears the context's sticky error word.
ts_status check_flags(ts_ctx* c, hipStream_t s) {
  uint32_t f = 0;
  hipError_t e = hipMemcpyAsync(&f, c->dflags, sizeof f, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(c, e, "error word readback");
  if (f == 0) return TS_OK;
  e = hipMemsetAsync(c->dflags, 0, sizeof f, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(c, e, "error word clear");
  return fail(c, TS_ERR_VERIFY_FAILED, flag_text(f));
}

// A radix job on the context's device with its sticky error word.
ts::RadixJob radix_job(ts_ctx* c, int kb, int mode, size_t n, char* meta, char* status, size_t status_cap) {
  c->last_meta = reinterpret_cast<const ts::SortMeta*>(meta);
  ts::RadixJob j;
  j.key_bytes = kb;
  j.mode = mode;
  j.n = n;
  j.meta = reinterpret_cast<ts::SortMeta*>(meta);
  j.status = status;
  j.status_cap = status_cap;
  j.error = c->dflags;
  j.num_cus = c->num_cus;
  j.atomic_rank = c->atomic_rank;
  return j;
}

// Workspace layouts.  meta | look-back status | arrays.
size_t sort_ws(size_t n, int kb, bool own_scratch) {
  size_t b = up(sizeof(ts::SortMeta));
  if (n > ts::kSmallMax) b += up(ts::status_bytes(n, kb));
  if (own_scratch && n > ts::kSmallMax) b += up(n * size_t(kb));
  return b;
}
size_t by_key_ws(size_t n, int kb, size_t gather_bytes) {
  size_t b = up(sizeof(ts::SortMeta));
  if (n > ts::kSmallMax) b += up(ts::status_bytes(n, kb)) + 2 * up(n * size_t(kb)) + 2 * up(n * 4);
  b += up(gather_bytes);
  return b;
}

void clock_reset(ts_ctx* c) { c->clk.count = 0; }

// ---- Sorting without allocation (ts_sort_with_buffer_static) ----
// The bookkeeping and look-back status of keys-only radix sorts of up to
// kStaticMaxKeys keys live in `scratch`, a file-scope device array of which
// each GPU holds its own copy for the life of the process.  Host code uses
// the array only through the per-device address hipGetSymbolAddress gives
// (static_scratch_setup, at context creation); the array's name on the host
// is not a device address.
constexpr size_t kStaticMaxKeys = size_t(1) << 28;
constexpr size_t static_tiles(size_t n, size_t tile) { return (n + tile - 1) / tile + size_t(ts::kSegments); }
constexpr size_t cmax(size_t a, size_t b) { return a > b ? a : b; }
// Keys-only shapes: the large tiles up to kStaticMaxKeys keys, the small
// tiles up to TS_SWEEP_SMALL_N (keys of 1, 2 and 4 bytes share the 4-byte
// shapes).
constexpr size_t kStaticTiles =
    cmax(cmax(static_tiles(kStaticMaxKeys, ts::SweepShape<4>::tile), static_tiles(kStaticMaxKeys, ts::SweepShape<8>::tile)),
         cmax(static_tiles(TS_SWEEP_SMALL_N, ts::SweepShape<4, false, false, true>::tile),
              static_tiles(TS_SWEEP_SMALL_N, ts::SweepShape<8, false, false, true>::tile)));
static_assert(kStaticMaxKeys < (size_t(1) << 30), "32-bit look-back words");
constexpr size_t kStaticMeta = (sizeof(ts::SortMeta) + 255) / 256 * 256;
constexpr size_t kStaticStatus = 2 * kStaticTiles * 256 * 4;
__device__ __align__(256) static unsigned char scratch[kStaticMeta + kStaticStatus];

std::mutex g_static_mu;
std::vector<void*> g_static_ptr;  // per device: the address of `scratch` there

// With `device` current: looks up the device's copy of `scratch` once.
hipError_t static_scratch_setup(int device) {
  std::lock_guard<std::mutex> lock(g_static_mu);
  if (size_t(device) < g_static_ptr.size() && g_static_ptr[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (size_t(device) >= g_static_ptr.size()) g_static_ptr.resize(size_t(device) + 1, nullptr);
  g_static_ptr[device] = p;
  return hipSuccess;
}

char* static_scratch(int device) {
  std::lock_guard<std::mutex> lock(g_static_mu);
  return size_t(device) < g_static_ptr.size() ? static_cast<char*>(g_static_ptr[device]) : nullptr;
}

__global__ void k_iota(uint32_t* out, size_t n) {
  const size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x, gs = size_t(gridDim.x) * blockDim.x;
  for (size_t i = g; i < n; i += gs) out[i] = uint32_t(i);
}

// Index of a key width in ts_ctx::seg_variant.
int width_index(int kb) { return kb == 1 ? 0 : kb == 2 ? 1 : kb == 4 ? 2 : 3; }

ts_status sort_impl(ts_ctx* c, ts_dtype dt, void* data, size_t n, void* buf, hipStream_t s) {
  const int kb = ts::dtype_bytes(dt), mode = ts::dtype_mode(dt);
  clock_reset(c);
  if (n <= 1) return TS_OK;
  if (n <= ts::kSmallMax) {
    hipError_t e = ts::small_sort(kb, mode, data, n, nullptr, 0, 0, nullptr, s, &c->clk);
    return e == hipSuccess ? TS_OK : hip_fail(c, e, "small_sort");
  }
  if (n <= ts::seg_lds_max_len(kb, false)) {
    // Still one workgroup's LDS: the keys-only segment kernel for this length
    // on a one-segment list (two launches instead of the radix sort's nine),
    // the list in the caller's buffer when it is 8-byte aligned.
    void* area = buf;
    if (!area || reinterpret_cast<uintptr_t>(area) % 8 != 0) {
      ts_status st = ensure(c, ts::kOneSegmentBytes);
      if (st != TS_OK) return st;
      area = c->ws;
    }
    hipError_t e = ts::lds_sort(kb, mode, data, n, area, c->seg_variant[width_index(kb)][0], c->atomic_rank,
                                c->dflags, s, &c->clk);
    return e == hipSuccess ? TS_OK : hip_fail(c, e, "lds_sort");
  }
  ts_status st = ensure(c, sort_ws(n, kb, buf == nullptr));
  if (st != TS_OK) return st;
  char* p = c->ws;
  const size_t sb = ts::status_bytes(n, kb);
  ts::RadixJob j = radix_job(c, kb, mode, n, p, p + up(sizeof(ts::SortMeta)), sb);
  p += up(sizeof(ts::SortMeta)) + up(sb);
  j.keys_a = data;
  j.keys_b = buf ? buf : p;
  j.divert = kb == 8 && n >= c->divert_min;
  hipError_t e = ts::radix_sort(j, s, &c->clk);
  return e == hipSuccess ? TS_OK : hip_fail(c, e, "radix_sort");
}

// A segment too long for one workgroup: the device radix sort by key with
// its order as values, then keys and values gathered into that order.
ts_status seg_device_sort(ts_ctx* c, int kb, int mode, unsigned char* keys, uint32_t* vals, size_t m,
                          hipStream_t s) {
  if (!vals && m > ts::kSmallMax) {  // keys only: the plain device sort in place
    ts_status st = ensure(c, sort_ws(m, kb, true));
    if (st != TS_OK) return st;
    const size_t sb = ts::status_bytes(m, kb);
    ts::RadixJob j = radix_job(c, kb, mode, m, c->ws, c->ws + up(sizeof(ts::SortMeta)), sb);
    j.keys_a = keys;
    j.keys_b = c->ws + up(sizeof(ts::SortMeta)) + up(sb);
    hipError_t e = ts::radix_sort(j, s, &c->clk);
    return e == hipSuccess ? TS_OK : hip_fail(c, e, "radix_sort");
  }
  const size_t gbytes = std::max(m * size_t(kb), vals ? m * 4 : size_t(0));
  ts_status st = ensure(c, by_key_ws(m, kb, gbytes) + up(m * 4));
  if (st != TS_OK) return st;
  char* p = c->ws;
  hipError_t e;
  uint32_t* order;
  char* gbuf;
  if (m <= ts::kSmallMax) {
    order = reinterpret_cast<uint32_t*>(p);
    gbuf = p + up(m * 4);
    e = ts::small_sort(kb, mode, nullptr, m, keys, size_t(kb), 0, order, s, &c->clk);
    if (e != hipSuccess) return hip_fail(c, e, "small_sort");
  } else {
    const size_t sb = ts::status_bytes(m, kb);
    ts::RadixJob j = radix_job(c, kb, mode, m, p, p + up(sizeof(ts::SortMeta)), sb);
    j.records = keys;
    j.stride = size_t(kb);
    j.key_offset = 0;
    p += up(sizeof(ts::SortMeta)) + up(sb);
    j.keys_a = p;
    p += up(m * size_t(kb));
    j.keys_b = p;
    p += up(m * size_t(kb));
    j.vals_a = reinterpret_cast<uint32_t*>(p);
    p += up(m * 4);
    j.vals_b = reinterpret_cast<uint32_t*>(p);
    p += up(m * 4);
    order = reinterpret_cast<uint32_t*>(p);
    p += up(m * 4);
    gbuf = p;
    j.final_vals = order;
    e = ts::radix_sort(j, s, &c->clk);
    if (e != hipSuccess) return hip_fail(c, e, "radix_sort");
  }
  e = ts::gather(order, keys, gbuf, m, size_t(kb), s, &c->clk);
  if (e == hipSuccess) e = hipMemcpyAsync(keys, gbuf, m * size_t(kb), hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess && vals) {
    e = ts::gather(order, vals, gbuf, m, 4, s, &c->clk);
    if (e == hipSuccess) e = hipMemcpyAsync(vals, gbuf, m * 4, hipMemcpyDeviceToDevice, s);
  }
  return e == hipSuccess ? TS_OK : hip_fail(c, e, "segment gather");
}

void set_variants(const ts_ctx* c, int kb, bool vals, ts::SegJob& j) {
  for (int k = 0; k < ts::kSegClasses; ++k) j.variant[k] = c->seg_variant[width_index(kb)][vals ? 1 : 0][k];
}

// Device arrays for sort_long_keys: room for `f` long segments per round.
size_t long_keys_bytes(size_t f) {
  return up(f * 20) + up(f * 16) + up(f * 8) + up(f * 32) + up((ts::kSegClasses + 2) * 4) +
         up(size_t(ts::kSegClasses) * 2 * f * 4);
}
// Pinned host arrays for sort_long_keys: a round's segments (starts,
// lengths, chunk starts), split counts, and the parts' starts and ends.
size_t long_keys_pin_bytes(size_t f) { return up(f * 20) + up(f * 16) + up(f * 32); }
// At most this many segments longer than `cap` fit in n keys.
size_t long_keys_max(size_t n, size_t cap) { return n / (cap + 1) + 1; }

// Reading a seg_long_bounds block (room for fmax pairs) into the pinned
// area, whose start then holds the header: start_long_bounds queues a copy of
// the header and the first kReadbackPairs pairs; once it is complete,
// finish_long_bounds copies the rest (if any) on stream s2, waits for it and
// lists the long segments as (start, length).
constexpr size_t kReadbackPairs = 8192;
hipError_t start_long_bounds(ts_ctx* c, const unsigned long long* back, size_t fmax, hipStream_t s) {
  return hipMemcpyAsync(c->pin, back, ts::kSegLongHeader + std::min(fmax, kReadbackPairs) * 16,
                        hipMemcpyDeviceToHost, s);
}
hipError_t finish_long_bounds(ts_ctx* c, const unsigned long long* back, size_t fmax, hipStream_t s2,
                              std::vector<std::pair<uint64_t, uint64_t>>* fat) {
  const size_t first = std::min(fmax, kReadbackPairs);
  const size_t nbig = std::min<size_t>(reinterpret_cast<const uint32_t*>(c->pin)[ts::kSegDevice], fmax);
  auto* pairs = reinterpret_cast<const unsigned long long*>(c->pin + ts::kSegLongHeader);
  if (nbig > first) {
    hipError_t e = hipMemcpyAsync(c->pin + ts::kSegLongHeader + first * 16, back + ts::kSegLongHeader / 8 + 2 * first,
                                  (nbig - first) * 16, hipMemcpyDeviceToHost, s2);
    if (e == hipSuccess) e = hipStreamSynchronize(s2);
    if (e != hipSuccess) return e;
  }
  fat->clear();
  fat->reserve(nbig);
  for (size_t i = 0; i < nbig; ++i) fat->emplace_back(pairs[2 * i], pairs[2 * i + 1]);
  return hipSuccess;
}

// Device area for the radix sorts of sort_long_keys: meta and look-back
// status for up to n keys.
size_t long_radix_bytes(size_t n, int kb) { return up(sizeof(ts::SortMeta)) + up(ts::status_bytes(n, kb)); }

// Keys-only segments too long for one workgroup (equal keys are equal bits,
// so no order among them is observable): rounds of three-way splits around a
// pivot; the parts that fit one workgroup go to the segmented LDS classes,
// longer parts are split again (all of a round's segments in the same
// launches).  What is left after the context's split_rounds rounds takes the
// device radix sort, with the segment's own stretch of `tmp` as its second
// buffer.  `tmp` holds n keys; `area` has long_keys_bytes(f) bytes,
// `radix_ws` long_radix_bytes(n, kb) bytes and `pin` long_keys_pin_bytes(f)
// bytes of pinned host memory, with f >= long_keys_max(n, cap).  Nothing else
// of the context's workspace is touched, so the caller may still be using
// it.  Returns the rounds used in *rounds.
ts_status sort_long_keys(ts_ctx* c, int kb, int mode, void* data, void* tmp, size_t n,
                         std::vector<std::pair<uint64_t, uint64_t>> segs, char* area, char* radix_ws, char* pin,
                         size_t f, hipStream_t s, uint32_t* rounds) {
  const uint32_t kRounds = uint32_t(c->split_rounds);
  const uint32_t cap = ts::seg_lds_max_len(kb, false);
  auto radix_part = [&](uint64_t lo, uint64_t len) {
    // radix_ws was sized for n keys, which covers every part of at most n.
    ts::RadixJob j = radix_job(c, kb, mode, len, radix_ws, radix_ws + up(sizeof(ts::SortMeta)), ts::status_bytes(n, kb));
    j.keys_a = static_cast<unsigned char*>(data) + lo * size_t(kb);
    j.keys_b = static_cast<unsigned char*>(tmp) + lo * size_t(kb);
    return ts::radix_sort(j, s, &c->clk);
  };
  size_t o = 0;
  auto region = [&](size_t bytes) {
    char* r = area + o;
    o += up(bytes);
    return r;
  };
  // Per round, packed so that each direction takes one copy: the segments
  // (starts[ns] | lengths[ns] | chunk starts[ns] as u32), the split counts
  // (below[ns] | above[ns]) and the parts (starts[m] | ends[m], m <= 2 ns).
  char* d_seg = region(f * 20);
  auto* d_ba = reinterpret_cast<unsigned long long*>(region(f * 16));
  auto* d_pivot = reinterpret_cast<unsigned long long*>(region(f * 8));
  auto* d_se = reinterpret_cast<unsigned long long*>(region(f * 32));
  auto* d_counts = reinterpret_cast<uint32_t*>(region((ts::kSegClasses + 2) * 4));
  auto* d_lists = reinterpret_cast<uint32_t*>(region(size_t(ts::kSegClasses) * 2 * f * 4));
  // Host side: each array is rewritten only after a stream synchronisation
  // that follows every copy reading it.
  size_t po = 0;
  auto pregion = [&](size_t bytes) {
    char* r = pin + po;
    po += up(bytes);
    return r;
  };
  char* h_seg = pregion(f * 20);
  auto* h_ba = reinterpret_cast<unsigned long long*>(pregion(f * 16));
  auto* h_se = reinterpret_cast<unsigned long long*>(pregion(f * 32));
  const uint32_t chunk = ts::split_chunk();
  uint32_t class_cap[ts::kSegDevice];
  for (int k = 0; k < ts::kSegDevice; ++k) class_cap[k] = ts::seg_class_cap(kb, false, k);
  uint32_t r = 0;
  for (; r < kRounds && !segs.empty(); ++r) {
    const size_t ns = segs.size();
    if (ns > f) return fail(c, TS_ERR_INVALID_ARG, "internal: too many long segments");
    auto* h_lo = reinterpret_cast<unsigned long long*>(h_seg);
    auto* h_chunks = reinterpret_cast<uint32_t*>(h_seg + ns * 16);
    uint32_t chunks = 0;
    for (size_t i = 0; i < ns; ++i) {
      h_lo[i] = segs[i].first;
      h_lo[ns + i] = segs[i].second;
      h_chunks[i] = chunks;
      chunks += uint32_t((segs[i].second + chunk - 1) / chunk);
    }
    hipError_t e = hipMemcpyAsync(d_seg, h_seg, ns * 20, hipMemcpyHostToDevice, s);
    ts::SplitJob sj;
    sj.key_bytes = kb;
    sj.mode = mode;
    sj.data = data;
    sj.tmp = tmp;
    sj.nseg = uint32_t(ns);
    sj.chunks = chunks;
    sj.lo = reinterpret_cast<const unsigned long long*>(d_seg);
    sj.len = reinterpret_cast<const unsigned long long*>(d_seg) + ns;
    sj.chunk_start = reinterpret_cast<const uint32_t*>(d_seg + ns * 16);
    sj.below = d_ba;
    sj.above = d_ba + ns;
    sj.pivot = d_pivot;
    if (e == hipSuccess) e = ts::split_segments(sj, s);
    if (e == hipSuccess) e = hipMemcpyAsync(h_ba, d_ba, ns * 16, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(c, e, "split");
    std::vector<std::pair<uint64_t, uint64_t>> next, fit;
    uint32_t mask = 0;  // size classes the fitting parts fall in
    for (size_t i = 0; i < ns; ++i) {
      const uint64_t lo = segs[i].first, len = segs[i].second, nb = h_ba[i], na = h_ba[ns + i];
      const std::pair<uint64_t, uint64_t> parts[2] = {{lo, nb}, {lo + len - na, na}};
      for (const auto& pt : parts) {
        if (pt.second < 2) continue;
        if (pt.second <= cap) {
          fit.push_back(pt);
          int k = 0;
          while (k + 1 < ts::kSegDevice && (class_cap[k] == 0 || pt.second > class_cap[k])) ++k;
          mask |= 1u << k;
        } else {
          next.push_back(pt);
        }
      }
    }
    if (!fit.empty()) {
      const size_t m = fit.size();
      for (size_t i = 0; i < m; ++i) {
        h_se[i] = fit[i].first;
        h_se[m + i] = fit[i].first + fit[i].second;
      }
      e = hipMemcpyAsync(d_se, h_se, m * 16, hipMemcpyHostToDevice, s);
      ts::SegJob gj;
      gj.key_bytes = kb;
      gj.mode = mode;
      gj.keys = data;
      gj.n = n;
      gj.offsets = d_se;
      gj.ends = d_se + m;
      gj.n_segments = m;
      gj.class_mask = mask;
      gj.counts = d_counts;
      gj.lists = d_lists;
      gj.num_cus = c->num_cus;
      gj.atomic_rank = c->atomic_rank;
      gj.error = c->dflags;
      set_variants(c, kb, false, gj);
      if (e == hipSuccess) e = ts::segmented_sort(gj, s, &c->clk);
      if (e != hipSuccess) return hip_fail(c, e, "segmented_sort");
      // The next round reuses d_se only after this sort (same stream).
    }
    segs.swap(next);
  }
  if (rounds) *rounds = r;
  hipError_t e = hipSuccess;
  for (const auto& sg : segs)
    if (e == hipSuccess) e = radix_part(sg.first, sg.second);
  // The last parts' copies read pinned memory that the next call may rewrite.
  if (e != hipSuccess) return hip_fail(c, e, "split rounds");
  return check_flags(c, s);
}

}  // namespace

extern "C" {

const char* ts_status_string(ts_status s) {
  switch (s) {
    case TS_OK: return "ok";
    case TS_ERR_BUFFER_TOO_SMALL: return "buffer too small";
    case TS_ERR_LENGTH_MISMATCH: return "length mismatch";
    case TS_ERR_TOO_LONG: return "too long for a u32 order";
    case TS_ERR_INVALID_ARG: return "invalid argument";
    case TS_ERR_HIP: return "HIP runtime error";
    case TS_ERR_VERIFY_FAILED: return "output verification failed";
    case TS_ERR_NO_DEVICE: return "no such device";
  }
  return "unknown status";
}

namespace {

// Runs rank_order_check once per device (first context on it) and keeps the
// answer: 1 when every atomic rank matched its ballot rank, 0 otherwise.
hipError_t atomic_rank_ok(int device, bool* ok) {
  static std::mutex mu;
  static std::vector<int8_t> known;
  std::lock_guard<std::mutex> lock(mu);
  if (size_t(device) >= known.size()) known.resize(size_t(device) + 1, -1);
  if (known[device] < 0) {
    DeviceScope ds(device);
    uint32_t* bad = nullptr;
    hipError_t e = hipMalloc(&bad, sizeof(uint32_t));
    if (e != hipSuccess) return e;
    uint32_t h = 1;
    e = hipMemset(bad, 0, sizeof(uint32_t));
    if (e == hipSuccess) e = ts::rank_order_check(bad, nullptr);
    if (e == hipSuccess) e = hipMemcpy(&h, bad, sizeof(uint32_t), hipMemcpyDeviceToHost);
    (void)hipFree(bad);
    if (e != hipSuccess) return e;
    known[device] = h == 0 ? 1 : 0;
  }
  *ok = known[device] == 1;
  return hipSuccess;
}

}  // namespace

ts_status ts_ctx_create(ts_ctx** out, int device) {
  if (!out) return TS_ERR_INVALID_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return TS_ERR_NO_DEVICE;
  bool ok = false;
  if (atomic_rank_ok(device, &ok) != hipSuccess) return TS_ERR_HIP;
  ts_ctx* c = new ts_ctx();
  c->device = device;
  c->atomic_rank_ok = ok;
  c->atomic_rank = ok;
  {
    DeviceScope ds(device);
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&c->dflags), 256);
    if (e == hipSuccess) e = hipMemset(c->dflags, 0, 256);
    if (e == hipSuccess) e = static_scratch_setup(device);
    if (e != hipSuccess) {
      if (c->dflags) (void)hipFree(c->dflags);
      delete c;
      return TS_ERR_HIP;
    }
  }
  // Keys-only segments of the largest class: the MSD shape (4-byte keys)
  // and its pipelined form (8-byte keys: the learned sort's buckets, 0.5 ms
  // faster over 10^9 keys than the plain form).
  c->seg_variant[2][0][8] = 2;
  c->seg_variant[3][0][8] = 1;
  // 4-byte keys with values, segments of 65-16384 keys: one MSD pass with
  // ranks inside the bins (the shapes the autotuner picks on MI355X for the
  // log-uniform batch of 10^5 segments; 2^BB bins = the class's capacity).
  c->seg_variant[2][1][3] = 1;
  c->seg_variant[2][1][4] = 3;
  c->seg_variant[2][1][5] = 3;
  c->seg_variant[2][1][6] = 2;
  c->seg_variant[2][1][7] = 3;
  c->seg_variant[2][1][8] = 2;
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
    c->num_cus = cus;
  *out = c;
  return TS_OK;
}

ts_status ts_ctx_set_ranking(ts_ctx* c, int method) {
  if (!c || (method != TS_RANK_BALLOT && method != TS_RANK_ATOMIC)) return TS_ERR_INVALID_ARG;
  if (method == TS_RANK_ATOMIC && !c->atomic_rank_ok)
    return fail(c, TS_ERR_INVALID_ARG, "atomic ranking: this device's LDS atomics failed the lane-order check");
  c->atomic_rank = method == TS_RANK_ATOMIC;
  return TS_OK;
}

int ts_ctx_ranking(const ts_ctx* c) { return c && c->atomic_rank ? TS_RANK_ATOMIC : TS_RANK_BALLOT; }

ts_status ts_ctx_set_split_rounds(ts_ctx* c, int rounds) {
  if (!c) return TS_ERR_INVALID_ARG;
  if (rounds < 0 || rounds > 64) return fail(c, TS_ERR_INVALID_ARG, "split rounds: 0 to 64");
  c->split_rounds = rounds;
  return TS_OK;
}

int ts_ctx_split_rounds(const ts_ctx* c) { return c ? c->split_rounds : -1; }

ts_status ts_ctx_set_divert_min(ts_ctx* c, size_t n) {
  if (!c) return TS_ERR_INVALID_ARG;
  c->divert_min = n;
  return TS_OK;
}

size_t ts_ctx_divert_min(const ts_ctx* c) { return c ? c->divert_min : 0; }

ts_status ts_ctx_check(ts_ctx* c, void* stream) {
  if (!c) return TS_ERR_INVALID_ARG;
  DeviceScope ds(c->device);
  return check_flags(c, static_cast<hipStream_t>(stream));
}

ts_status ts_ctx_last_sort_info(ts_ctx* c, ts_sort_info* info, void* stream) {
  if (!c || !info) return TS_ERR_INVALID_ARG;
  std::memset(info, 0, sizeof *info);
  const ts::SortMeta* m = c->last_meta;
  if (!m) return TS_OK;
  DeviceScope ds(c->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  uint32_t w[8];
  hipError_t e = hipMemcpyAsync(&w[0], &m->all_live, 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(&w[1], &m->n_live, 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(&w[2], &m->diverted, 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(&w[3], &m->finish_shift, 8, hipMemcpyDeviceToHost, s);  // shift, fallback
  if (e == hipSuccess) e = hipMemcpyAsync(&w[5], &m->reverse, 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(c, e, "sort info readback");
  info->live_digits = w[0];
  info->executed_passes = w[1];
  info->skipped_digits = w[2];
  info->finish_shift = w[3];
  info->fallback = w[4];
  info->reversed = w[5];
  return TS_OK;
}

size_t ts_plan_status_bytes(ts_dtype dt, size_t n, int kind) {
  if (!valid_dtype(dt)) return 0;
  const int kb = ts::dtype_bytes(dt);
  switch (kind) {
    case -1: return ts::status_bytes(n, kb);
    case ts::kSweepKeys: return ts::status_bytes_used(n, kb, false, false);
    case ts::kSweepVals: return ts::status_bytes_used(n, kb, false, true);
    case ts::kSweepLearned: return kb >= 4 ? ts::status_bytes_used(n, kb, true, false) : 0;
    case ts::kSweepRec8Vals: return ts::status_bytes_used(n, kb, false, true, 2);
    case ts::kSweepRec16Vals: return ts::status_bytes_used(n, kb, false, true, 4);
    case ts::kSweepRec8: return ts::status_bytes_used(n, kb, false, false, 2);
    case ts::kSweepRec16: return ts::status_bytes_used(n, kb, false, false, 4);
    default: return 0;
  }
}

void ts_ctx_destroy(ts_ctx* c) {
  if (!c) return;
  DeviceScope ds(c->device);
  if (c->ws) hipFree(c->ws);
  if (c->dflags) hipFree(c->dflags);
  if (c->hdev) hipFree(c->hdev);
  if (c->pin) hipHostFree(c->pin);
  if (c->aux_done) hipEventDestroy(c->aux_done);
  if (c->back_ready) hipEventDestroy(c->back_ready);
  if (c->aux) hipStreamDestroy(c->aux);
  if (c->clk.made)
    for (int i = 0; i < 2 * ts::KernelClock::kMax; ++i) hipEventDestroy(c->clk.ev[i]);
  delete c;
}

ts_status ts_ctx_reserve(ts_ctx* c, ts_dtype dt, size_t n, size_t record_bytes) {
  if (!c || !valid_dtype(dt)) return TS_ERR_INVALID_ARG;
  DeviceScope ds(c->device);
  const int kb = ts::dtype_bytes(dt);
  size_t need = std::max(sort_ws(n, kb, true), by_key_ws(n, kb, n * record_bytes));
  return ensure(c, need);
}

size_t ts_ctx_workspace_bytes(const ts_ctx* c) { return c ? c->ws_bytes : 0; }

const char* ts_ctx_last_error(const ts_ctx* c) { return c ? c->err.c_str() : "null context"; }

ts_status ts_sort(ts_ctx* c, ts_dtype dt, void* data, size_t n, void* stream) {
  if (!c) return TS_ERR_INVALID_ARG;
  if (!valid_dtype(dt)) return fail(c, TS_ERR_INVALID_ARG, "unknown dtype");
  if (n > 1 && !data) return fail(c, TS_ERR_INVALID_ARG, "null data");
  DeviceScope ds(c->device);
  return sort_impl(c, dt, data, n, nullptr, static_cast<hipStream_t>(stream));
}

ts_status ts_sort_with_buffer(ts_ctx* c, ts_dtype dt, void* data, size_t n, void* buf, size_t buf_len,
                              void* stream) {
  if (!c) return TS_ERR_INVALID_ARG;
  if (!valid_dtype(dt)) return fail(c, TS_ERR_INVALID_ARG, "unknown dtype");
  if (buf_len < n)  // dispatch.rs:78-83
    return fail(c, TS_ERR_BUFFER_TOO_SMALL,
                "buffer too small: need " + std::to_string(n) + ", got " + std::to_string(buf_len));
  if (n > 1 && (!data || !buf)) return fail(c, TS_ERR_INVALID_ARG, "null data or buffer");
  DeviceScope ds(c->device);
  return sort_impl(c, dt, data, n, buf, static_cast<hipStream_t>(stream));
}

ts_status ts_sort_with_buffer_static(ts_ctx* c, ts_dtype dt, void* data, size_t n, void* buf, size_t buf_len,
                                     void* stream) {
  if (!c) return TS_ERR_INVALID_ARG;
  if (!valid_dtype(dt)) return fail(c, TS_ERR_INVALID_ARG, "unknown dtype");
  if (buf_len < n)  // dispatch.rs:78-83
    return fail(c, TS_ERR_BUFFER_TOO_SMALL,
                "buffer too small: need " + std::to_string(n) + ", got " + std::to_string(buf_len));
  if (n > kStaticMaxKeys)
    return fail(c, TS_ERR_INVALID_ARG,
                "ts_sort_with_buffer_static sorts at most 2^28 keys, got " + std::to_string(n));
  if (n > 1 && (!data || !buf)) return fail(c, TS_ERR_INVALID_ARG, "null data or buffer");
  DeviceScope ds(c->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int kb = ts::dtype_bytes(dt), mode = ts::dtype_mode(dt);
  clock_reset(c);
  if (n <= 1) return TS_OK;
  if (n <= ts::kSmallMax) {  // one workgroup in LDS: no scratch at all
    hipError_t e = ts::small_sort(kb, mode, data, n, nullptr, 0, 0, nullptr, s, &c->clk);
    return e == hipSuccess ? TS_OK : hip_fail(c, e, "small_sort");
  }
  char* p = static_scratch(c->device);
  if (!p) return fail(c, TS_ERR_HIP, "no static scratch address for this device");
  if (ts::status_bytes_used(n, kb, false, false) > kStaticStatus)
    return fail(c, TS_ERR_INVALID_ARG, "internal: static look-back status too small");
  ts::RadixJob j = radix_job(c, kb, mode, n, p, p + kStaticMeta, kStaticStatus);
  j.keys_a = data;
  j.keys_b = buf;
  j.divert = kb == 8 && n >= c->divert_min;
  hipError_t e = ts::radix_sort(j, s, &c->clk);
  return e == hipSuccess ? TS_OK : hip_fail(c, e, "radix_sort");
}

ts_status ts_sort_host(ts_ctx* c, ts_dtype dt, void* host_data, size_t n, void* stream) {
  if (!c) return TS_ERR_INVALID_ARG;
  if (!valid_dtype(dt)) return fail(c, TS_ERR_INVALID_ARG, "unknown dtype");
  if (n <= 1) return TS_OK;
  if (!host_data) return fail(c, TS_ERR_INVALID_ARG, "null data");
  DeviceScope ds(c->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t bytes = n * size_t(ts::dtype_bytes(dt));
  if (bytes > c->hdev_bytes) {
    if (c->hdev) {
      hipError_t e = hipFree(c->hdev);
      c->hdev = nullptr;
      c->hdev_bytes = 0;
      if (e != hipSuccess) return hip_fail(c, e, "hipFree");
    }
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&c->hdev), bytes);
    if (e != hipSuccess) {
      c->hdev = nullptr;
      return hip_fail(c, e, "hipMalloc(host-sort buffer)");
    }
    c->hdev_bytes = bytes;
  }
  // Copies in chunks so that a pinned slice streams at full link rate and a
  // pageable one goes through the runtime's staging piece by piece.
  constexpr size_t kChunk = size_t(64) << 20;
  char* h = static_cast<char*>(host_data);
  for (size_t o = 0; o < bytes; o += kChunk) {
    hipError_t e = hipMemcpyAsync(c->hdev + o, h + o, std::min(kChunk, bytes - o), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_fail(c, e, "hipMemcpyAsync(host to device)");
  }
  ts_status st = sort_impl(c, dt, c->hdev, n, nullptr, s);
  if (st != TS_OK) return st;
  for (size_t o = 0; o < bytes; o += kChunk) {
    hipError_t e = hipMemcpyAsync(h + o, c->hdev + o, std::min(kChunk, bytes - o), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return hip_fail(c, e, "hipMemcpyAsync(device to host)");
  }
  return check_flags(c, s);
}

namespace {
// By-key sorts of 8- and 16-byte records (aligned to their size) carry the
// records through the radix passes instead of gathering them afterwards: a
// random gather reads a whole 128-byte line per record, the passes read and
// write them in runs.
bool carries(const void* records, size_t n, size_t stride) {
  return n > ts::kSmallMax && (stride == 8 || stride == 16) && reinterpret_cast<uintptr_t>(records) % stride == 0;
}
size_t carry_ws(size_t n, int kb, size_t stride, bool order) {
  return up(sizeof(ts::SortMeta)) + up(ts::status_bytes_used(n, kb, false, order, int(stride / 4))) +
         up(n * stride) + (order ? 2 * up(n * 4) : 0);
}
// The carried sort with its workspace at p (carry_ws bytes): records end in
// stable key order in place, and the order is written when `order` is set.
ts_status sort_carried(ts_ctx* c, int kb, int mode, void* records, size_t n, size_t stride, size_t key_offset,
                       uint32_t* order, char* p, hipStream_t s) {
  const int rw = int(stride / 4);
  const size_t sb = ts::status_bytes_used(n, kb, false, order != nullptr, rw);
  ts::RadixJob j = radix_job(c, kb, mode, n, p, p + up(sizeof(ts::SortMeta)), sb);
  p += up(sizeof(ts::SortMeta)) + up(sb);
  j.records = static_cast<const unsigned char*>(records);
  j.stride = stride;
  j.key_offset = key_offset;
  j.rec_words = rw;
  j.recs_b = p;
  p += up(n * stride);
  if (order) {
    j.vals_a = reinterpret_cast<uint32_t*>(p);
    p += up(n * 4);
    j.vals_b = reinterpret_cast<uint32_t*>(p);
    j.final_vals = order;
  }
  hipError_t e = ts::radix_sort(j, s, &c->clk);
  return e == hipSuccess ? TS_OK : hip_fail(c, e, "radix_sort (records carried)");
}

// A stable sort of n records by keys read at keys + i * key_stride +
// key_offset (inside the records or beside them), with no order: the by-key
// sort of slices longer than a u32 order can index (by_key.rs:190-194).
// Bare keys sort by themselves (equal keys are equal bits); 8- and 16-byte
// records carrying their keys are carried; anything else sorts 16-byte
// (key, u64 position) records and gathers by their positions.
ts_status sort_unordered(ts_ctx* c, ts_dtype kt, const void* keys, size_t key_stride, size_t key_offset, void* records,
                         size_t n, size_t record_bytes, hipStream_t s) {
  const int kb = ts::dtype_bytes(kt), mode = ts::dtype_mode(kt);
  const bool inside = keys == records && key_stride == record_bytes;
  if (inside && record_bytes == size_t(kb)) return sort_impl(c, kt, records, n, nullptr, s);
  if (n <= ts::kSmallMax) {  // one workgroup's (key, position) sort, then the gather
    ts_status st = ensure(c, up(n * 4) + up(n * record_bytes));
    if (st != TS_OK) return st;
    uint32_t* order = reinterpret_cast<uint32_t*>(c->ws);
    char* gbuf = c->ws + up(n * 4);
    hipError_t e = ts::small_sort(kb, mode, nullptr, n, static_cast<const unsigned char*>(keys), key_stride,
                                  key_offset, order, s, &c->clk);
    if (e == hipSuccess) e = ts::gather(order, records, gbuf, n, record_bytes, s, &c->clk);
    if (e == hipSuccess) e = hipMemcpyAsync(records, gbuf, n * record_bytes, hipMemcpyDeviceToDevice, s);
    return e == hipSuccess ? TS_OK : hip_fail(c, e, "small by-key sort");
  }
  if (inside && carries(records, n, record_bytes)) {
    ts_status st = ensure(c, carry_ws(n, kb, record_bytes, false));
    if (st != TS_OK) return st;
    return sort_carried(c, kb, mode, records, n, record_bytes, key_offset, nullptr, c->ws, s);
  }
  const size_t sort_bytes = carry_ws(n, kb, 16, false);
  ts_status st = ensure(c, sort_bytes + up(n * 16) + up(n * record_bytes));
  if (st != TS_OK) return st;
  char* packed = c->ws + sort_bytes;
  char* gbuf = packed + up(n * 16);
  hipError_t e = ts::pack_key_pos(kb, keys, key_stride, key_offset, packed, n, s);
  if (e != hipSuccess) return hip_fail(c, e, "pack_key_pos");
  st = sort_carried(c, kb, mode, packed, n, 16, 0, nullptr, c->ws, s);
  if (st != TS_OK) return st;
  e = ts::gather_packed(packed, records, gbuf, n, record_bytes, s, &c->clk);
  if (e != hipSuccess) return hip_fail(c, e, "gather");
  e = hipMemcpyAsync(records, gbuf, n * record_bytes, hipMemcpyDeviceToDevice, s);
  return e == hipSuccess ? TS_OK : hip_fail(c, e, "hipMemcpyAsync");
}

}  // namespace

ts_status ts_sort_by_key(ts_ctx* c, ts_dtype kt, void* records, size_t n, size_t stride, size_t key_offset,
                         uint32_t* order, int gather_records, void* stream) {
  if (!c) return TS_ERR_INVALID_ARG;
  if (!valid_dtype(kt)) return fail(c, TS_ERR_INVALID_ARG, "unknown key dtype");
  const int kb = ts::dtype_bytes(kt), mode = ts::dtype_mode(kt);
  if (stride < size_t(kb) || key_offset + size_t(kb) > stride)
    return fail(c, TS_ERR_INVALID_ARG, "key does not fit inside the record stride");
  if ((reinterpret_cast<uintptr_t>(records) + key_offset) % size_t(kb) != 0 || stride % size_t(kb) != 0)
    return fail(c, TS_ERR_INVALID_ARG, "record keys must be aligned to the key width");
  if (n > size_t(UINT32_MAX)) {  // by_key.rs:190-194: a stable sort, and no order
    if (!gather_records)
      return fail(c, TS_ERR_TOO_LONG, "more than u32::MAX elements: no order, so only a sort that moves the records");
    if (!records) return fail(c, TS_ERR_INVALID_ARG, "null records");
    DeviceScope ds(c->device);
    clock_reset(c);
    return sort_unordered(c, kt, records, stride, key_offset, records, n, stride, static_cast<hipStream_t>(stream));
  }
  if (n > 0 && (!records || !order)) return fail(c, TS_ERR_INVALID_ARG, "null records or order");
  DeviceScope ds(c->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  clock_reset(c);
  if (n == 0) return TS_OK;
  if (gather_records && carries(records, n, stride)) {
    ts_status st = ensure(c, carry_ws(n, kb, stride, true));
    if (st != TS_OK) return st;
    return sort_carried(c, kb, mode, records, n, stride, key_offset, order, c->ws, s);
  }
  const unsigned char* rec = static_cast<const unsigned char*>(records);
  const size_t gbytes = gather_records ? n * stride : 0;
  ts_status st = ensure(c, by_key_ws(n, kb, gbytes));
  if (st != TS_OK) return st;
  char* p = c->ws;
  hipError_t e;
  if (n < 2) {  // by_key.rs:185-189
    hipLaunchKernelGGL(k_iota, dim3(1), dim3(64), 0, s, order, n);
    return TS_OK;
  }
  char* gbuf;
  if (n <= ts::kSmallMax) {
    gbuf = p + up(sizeof(ts::SortMeta));
    e = ts::small_sort(kb, mode, nullptr, n, rec, stride, key_offset, order, s, &c->clk);
    if (e != hipSuccess) return hip_fail(c, e, "small_sort");
  } else {
    const size_t sb = ts::status_bytes(n, kb);
    ts::RadixJob j = radix_job(c, kb, mode, n, p, p + up(sizeof(ts::SortMeta)), sb);
    j.records = rec;
    j.stride = stride;
    j.key_offset = key_offset;
    p += up(sizeof(ts::SortMeta)) + up(sb);
    j.keys_a = p;
    p += up(n * size_t(kb));
    j.keys_b = p;
    p += up(n * size_t(kb));
    j.vals_a = reinterpret_cast<uint32_t*>(p);
    p += up(n * 4);
    j.vals_b = reinterpret_cast<uint32_t*>(p);
    p += up(n * 4);
    j.final_vals = order;
    gbuf = p;
    e = ts::radix_sort(j, s, &c->clk);
    if (e != hipSuccess) return hip_fail(c, e, "radix_sort");
  }
  if (gather_records) {  // move_into_order (by_key.rs:315-334): gather, then copy back
    e = ts::gather(order, records, gbuf, n, stride, s, &c->clk);
    if (e != hipSuccess) return hip_fail(c, e, "gather");
    e = hipMemcpyAsync(records, gbuf, n * stride, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return hip_fail(c, e, "hipMemcpyAsync");
  }
  return TS_OK;
}

ts_status ts_sort_by_key_unordered(ts_ctx* c, ts_dtype kt, const void* keys, size_t key_stride, size_t key_offset,
                                   void* records, size_t n, size_t record_bytes, void* stream) {
  if (!c) return TS_ERR_INVALID_ARG;
  if (!valid_dtype(kt)) return fail(c, TS_ERR_INVALID_ARG, "unknown key dtype");
  const int kb = ts::dtype_bytes(kt);
  if (key_stride < size_t(kb) || key_offset + size_t(kb) > key_stride)
    return fail(c, TS_ERR_INVALID_ARG, "key does not fit inside the key stride");
  if ((reinterpret_cast<uintptr_t>(keys) + key_offset) % size_t(kb) != 0 || key_stride % size_t(kb) != 0)
    return fail(c, TS_ERR_INVALID_ARG, "keys must be aligned to the key width");
  if (record_bytes == 0) return fail(c, TS_ERR_INVALID_ARG, "zero-sized records");
  if (n > 0 && (!keys || !records)) return fail(c, TS_ERR_INVALID_ARG, "null keys or records");
  DeviceScope ds(c->device);
  clock_reset(c);
  if (n < 2) return TS_OK;
  return sort_unordered(c, kt, keys, key_stride, key_offset, records, n, record_bytes,
                        static_cast<hipStream_t>(stream));
}

ts_status ts_apply_order(ts_ctx* c, const uint32_t* order, size_t order_len, void* other, size_t n,
                         size_t elem_bytes, void* stream) {
  if (!c) return TS_ERR_INVALID_ARG;
  if (order_len != n)  // by_key.rs:137-141
    return fail(c, TS_ERR_LENGTH_MISMATCH, "apply_order: the slice's length differs from the last sorted slice's");
  if (elem_bytes == 0) return fail(c, TS_ERR_INVALID_ARG, "zero-sized elements");
  if (n > 0 && (!order || !other)) return fail(c, TS_ERR_INVALID_ARG, "null order or slice");
  DeviceScope ds(c->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  clock_reset(c);
  if (n == 0) return TS_OK;
  ts_status st = ensure(c, up(n * elem_bytes));
  if (st != TS_OK) return st;
  hipError_t e = ts::gather(order, other, c->ws, n, elem_bytes, s, &c->clk);
  if (e != hipSuccess) return hip_fail(c, e, "gather");
  e = hipMemcpyAsync(other, c->ws, n * elem_bytes, hipMemcpyDeviceToDevice, s);
  return e == hipSuccess ? TS_OK : hip_fail(c, e, "hipMemcpyAsync");
}

ts_status ts_segmented_sort_pairs(ts_ctx* c, ts_dtype dt, void* keys, uint32_t* vals, size_t n,
                                  const uint64_t* offsets, size_t n_segments, size_t max_segment_len,
                                  void* stream) {
  if (!c) return TS_ERR_INVALID_ARG;
  if (!valid_dtype(dt)) return fail(c, TS_ERR_INVALID_ARG, "unknown dtype");
  clock_reset(c);
  if (n_segments == 0) return TS_OK;
  if (!offsets) return fail(c, TS_ERR_INVALID_ARG, "null offsets");
  if (n_segments >= size_t(UINT32_MAX)) return fail(c, TS_ERR_TOO_LONG, "more than u32::MAX - 1 segments");
  const int kb = ts::dtype_bytes(dt), mode = ts::dtype_mode(dt);
  if (n > 0 && !keys) return fail(c, TS_ERR_INVALID_ARG, "null keys");
  if (reinterpret_cast<uintptr_t>(keys) % size_t(kb) != 0 || reinterpret_cast<uintptr_t>(vals) % 4 != 0)
    return fail(c, TS_ERR_INVALID_ARG, "keys and values must be aligned to their widths");
  DeviceScope ds(c->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t count_bytes = up((ts::kSegClasses + 2) * sizeof(uint32_t));
  const size_t list_bytes = up(size_t(ts::kSegClasses) * n_segments * sizeof(uint32_t));
  const uint32_t cap = ts::seg_lds_max_len(kb, vals != nullptr);
  const size_t fmax = long_keys_max(n, cap);  // valid offsets are non-decreasing: segments are disjoint
  const size_t back_bytes = ts::kSegLongHeader + fmax * 16;
  ts_status st = ensure(c, count_bytes + list_bytes + up(back_bytes));
  if (st != TS_OK) return st;
  ts::SegJob j;
  j.key_bytes = kb;
  j.mode = mode;
  j.keys = keys;
  j.vals = vals;
  j.n = n;
  j.offsets = reinterpret_cast<const unsigned long long*>(offsets);
  j.n_segments = n_segments;
  j.counts = reinterpret_cast<uint32_t*>(c->ws);
  j.lists = reinterpret_cast<uint32_t*>(c->ws + count_bytes);
  j.num_cus = c->num_cus;
  j.atomic_rank = c->atomic_rank;
  j.error = c->dflags;
  const bool async = max_segment_len != 0 && max_segment_len <= ts::seg_lds_max_len(kb, vals != nullptr);
  // Without a host wait, a violated bound or bad offsets are recorded in the
  // sticky error word for the next ts_ctx_check (or call that waits).
  j.max_len = async ? max_segment_len : 0;
  set_variants(c, kb, vals != nullptr, j);
  hipError_t e = ts::segmented_sort(j, s, &c->clk);
  if (e != hipSuccess) return hip_fail(c, e, "segmented_sort");
  if (async) return TS_OK;

  // Unknown or long segments: read the class counts and the long segments'
  // bounds back (one copy unless they are many).
  st = ensure_pinned(c, up(back_bytes) + (vals ? 0 : long_keys_pin_bytes(fmax)));
  if (st != TS_OK) return st;
  auto* back = reinterpret_cast<unsigned long long*>(c->ws + count_bytes + list_bytes);
  std::vector<std::pair<uint64_t, uint64_t>> fat;
  e = ts::seg_long_bounds(j, back, uint32_t(fmax), nullptr, 0, s);
  if (e == hipSuccess) e = start_long_bounds(c, back, fmax, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e == hipSuccess) e = finish_long_bounds(c, back, fmax, s, &fat);
  if (e != hipSuccess) return hip_fail(c, e, "segment bounds");
  if (reinterpret_cast<const uint32_t*>(c->pin)[ts::kSegClasses]) {
    (void)check_flags(c, s);  // the bit this call set is answered here
    return fail(c, TS_ERR_INVALID_ARG, "segment offsets must be non-decreasing and at most n");
  }
  if (fat.empty()) return check_flags(c, s);
  if (!vals) {
    // Keys only: three-way splits, then the LDS classes (the workspace's
    // class lists are no longer needed, so it may be regrown).
    st = ensure(c, up(n * size_t(kb)) + long_keys_bytes(fmax) + long_radix_bytes(n, kb));
    if (st != TS_OK) return st;
    char* area = c->ws + up(n * size_t(kb));
    return sort_long_keys(c, kb, mode, keys, c->ws, n, std::move(fat), area, area + long_keys_bytes(fmax),
                          c->pin + up(back_bytes), fmax, s, nullptr);
  }
  for (const auto& sg : fat) {
    st = seg_device_sort(c, kb, mode, static_cast<unsigned char*>(keys) + sg.first * size_t(kb),
                         vals + sg.first, size_t(sg.second), s);
    if (st != TS_OK) return st;
  }
  return TS_OK;
}

void ts_learned_cfg_default(ts_learned_cfg* cfg) {
  if (!cfg) return;
  cfg->buckets_log2 = 16;
  cfg->sample_per_mille = 5;
  cfg->train = 1;
  cfg->iters = 40;
  cfg->lr = 0.02f;
  cfg->points = 16384;
  cfg->verify = 1;
}

ts_status ts_learned_sort(ts_ctx* c, ts_dtype dt, void* data, size_t n, const ts_learned_cfg* cfg_in,
                          ts_learned_report* report, void* stream) {
  if (!c) return TS_ERR_INVALID_ARG;
  if (!valid_dtype(dt)) return fail(c, TS_ERR_INVALID_ARG, "unknown dtype");
  const int kb = ts::dtype_bytes(dt), mode = ts::dtype_mode(dt);
  if (kb != 4 && kb != 8) return fail(c, TS_ERR_INVALID_ARG, "the learned sort takes 4- and 8-byte keys");
  ts_learned_cfg cfg;
  ts_learned_cfg_default(&cfg);
  if (cfg_in) cfg = *cfg_in;
  if (cfg.buckets_log2 < 1 || cfg.buckets_log2 > 24 || cfg.sample_per_mille < 1 || cfg.sample_per_mille > 1000 ||
      cfg.points < 1 || !(cfg.lr > 0.f))
    return fail(c, TS_ERR_INVALID_ARG, "invalid learned-sort configuration");
  if (report) std::memset(report, 0, sizeof *report);
  if (n > 0 && !data) return fail(c, TS_ERR_INVALID_ARG, "null data");
  DeviceScope ds(c->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  clock_reset(c);
  if (n <= 1) return TS_OK;

  const uint32_t B = 1u << cfg.buckets_log2;
  size_t S = n / 1000 * cfg.sample_per_mille + (n % 1000) * cfg.sample_per_mille / 1000;
  S = std::min(n, std::max<size_t>(S, std::min<size_t>(n, 1024)));
  const uint32_t T = uint32_t(std::min<size_t>(S, cfg.points));
  // meta | status | scratch | sample x2 | u | t | theta | grad | loss | model | offsets | seg counts | seg lists
  // | long-bucket arrays
  const size_t cap = ts::seg_lds_max_len(kb, false);
  const size_t fmax = long_keys_max(n, cap);
  size_t off[17];
  size_t b = 0;
  auto take = [&](int i, size_t bytes) { off[i] = b; b += up(bytes); };
  // The look-back status serves the sample's sort (8-byte keys) and the
  // partition.
  const size_t sb = std::max(ts::status_bytes(n, kb), ts::status_bytes(S, 8));
  take(0, sizeof(ts::SortMeta));
  take(1, sb);
  take(2, n * size_t(kb));
  take(3, S * 8);
  take(4, S * 8);
  take(5, size_t(T) * 4);
  take(6, size_t(T) * 4);
  take(7, ts::kKnots * 4);
  take(8, ts::kKnots * 4 + (ts::kKnots + 1) * 8);  // gradient, then the knots' first sample indices
  take(9, 4 * 4 + 2 * 8);  // loss at the start / after fitting / of uniform knots, then the largest bucket
  take(10, sizeof(ts::LearnedModel));
  take(11, (size_t(B) + 1) * 8);
  take(12, (ts::kSegClasses + 2) * 4);
  take(13, size_t(ts::kSegClasses) * B * 4);
  take(14, long_keys_bytes(fmax));
  const size_t back_bytes = ts::kSegLongHeader + fmax * 16;
  take(15, back_bytes);
  take(16, long_radix_bytes(n, kb));
  ts_status st = ensure(c, b);
  if (st == TS_OK) st = ensure_pinned(c, up(back_bytes) + long_keys_pin_bytes(fmax));
  if (st != TS_OK) return st;
  char* w = c->ws;
  ts::LearnedJob j;
  j.key_bytes = kb;
  j.mode = mode;
  j.data = data;
  j.scratch = w + off[2];
  j.n = n;
  j.sample = S;
  j.sample_a = reinterpret_cast<unsigned long long*>(w + off[3]);
  j.sample_b = reinterpret_cast<unsigned long long*>(w + off[4]);
  j.points = T;
  j.u = reinterpret_cast<float*>(w + off[5]);
  j.t = reinterpret_cast<float*>(w + off[6]);
  j.theta = reinterpret_cast<float*>(w + off[7]);
  j.grad = reinterpret_cast<float*>(w + off[8]);
  j.knot_first = reinterpret_cast<unsigned long long*>(w + off[8] + ts::kKnots * 4);
  j.loss = reinterpret_cast<float*>(w + off[9]);
  j.stats = reinterpret_cast<unsigned long long*>(w + off[9] + 16);
  j.train = cfg.train != 0;
  j.iters = int(cfg.iters);
  j.lr = cfg.lr;
  j.model = reinterpret_cast<ts::LearnedModel*>(w + off[10]);
  j.bucket_shift = 32 - cfg.buckets_log2;
  j.buckets = B;
  j.offsets = reinterpret_cast<unsigned long long*>(w + off[11]);
  j.meta = reinterpret_cast<ts::SortMeta*>(w + off[0]);
  c->last_meta = j.meta;
  j.status = w + off[1];
  j.status_cap = sb;
  j.error = c->dflags;
  j.num_cus = c->num_cus;
  j.atomic_rank = c->atomic_rank;
  // The judge's reference: the input keys' multiset fingerprint, taken by
  // the partition's histogram read.
  hipError_t e = hipSuccess;
  if (cfg.verify) {
    j.fingerprint = fp_slot(c, kFpIn);
    e = ts::fingerprint_zero(j.fingerprint, s);
  }
  if (e == hipSuccess) e = ts::learned_partition(j, s, &c->clk);
  if (e != hipSuccess) return hip_fail(c, e, "learned_partition");

  ts::SegJob sj;
  sj.key_bytes = kb;
  sj.mode = mode;
  sj.keys = data;
  sj.n = n;
  sj.offsets = j.offsets;
  sj.n_segments = B;
  sj.counts = reinterpret_cast<uint32_t*>(w + off[12]);
  sj.lists = reinterpret_cast<uint32_t*>(w + off[13]);
  sj.num_cus = c->num_cus;
  sj.atomic_rank = c->atomic_rank;
  sj.error = c->dflags;
  set_variants(c, kb, false, sj);
  sj.stage = 1;  // classify only: the readback block is queued between the classify and the bucket sort
  e = ts::segmented_sort(sj, s, &c->clk);
  if (e != hipSuccess) return hip_fail(c, e, "segmented_sort");

  // One readback: the class counts, the fit's losses (4 words from j.loss)
  // and the largest bucket (the 2 words of j.stats after them) in the
  // header, then the overflowing buckets' bounds.
  std::vector<std::pair<uint64_t, uint64_t>> fat;
  auto* back = reinterpret_cast<unsigned long long*>(w + off[15]);
  // The buckets that fit a workgroup are sorted on the caller's stream
  // (queued before the host waits for the readback) while the overflowing
  // ones (disjoint ranges, their own scratch) go through the split rounds on
  // the second stream, whose host waits then overlap that sort.
  if (!c->aux) e = hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking);
  if (e == hipSuccess && !c->aux_done) e = hipEventCreateWithFlags(&c->aux_done, hipEventDisableTiming);
  if (e == hipSuccess && !c->back_ready) e = hipEventCreateWithFlags(&c->back_ready, hipEventDisableTiming);
  if (e != hipSuccess) return hip_fail(c, e, "auxiliary stream");
  e = ts::seg_long_bounds(sj, back, uint32_t(fmax), reinterpret_cast<const uint32_t*>(j.loss), 6, s);
  if (e == hipSuccess) e = start_long_bounds(c, back, fmax, s);
  if (e == hipSuccess) e = hipEventRecord(c->back_ready, s);
  if (e != hipSuccess) return hip_fail(c, e, "learned buckets");
  sj.stage = 2;
  e = ts::segmented_sort(sj, s, &c->clk);
  if (e != hipSuccess) return hip_fail(c, e, "segmented_sort");
  e = hipEventSynchronize(c->back_ready);
  if (e == hipSuccess) e = hipStreamWaitEvent(c->aux, c->back_ready, 0);
  if (e == hipSuccess) e = finish_long_bounds(c, back, fmax, c->aux, &fat);
  if (e != hipSuccess) return hip_fail(c, e, "learned buckets");
  const uint32_t* h = reinterpret_cast<const uint32_t*>(c->pin);
  const uint32_t nbig = h[ts::kSegDevice];
  float loss[3] = {0.f, 0.f, 0.f};
  unsigned long long largest = 0;
  if (j.train) std::memcpy(loss, h + ts::kSegClasses + 2, sizeof loss);
  std::memcpy(&largest, h + ts::kSegClasses + 2 + 4, sizeof largest);
  uint64_t okeys = 0;
  for (const auto& sg : fat) okeys += sg.second;
  uint32_t rounds = 0;
  if (!fat.empty()) {
    st = sort_long_keys(c, kb, mode, data, j.scratch, n, std::move(fat), w + off[14], w + off[16],
                        c->pin + up(back_bytes), fmax, c->aux, &rounds);
    if (st != TS_OK) return st;
    e = hipEventRecord(c->aux_done, c->aux);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, c->aux_done, 0);
    if (e != hipSuccess) return hip_fail(c, e, "auxiliary stream join");
  }
  uint32_t verified = 0;
  if (cfg.verify) {
    // The judge: every output key in order, and the same multiset as the
    // input's (fingerprint taken before the partition).
    uint64_t want[3] = {0, 0, 0};
    e = hipMemcpyAsync(want, fp_slot(c, kFpIn), sizeof want, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return hip_fail(c, e, "verify");
    // verify_output reads `want` after waiting for the stream, by when the
    // copy above has landed.  (Checking the buckets that fit a workgroup
    // while the overflowing ones are still split on the second stream was
    // measured slower: the check then competes with the overflow rounds,
    // which end the sort.)
    st = verify_output(c, kb, mode, data, n, want, s);
    if (st != TS_OK) return st;
    verified = 1;
  }
  st = check_flags(c, s);
  if (st != TS_OK) return st;
  if (report) {
    report->sample = S;
    report->buckets = B;
    report->overflow_buckets = nbig;
    report->split_rounds = rounds;
    report->overflow_keys = okeys;
    report->largest_bucket = largest;
    report->loss_before = loss[0];
    report->loss_after = loss[1];
    report->loss_uniform = loss[2];
    report->verified = verified;
  }
  return TS_OK;
}

ts_status ts_verify_sorted(ts_ctx* c, ts_dtype dt, const void* data, size_t n, const uint64_t* expect_sum_xor,
                           void* stream) {
  if (!c) return TS_ERR_INVALID_ARG;
  if (!valid_dtype(dt)) return fail(c, TS_ERR_INVALID_ARG, "unknown dtype");
  if (n && !data) return fail(c, TS_ERR_INVALID_ARG, "null data");
  DeviceScope ds(c->device);
  return verify_output(c, ts::dtype_bytes(dt), ts::dtype_mode(dt), data, n, expect_sum_xor,
                       static_cast<hipStream_t>(stream));
}

ts_status ts_pwl_fit(ts_ctx* c, const float* u, const float* t, size_t T, float* theta, float* grad, float* loss2,
                     int iters, float lr, void* stream) {
  if (!c) return TS_ERR_INVALID_ARG;
  if (T == 0 || T > size_t(UINT32_MAX) || !u || !t || !theta || !grad || !loss2 || iters < 0 || !(lr > 0.f))
    return fail(c, TS_ERR_INVALID_ARG, "invalid fit arguments");
  DeviceScope ds(c->device);
  hipError_t e = ts::pwl_fit(u, t, uint32_t(T), theta, grad, loss2, iters, lr, static_cast<hipStream_t>(stream));
  return e == hipSuccess ? TS_OK : hip_fail(c, e, "pwl_fit");
}

int ts_segmented_variants(ts_dtype dt, int with_values, int size_class) {
  if (!valid_dtype(dt)) return 0;
  return ts::seg_variants(ts::dtype_bytes(dt), with_values != 0, size_class);
}

const char* ts_segmented_variant_name(ts_dtype dt, int with_values, int size_class, int variant) {
  if (!valid_dtype(dt)) return nullptr;
  return ts::seg_variant_name(ts::dtype_bytes(dt), with_values != 0, size_class, variant);
}

ts_status ts_segmented_set_variant(ts_ctx* c, ts_dtype dt, int with_values, int size_class, int variant) {
  if (!c) return TS_ERR_INVALID_ARG;
  if (!valid_dtype(dt)) return fail(c, TS_ERR_INVALID_ARG, "unknown dtype");
  const int kb = ts::dtype_bytes(dt);
  if (variant < 0 || variant >= ts::seg_variants(kb, with_values != 0, size_class))
    return fail(c, TS_ERR_INVALID_ARG, "no such size class or variant");
  c->seg_variant[width_index(kb)][with_values ? 1 : 0][size_class] = uint8_t(variant);
  return TS_OK;
}

int ts_segmented_get_variant(const ts_ctx* c, ts_dtype dt, int with_values, int size_class) {
  if (!c || !valid_dtype(dt) || ts::seg_variants(ts::dtype_bytes(dt), with_values != 0, size_class) == 0) return -1;
  return c->seg_variant[width_index(ts::dtype_bytes(dt))][with_values ? 1 : 0][size_class];
}

size_t ts_segmented_lds_max_len(ts_dtype dt, int with_values) {
  if (!valid_dtype(dt)) return 0;
  return ts::seg_lds_max_len(ts::dtype_bytes(dt), with_values != 0);
}

}  // extern "C"

struct ts_sbk {
  ts_ctx* ctx = nullptr;
  uint32_t* order = nullptr;
  size_t cap = 0, len = 0;
};

extern "C" {

ts_status ts_sbk_create(ts_sbk** out, int device) {
  if (!out) return TS_ERR_INVALID_ARG;
  *out = nullptr;
  ts_ctx* c = nullptr;
  ts_status st = ts_ctx_create(&c, device);
  if (st != TS_OK) return st;
  ts_sbk* b = new ts_sbk();
  b->ctx = c;
  *out = b;
  return TS_OK;
}

void ts_sbk_destroy(ts_sbk* b) {
  if (!b) return;
  {
    DeviceScope ds(b->ctx->device);
    if (b->order) hipFree(b->order);
  }
  ts_ctx_destroy(b->ctx);
  delete b;
}

size_t ts_sbk_capacity(const ts_sbk* b) { return b ? b->cap : 0; }

const uint32_t* ts_sbk_order(const ts_sbk* b, size_t* len) {
  if (len) *len = b ? b->len : 0;
  return b ? b->order : nullptr;
}

const char* ts_sbk_last_error(const ts_sbk* b) { return b ? ts_ctx_last_error(b->ctx) : "null buffer"; }

ts_status ts_sbk_sort(ts_sbk* b, ts_dtype kt, void* records, size_t n, size_t stride, size_t key_offset,
                      int gather_records, void* stream) {
  if (!b) return TS_ERR_INVALID_ARG;
  if (n > size_t(UINT32_MAX))
    return fail(b->ctx, TS_ERR_TOO_LONG, "more than u32::MAX elements: order() stays empty");
  if (n > b->cap) {
    DeviceScope ds(b->ctx->device);
    if (b->order) {
      hipError_t e = hipFree(b->order);
      b->order = nullptr;
      b->cap = 0;
      if (e != hipSuccess) return hip_fail(b->ctx, e, "hipFree(order)");
    }
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&b->order), n * sizeof(uint32_t));
    if (e != hipSuccess) {
      b->order = nullptr;
      return hip_fail(b->ctx, e, "hipMalloc(order)");
    }
    b->cap = n;
  }
  ts_status st = ts_sort_by_key(b->ctx, kt, records, n, stride, key_offset, b->order, gather_records, stream);
  b->len = st == TS_OK ? n : 0;
  return st;
}

ts_status ts_sbk_apply_order(ts_sbk* b, void* other, size_t n, size_t elem_bytes, void* stream) {
  if (!b) return TS_ERR_INVALID_ARG;
  return ts_apply_order(b->ctx, b->order, b->len, other, n, elem_bytes, stream);
}

ts_status ts_fill_splitmix64(ts_ctx* c, void* out, size_t n, uint64_t seed, int kind, void* stream) {
  if (!c || kind < 0 || kind > 2 || (n && !out)) return TS_ERR_INVALID_ARG;
  DeviceScope ds(c->device);
  hipError_t e = ts::fill_splitmix64(out, n, seed, kind, static_cast<hipStream_t>(stream));
  return e == hipSuccess ? TS_OK : hip_fail(c, e, "fill_splitmix64");
}

ts_status ts_checksum(ts_ctx* c, ts_dtype dt, const void* data, size_t n, uint64_t out_sum_xor[2],
                      uint64_t* out_unsorted, void* stream) {
  if (!c || !valid_dtype(dt) || !out_sum_xor || !out_unsorted || (n && !data)) return TS_ERR_INVALID_ARG;
  DeviceScope ds(c->device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  ts_status st = ensure(c, up(sizeof(ts::SortMeta)));
  if (st != TS_OK) return st;
  ts::SortMeta* m = reinterpret_cast<ts::SortMeta*>(c->ws);
  hipError_t e = ts::checksum(ts::dtype_bytes(dt), ts::dtype_mode(dt), data, n, m->check, s);
  if (e != hipSuccess) return hip_fail(c, e, "checksum");
  unsigned long long h[4];
  e = hipMemcpyAsync(h, m->check, sizeof(h), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(c, e, "checksum readback");
  out_sum_xor[0] = h[0];
  out_sum_xor[1] = h[1];
  *out_unsorted = h[2];
  return TS_OK;
}

ts_status ts_ctx_set_profiling(ts_ctx* c, int enable) {
  if (!c) return TS_ERR_INVALID_ARG;
  DeviceScope ds(c->device);
  if (enable && !c->clk.made) {
    for (int i = 0; i < 2 * ts::KernelClock::kMax; ++i)
      if (hipEventCreate(&c->clk.ev[i]) != hipSuccess) return fail(c, TS_ERR_HIP, "hipEventCreate");
    c->clk.made = true;
  }
  c->clk.on = enable != 0;
  c->clk.count = 0;
  return TS_OK;
}

int ts_ctx_kernel_times(const ts_ctx* c, const char** names, float* ms, double* bytes, int cap) {
  if (!c || !c->clk.on) return 0;
  DeviceScope ds(c->device);
  const int k = std::min(cap, c->clk.count);
  for (int i = 0; i < k; ++i) {
    float t = 0.f;
    hipEventSynchronize(c->clk.ev[2 * i + 1]);
    hipEventElapsedTime(&t, c->clk.ev[2 * i], c->clk.ev[2 * i + 1]);
    if (names) names[i] = c->clk.name[i];
    if (ms) ms[i] = t;
    if (bytes) bytes[i] = c->clk.bytes[i];
  }
  return k;
}

}  // extern "C"
