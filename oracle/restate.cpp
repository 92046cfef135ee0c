// CPU restatement of turbosort's sort path -- TEST INFRASTRUCTURE ONLY.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// this library, as the checker and as the timed CPU baseline.  The product
// (libtsort.so) never links it.  Citations are to the reference crate
// (turbosort 0.2.2); the algorithm steps are restated, the code is this
// project's own.
#include "restate.h"

#include <algorithm>
#include <cstring>
#include <thread>
#include <type_traits>
#include <vector>

namespace {

// ---------------------------------------------------------------- key model
// key.rs:115-198.  Each sortable type maps to an unsigned key of its width
// whose unsigned order is the type's total order.

template <class T> struct KeyOf;

template <class U> struct UnsignedIdentity {
  using Key = U;
  static Key to(U x) { return x; }       // key.rs:122
  static U from(Key k) { return k; }     // key.rs:125
};
template <> struct KeyOf<uint8_t> : UnsignedIdentity<uint8_t> {};
template <> struct KeyOf<uint16_t> : UnsignedIdentity<uint16_t> {};
template <> struct KeyOf<uint32_t> : UnsignedIdentity<uint32_t> {};
template <> struct KeyOf<uint64_t> : UnsignedIdentity<uint64_t> {};

template <class S, class U> struct SignedFlip {
  using Key = U;
  static constexpr U kSign = U(U(1) << (sizeof(U) * 8 - 1));
  static Key to(S x) { return U(U(x) ^ kSign); }     // key.rs:144
  static S from(Key k) { return S(U(k ^ kSign)); }   // key.rs:149
};
template <> struct KeyOf<int8_t> : SignedFlip<int8_t, uint8_t> {};
template <> struct KeyOf<int16_t> : SignedFlip<int16_t, uint16_t> {};
template <> struct KeyOf<int32_t> : SignedFlip<int32_t, uint32_t> {};
template <> struct KeyOf<int64_t> : SignedFlip<int64_t, uint64_t> {};

template <> struct KeyOf<float> {
  using Key = uint32_t;
  static Key to(float x) {  // key.rs:174: positive -> set sign bit, negative -> flip all
    uint32_t b;
    std::memcpy(&b, &x, 4);
    uint32_t spread = uint32_t(int32_t(b) >> 31);
    return b ^ (spread | 0x80000000u);
  }
  static float from(Key k) {  // key.rs:180
    uint32_t spread = uint32_t(int32_t(k) >> 31);
    uint32_t b = k ^ (~spread | 0x80000000u);
    float x;
    std::memcpy(&x, &b, 4);
    return x;
  }
};
template <> struct KeyOf<double> {
  using Key = uint64_t;
  static Key to(double x) {  // key.rs:191
    uint64_t b;
    std::memcpy(&b, &x, 8);
    uint64_t spread = uint64_t(int64_t(b) >> 63);
    return b ^ (spread | 0x8000000000000000ull);
  }
  static double from(Key k) {  // key.rs:196
    uint64_t spread = uint64_t(int64_t(k) >> 63);
    uint64_t b = k ^ (~spread | 0x8000000000000000ull);
    double x;
    std::memcpy(&x, &b, 8);
    return x;
  }
};

template <class T> using KeyT = typename KeyOf<T>::Key;
template <class T> inline KeyT<T> rk(const T& x) { return KeyOf<T>::to(x); }
template <class K> inline unsigned byte_digit(K k, unsigned pass) {  // key.rs:86-88
  return unsigned((uint64_t(k) >> (8 * pass)) & 0xFF);
}

// ------------------------------------------------------------ small helpers

// arch/scalar.rs:13-24: insertion sort on radix keys.
template <class T> void insertion_sort(T* s, size_t n) {
  for (size_t i = 1; i < n; ++i) {
    T e = s[i];
    auto k = rk(e);
    size_t j = i;
    while (j > 0 && rk(s[j - 1]) > k) {
      s[j] = s[j - 1];
      --j;
    }
    s[j] = e;
  }
}

// arch/scalar.rs:31-33: unstable comparison sort on radix keys.  Equal radix
// keys are equal bit patterns, so which comparison sort runs is invisible.
template <class T> void comparison_sort(T* s, size_t n) {
  std::sort(s, s + n, [](const T& a, const T& b) { return rk(a) < rk(b); });
}

// presorted.rs:15-65: finish non-decreasing input, or reverse non-increasing.
template <class T> bool finish_presorted(T* s, size_t n) {
  if (n < 2) return true;
  if (rk(s[0]) <= rk(s[1])) {
    for (size_t i = 0; i + 1 < n; ++i)
      if (!(rk(s[i]) <= rk(s[i + 1]))) return false;
    return true;
  }
  for (size_t i = 0; i + 1 < n; ++i)
    if (!(rk(s[i]) >= rk(s[i + 1]))) return false;
  std::reverse(s, s + n);
  return true;
}

// divert.rs:95-104: log2 in eighths of a bit, never above the true value.
uint32_t log2_eighths(uint64_t x) {
  uint32_t whole = 63u - uint32_t(__builtin_clzll(x));
  uint64_t frac = whole >= 3 ? (x >> (whole - 3)) & 7 : (x << (3 - whole)) & 7;
  return 8 * whole + uint32_t(frac);
}

// divert.rs:19-54.
constexpr uint32_t kSlackBits = 2;
constexpr size_t kShortRun = 16;
constexpr size_t kMinSkip = 3;

template <class F> size_t skippable_digits(size_t n_live, size_t len, F max_bucket) {
  uint32_t log_len = log2_eighths(len);
  uint32_t need = log_len + 8 * kSlackBits;
  uint32_t have = 0;
  for (size_t i = n_live; i-- > kMinSkip;) {
    have += log_len - log2_eighths(max_bucket(i));
    if (have >= need) return i;
  }
  return 0;
}

// divert.rs:60-90: sort each run that agrees on every bit from `shift` up.
template <class T> void sort_run(T* s, size_t n) {
  if (n > kShortRun)
    comparison_sort(s, n);
  else if (n > 1)
    insertion_sort(s, n);
}
template <class T> void finish_runs(T* s, size_t n, uint32_t shift) {
  if (n == 0) return;
  auto prefix = [shift](const T& x) { return uint64_t(rk(x)) >> shift; };
  size_t start = 0;
  uint64_t run = prefix(s[0]);
  for (size_t i = 1; i < n; ++i) {
    uint64_t p = prefix(s[i]);
    if (p != run) {
      sort_run(s + start, i - start);
      start = i;
      run = p;
    }
  }
  sort_run(s + start, n - start);
}

// histogram.rs:55-68: the bits in which some key differs from the first key,
// or "none" (false) once every byte has differed.
template <class T> bool varying_bits(const T* s, size_t n, uint64_t* out) {
  const unsigned bytes = sizeof(KeyT<T>);
  uint64_t first = uint64_t(rk(s[0]));
  uint64_t diff = 0;
  for (size_t b = 0; b < n; b += 64) {
    size_t e = std::min(n, b + 64);
    for (size_t i = b; i < e; ++i) diff |= uint64_t(rk(s[i])) ^ first;
    bool every = true;
    for (unsigned p = 0; p < bytes; ++p) every &= ((diff >> (8 * p)) & 0xFF) != 0;
    if (every) return false;
  }
  *out = diff;
  return true;
}

// radix/mod.rs:123-142: counting sort for one-byte keys.
template <class T> void counting_sort(T* s, size_t n) {
  size_t counts[256] = {0};
  for (size_t i = 0; i < n; ++i) counts[rk(s[i])]++;
  size_t at = 0;
  for (unsigned d = 0; d < 256; ++d) {
    T v = KeyOf<T>::from(KeyT<T>(d));
    for (size_t c = 0; c < counts[d]; ++c) s[at++] = v;
  }
}

// -------------------------------------------------------- serial radix cores

// radix/mod.rs:205-271: byte-digit LSD with constant-digit skipping and divert.
template <class T> void sort_core_bytes(T* s, size_t n, T* scratch) {
  constexpr unsigned P = sizeof(KeyT<T>);
  uint64_t varying = 0;
  bool some = varying_bits(s, n, &varying);
  bool counted[8] = {false};
  for (unsigned p = 0; p < P; ++p) counted[p] = !some || ((varying >> (8 * p)) & 0xFF) != 0;
  std::vector<size_t> hist(P * 256, 0);
  for (size_t i = 0; i < n; ++i) {
    auto k = rk(s[i]);
    for (unsigned p = 0; p < P; ++p)
      if (counted[p]) hist[p * 256 + byte_digit(k, p)]++;
  }
  auto first = rk(s[0]);
  for (unsigned p = 0; p < P; ++p)  // histogram.rs:103-108: an uncounted row holds every key
    if (!counted[p]) hist[p * 256 + byte_digit(first, p)] = n;

  unsigned live[8];
  size_t n_live = 0;
  for (unsigned p = 0; p < P; ++p)
    if (hist[p * 256 + byte_digit(first, p)] < n) live[n_live++] = p;
  size_t skip = skippable_digits(n_live, n, [&](size_t i) {
    return uint64_t(*std::max_element(&hist[live[i] * 256], &hist[live[i] * 256] + 256));
  });

  bool in_scratch = false;
  for (size_t li = skip; li < n_live; ++li) {
    unsigned p = live[li];
    size_t off[256];
    size_t sum = 0;
    for (unsigned d = 0; d < 256; ++d) {  // prefix_sum.rs:17-24
      off[d] = sum;
      sum += hist[p * 256 + d];
    }
    const T* src = in_scratch ? scratch : s;
    T* dst = in_scratch ? s : scratch;
    for (size_t i = 0; i < n; ++i) dst[off[byte_digit(rk(src[i]), p)]++] = src[i];  // scatter.rs:26-40
    in_scratch = !in_scratch;
  }
  if (in_scratch) std::memcpy(s, scratch, n * sizeof(T));
  if (skip > 0) finish_runs(s, n, 8 * live[skip]);
}

// Digit helpers for an arbitrary width (key.rs:91-95).
template <class K> inline size_t wide_digit(K k, uint32_t shift, size_t mask) {
  return size_t(uint64_t(k) >> shift) & mask;
}

// radix/mod.rs:518-600: wide-digit LSD (BINS bins, PASSES passes).
template <class T, size_t BINS, size_t PASSES> void sort_core_wide(T* s, size_t n, T* scratch) {
  const uint32_t bits = uint32_t(__builtin_ctzll(BINS));
  const size_t mask = BINS - 1;
  uint64_t varying = 0;
  bool some = varying_bits(s, n, &varying);
  bool counted[PASSES];
  for (size_t p = 0; p < PASSES; ++p) counted[p] = !some || ((varying >> (p * bits)) & mask) != 0;
  auto first = rk(s[0]);
  std::vector<size_t> hist(PASSES * BINS, 0);
  for (size_t i = 0; i < n; ++i) {
    auto k = rk(s[i]);
    for (size_t p = 0; p < PASSES; ++p)
      if (counted[p]) hist[p * BINS + wide_digit(k, uint32_t(p * bits), mask)]++;
  }
  for (size_t p = 0; p < PASSES; ++p)
    if (!counted[p]) hist[p * BINS + wide_digit(first, uint32_t(p * bits), mask)] = n;
  size_t live[PASSES];
  size_t n_live = 0;
  for (size_t p = 0; p < PASSES; ++p)
    if (hist[p * BINS + wide_digit(first, uint32_t(p * bits), mask)] < n) live[n_live++] = p;
  size_t skip = skippable_digits(n_live, n, [&](size_t i) {
    return uint64_t(*std::max_element(&hist[live[i] * BINS], &hist[live[i] * BINS] + BINS));
  });
  std::vector<size_t> off(BINS);
  bool in_scratch = false;
  for (size_t li = skip; li < n_live; ++li) {
    size_t p = live[li];
    size_t sum = 0;
    for (size_t d = 0; d < BINS; ++d) {
      off[d] = sum;
      sum += hist[p * BINS + d];
    }
    const T* src = in_scratch ? scratch : s;
    T* dst = in_scratch ? s : scratch;
    uint32_t shift = uint32_t(p * bits);
    for (size_t i = 0; i < n; ++i) dst[off[wide_digit(rk(src[i]), shift, mask)]++] = src[i];
    in_scratch = !in_scratch;
  }
  if (in_scratch) std::memcpy(s, scratch, n * sizeof(T));
  if (skip > 0) finish_runs(s, n, uint32_t(live[skip] * bits));
}

// scatter.rs:189-198 (x86_64): destinations of at least 2 MiB stream.
template <class T> bool streams(size_t n) { return n * sizeof(T) >= (size_t(2) << 20); }

// radix/mod.rs:180-198.
template <class T, bool WIDE> void sort_raw(T* s, size_t n, T* scratch) {
  constexpr size_t B = sizeof(KeyT<T>);
  if (B == 8 && WIDE && n >= 65536 && !streams<T>(n)) {
    sort_core_wide<T, 2048, 6>(s, n, scratch);
    return;
  }
  sort_core_bytes<T>(s, n, scratch);
}

// dispatch.rs:31-67 and :77-111 (with_buffer chooses WIDE = false).
template <class T, bool WIDE> void dispatch_sort(T* s, size_t n, T* buffer) {
  if (n <= 1) return;
  if (n <= 16) {  // tiny tier: networks or core sort, same unique output
    comparison_sort(s, n);
    return;
  }
  if (finish_presorted(s, n)) return;
  if (sizeof(KeyT<T>) == 1 && n >= 48) {  // COUNTING_MIN, dispatch.rs:25
    counting_sort(s, n);
    return;
  }
  if (n <= 512) {
    comparison_sort(s, n);
    return;
  }
  if (sizeof(KeyT<T>) == 1) {  // radix/mod.rs:65-68
    counting_sort(s, n);
    return;
  }
  if (buffer) {
    sort_raw<T, WIDE>(s, n, buffer);
  } else {
    std::vector<T> scratch(n);
    sort_raw<T, WIDE>(s, n, scratch.data());
  }
}

// ------------------------------------------------------------ parallel core

template <class F> void parallel_for(size_t count, int threads, F&& fn) {
  if (count == 0) return;
  size_t workers = std::min<size_t>(count, size_t(std::max(threads, 1)));
  if (workers == 1) {
    for (size_t t = 0; t < count; ++t) fn(t);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve(workers);
  for (size_t w = 0; w < workers; ++w)
    pool.emplace_back([&, w]() {
      for (size_t t = w; t < count; t += workers) fn(t);
    });
  for (auto& th : pool) th.join();
}

// parallel.rs:267-291.
template <class T> void finish_runs_parallel(T* s, size_t n, uint32_t shift, size_t parts, int threads) {
  auto prefix = [shift](const T& x) { return uint64_t(rk(x)) >> shift; };
  std::vector<size_t> cuts;
  cuts.push_back(0);
  for (size_t t = 1; t < parts; ++t) {
    size_t cut = std::max(t * n / parts, cuts[t - 1]);
    while (cut > 0 && cut < n && prefix(s[cut]) == prefix(s[cut - 1])) ++cut;
    cuts.push_back(cut);
  }
  cuts.push_back(n);
  parallel_for(parts, threads, [&](size_t t) { finish_runs(s + cuts[t], cuts[t + 1] - cuts[t], shift); });
}

// parallel.rs:81-263.
template <class T, size_t BINS, size_t PASSES> void parallel_core(T* s, size_t n, T* scratch, int threads) {
  const uint32_t bits = uint32_t(__builtin_ctzll(BINS));
  const size_t mask = BINS - 1;
  const size_t chunk = (n + size_t(threads) - 1) / size_t(threads);
  const size_t num_chunks = (n + chunk - 1) / chunk;
  auto range = [&](size_t t) { return std::make_pair(t * chunk, std::min((t + 1) * chunk, n)); };

  // Step 0: varying bits, serial head then a parallel OR of the rest.
  const size_t head = std::min<size_t>(n, 4096);
  uint64_t varying = 0;
  bool some = varying_bits(s, head, &varying);
  if (some) {
    uint64_t first = uint64_t(rk(s[0]));
    size_t rest = n - head;
    size_t rest_chunks = (rest + chunk - 1) / chunk;
    std::vector<uint64_t> part(rest_chunks, 0);
    parallel_for(rest_chunks, threads, [&](size_t t) {
      size_t b = head + t * chunk, e = std::min(n, b + chunk);
      uint64_t acc = 0;
      for (size_t i = b; i < e; ++i) acc |= uint64_t(rk(s[i])) ^ first;
      part[t] = acc;
    });
    for (uint64_t v : part) varying |= v;
  }
  std::vector<size_t> counted;
  for (size_t p = 0; p < PASSES; ++p)
    if (!some || ((varying >> (p * bits)) & mask) != 0) counted.push_back(p);

  // Step 1: per-chunk histograms of the counted digits.
  std::vector<std::vector<size_t>> all_hists(num_chunks);
  parallel_for(num_chunks, threads, [&](size_t t) {
    auto [b, e] = range(t);
    std::vector<size_t> h(PASSES * BINS, 0);
    for (size_t i = b; i < e; ++i) {
      auto k = rk(s[i]);
      for (size_t p : counted) h[p * BINS + wide_digit(k, uint32_t(p * bits), mask)]++;
    }
    all_hists[t] = std::move(h);
  });
  std::vector<size_t> global(PASSES * BINS, 0);
  for (auto& h : all_hists)
    for (size_t i = 0; i < global.size(); ++i) global[i] += h[i];

  std::vector<size_t> live;  // histogram.rs:129-131: trivial if one bin holds all
  for (size_t p : counted) {
    bool trivial = std::find(&global[p * BINS], &global[p * BINS] + BINS, n) != &global[p * BINS] + BINS;
    if (!trivial) live.push_back(p);
  }
  size_t skip = skippable_digits(live.size(), n, [&](size_t i) {
    return uint64_t(*std::max_element(&global[live[i] * BINS], &global[live[i] * BINS] + BINS));
  });
  if (skip >= live.size()) return;
  const size_t first_pass = live[skip];

  std::vector<std::vector<size_t>> chunk_hists(num_chunks);
  for (size_t t = 0; t < num_chunks; ++t)
    chunk_hists[t].assign(all_hists[t].begin() + first_pass * BINS, all_hists[t].begin() + (first_pass + 1) * BINS);
  all_hists.clear();

  bool in_scratch = false;
  for (size_t li = skip; li < live.size(); ++li) {
    const size_t p = live[li];
    const uint32_t shift = uint32_t(p * bits);
    const T* src = in_scratch ? scratch : s;
    T* dst = in_scratch ? s : scratch;
    if (li > skip) {
      parallel_for(num_chunks, threads, [&](size_t t) {
        auto [b, e] = range(t);
        std::vector<size_t> h(BINS, 0);
        for (size_t i = b; i < e; ++i) h[wide_digit(rk(src[i]), shift, mask)]++;
        chunk_hists[t] = std::move(h);
      });
    }
    std::vector<size_t> running(BINS);
    size_t sum = 0;
    for (size_t d = 0; d < BINS; ++d) {
      running[d] = sum;
      sum += global[p * BINS + d];
    }
    std::vector<std::vector<size_t>> offs(num_chunks);
    for (size_t t = 0; t < num_chunks; ++t) {
      offs[t] = running;
      for (size_t d = 0; d < BINS; ++d) running[d] += chunk_hists[t][d];
    }
    parallel_for(num_chunks, threads, [&](size_t t) {
      auto [b, e] = range(t);
      auto& off = offs[t];
      for (size_t i = b; i < e; ++i) {
        T v = src[i];
        dst[off[wide_digit(rk(v), shift, mask)]++] = v;
      }
    });
    in_scratch = !in_scratch;
  }
  if (in_scratch)
    parallel_for(num_chunks, threads, [&](size_t t) {
      auto [b, e] = range(t);
      std::memcpy(s + b, scratch + b, (e - b) * sizeof(T));
    });
  if (skip > 0) finish_runs_parallel(s, n, uint32_t(first_pass * bits), num_chunks, threads);
}

// parallel.rs:46-73.
template <class T> void sort_parallel(T* s, size_t n, int threads) {
  if (n < 131072) {
    dispatch_sort<T, true>(s, n, nullptr);
    return;
  }
  if (finish_presorted(s, n)) return;
  constexpr size_t B = sizeof(KeyT<T>);
  if (B == 1) {
    counting_sort(s, n);
    return;
  }
  std::vector<T> scratch(n);
  if (B == 2) parallel_core<T, 256, 2>(s, n, scratch.data(), threads);
  if (B == 4) parallel_core<T, 256, 4>(s, n, scratch.data(), threads);
  if (B == 8) parallel_core<T, 2048, 6>(s, n, scratch.data(), threads);
}

// ------------------------------------------------------------------ by key

template <class K> struct Entry {
  K key;
  uint32_t pos;
};

// by_key.rs:239-249.
template <class K> void entries_insertion_sort(Entry<K>* e, size_t n) {
  for (size_t i = 1; i < n; ++i) {
    Entry<K> x = e[i];
    size_t j = i;
    while (j > 0 && e[j - 1].key > x.key) {
      e[j] = e[j - 1];
      --j;
    }
    e[j] = x;
  }
}

// by_key.rs:254-310.
template <class K> void sort_entries(std::vector<Entry<K>>& entries, std::vector<Entry<K>>& scratch) {
  const size_t len = entries.size();
  constexpr unsigned P = sizeof(K);
  std::vector<size_t> hist(8 * 256, 0);
  for (auto& e : entries)
    for (unsigned b = 0; b < P; ++b) hist[b * 256 + byte_digit(e.key, b)]++;
  K first = entries[0].key;
  unsigned live[8];
  size_t n_live = 0;
  for (unsigned b = 0; b < P; ++b)
    if (hist[b * 256 + byte_digit(first, b)] < len) live[n_live++] = b;
  size_t skip = skippable_digits(n_live, len, [&](size_t i) {
    return uint64_t(*std::max_element(&hist[live[i] * 256], &hist[live[i] * 256] + 256));
  });
  scratch.assign(len, Entry<K>{K(0), 0});
  for (size_t li = skip; li < n_live; ++li) {
    unsigned b = live[li];
    size_t off[256], sum = 0;
    for (unsigned d = 0; d < 256; ++d) {
      off[d] = sum;
      sum += hist[b * 256 + d];
    }
    for (auto& e : entries) scratch[off[byte_digit(e.key, b)]++] = e;
    std::swap(entries, scratch);
  }
  if (skip > 0) {
    uint32_t shift = 8 * live[skip];
    size_t start = 0;
    for (size_t i = 1; i <= len; ++i) {
      if (i == len || (uint64_t(entries[i].key) >> shift) != (uint64_t(entries[start].key) >> shift)) {
        if (i - start > 1)
          std::sort(entries.begin() + start, entries.begin() + i, [](const Entry<K>& a, const Entry<K>& b) {
            return a.key < b.key || (a.key == b.key && a.pos < b.pos);
          });
        start = i;
      }
    }
  }
}

// by_key.rs:175-235, order part only (the element move is tso_apply_order).
template <class T> int order_by_key(const T* vals, size_t n, uint32_t* order) {
  using K = KeyT<T>;
  if (n < 2) {
    for (size_t i = 0; i < n; ++i) order[i] = uint32_t(i);
    return TSO_OK;
  }
  if (n > size_t(UINT32_MAX)) return TSO_ERR_TOO_LONG;
  std::vector<Entry<K>> entries(n);
  for (size_t i = 0; i < n; ++i) entries[i] = Entry<K>{rk(vals[i]), uint32_t(i)};
  bool asc = true, desc = true;
  for (size_t i = 0; i + 1 < n; ++i) {
    asc &= entries[i].key <= entries[i + 1].key;
    desc &= entries[i].key >= entries[i + 1].key;
  }
  if (asc) {
    for (size_t i = 0; i < n; ++i) order[i] = uint32_t(i);
    return TSO_OK;
  }
  if (desc) {  // by_key.rs:207-226: the run entries[start..i) lands at [n-i, n-start)
    size_t start = 0;
    for (size_t i = 1; i <= n; ++i) {
      if (i == n || entries[i].key != entries[start].key) {
        for (size_t j = 0; j < i - start; ++j) order[n - i + j] = uint32_t(start + j);
        start = i;
      }
    }
    return TSO_OK;
  }
  if (n < 64) {  // INSERTION_MAX, by_key.rs:31
    entries_insertion_sort(entries.data(), n);
  } else {
    std::vector<Entry<K>> scratch;
    sort_entries(entries, scratch);
  }
  for (size_t i = 0; i < n; ++i) order[i] = entries[i].pos;
  return TSO_OK;
}

template <template <class> class Fn, class... A> int by_dtype(int dtype, A&&... a) {
  switch (dtype) {
    case TSO_U8: return Fn<uint8_t>::run(a...);
    case TSO_U16: return Fn<uint16_t>::run(a...);
    case TSO_U32: return Fn<uint32_t>::run(a...);
    case TSO_U64: return Fn<uint64_t>::run(a...);
    case TSO_I8: return Fn<int8_t>::run(a...);
    case TSO_I16: return Fn<int16_t>::run(a...);
    case TSO_I32: return Fn<int32_t>::run(a...);
    case TSO_I64: return Fn<int64_t>::run(a...);
    case TSO_F32: return Fn<float>::run(a...);
    case TSO_F64: return Fn<double>::run(a...);
    default: return TSO_ERR_INVALID_ARG;
  }
}

template <class T> struct SortFn {
  static int run(void* d, size_t n) {
    dispatch_sort<T, true>(static_cast<T*>(d), n, nullptr);
    return TSO_OK;
  }
};
template <class T> struct SortBufFn {
  static int run(void* d, size_t n, void* buf) {
    dispatch_sort<T, false>(static_cast<T*>(d), n, static_cast<T*>(buf));
    return TSO_OK;
  }
};
template <class T> struct SortParFn {
  static int run(void* d, size_t n, int threads) {
    sort_parallel<T>(static_cast<T*>(d), n, threads);
    return TSO_OK;
  }
};
template <class T> struct OrderFn {
  static int run(const void* k, size_t n, uint32_t* order) {
    return order_by_key<T>(static_cast<const T*>(k), n, order);
  }
};

// chunks_exact_mut(size).for_each(sort_by_key) over uneven slices
// (benches/sort_bench.rs:229-250 pattern): per segment the by-key order
// (by_key.rs:175-235), then keys and values moved into it (by_key.rs:315-334).
// Segments are independent, so `threads` std::threads take them in turn.
template <class T> struct SegFn {
  static int run(void* keys, uint32_t* vals, const uint64_t* off, size_t nseg, int threads) {
    T* k = static_cast<T*>(keys);
    parallel_for(nseg, threads, [&](size_t s) {
      const size_t a = size_t(off[s]), m = size_t(off[s + 1] - off[s]);
      if (m < 2) return;
      std::vector<uint32_t> order(m);
      order_by_key<T>(k + a, m, order.data());
      std::vector<T> kk(m);
      std::vector<uint32_t> vv(m);
      for (size_t i = 0; i < m; ++i) {
        kk[i] = k[a + order[i]];
        if (vals) vv[i] = vals[a + order[i]];
      }
      std::memcpy(k + a, kk.data(), m * sizeof(T));
      if (vals) std::memcpy(vals + a, vv.data(), m * sizeof(uint32_t));
    });
    return TSO_OK;
  }
};

}  // namespace

extern "C" {

int tso_segmented_sort_pairs(int dtype, void* keys, uint32_t* vals, const uint64_t* offsets, size_t n_segments,
                             int threads) {
  return by_dtype<SegFn>(dtype, keys, vals, offsets, n_segments, threads);
}

int tso_sort(int dtype, void* data, size_t n) { return by_dtype<SortFn>(dtype, data, n); }

int tso_sort_with_buffer(int dtype, void* data, size_t n, void* buf, size_t buf_len) {
  if (buf_len < n) return TSO_ERR_BUFFER_TOO_SMALL;
  return by_dtype<SortBufFn>(dtype, data, n, buf);
}

int tso_sort_parallel(int dtype, void* data, size_t n, int threads) {
  if (threads < 1) threads = 1;
  return by_dtype<SortParFn>(dtype, data, n, threads);
}

int tso_sort_by_key_order(int key_dtype, const void* keys, size_t n, uint32_t* order) {
  return by_dtype<OrderFn>(key_dtype, keys, n, order);
}

int tso_apply_order(const uint32_t* order, size_t order_len, void* data, size_t n, size_t elem_bytes) {
  if (order_len != n) return TSO_ERR_LENGTH_MISMATCH;
  std::vector<unsigned char> moved(n * elem_bytes);
  const unsigned char* src = static_cast<const unsigned char*>(data);
  for (size_t i = 0; i < n; ++i) std::memcpy(&moved[i * elem_bytes], src + size_t(order[i]) * elem_bytes, elem_bytes);
  std::memcpy(data, moved.data(), n * elem_bytes);
  return TSO_OK;
}

uint32_t tso_log2_eighths(uint64_t x) { return x ? log2_eighths(x) : 0; }

size_t tso_skippable_digits(size_t n_live, size_t len, const uint64_t* max_bucket) {
  return skippable_digits(n_live, len, [&](size_t i) { return max_bucket[i]; });
}

int tso_varying_bits(const uint64_t* keys, size_t n, int key_bytes, uint64_t* out) {
  uint64_t first = keys[0];
  uint64_t diff = 0;
  for (size_t b = 0; b < n; b += 64) {
    size_t e = std::min(n, b + 64);
    for (size_t i = b; i < e; ++i) diff |= keys[i] ^ first;
    bool every = true;
    for (int p = 0; p < key_bytes; ++p) every &= ((diff >> (8 * p)) & 0xFF) != 0;
    if (every) return 0;
  }
  *out = diff;
  return 1;
}

void tso_finish_runs_u16(uint16_t* keys, size_t n, uint32_t shift) { finish_runs(keys, n, shift); }

void tso_splitmix64_fill(uint64_t seed, size_t n, uint64_t* out) {
  for (size_t i = 0; i < n; ++i) {
    uint64_t z = seed + uint64_t(i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    out[i] = z ^ (z >> 31);
  }
}

}  // extern "C"
