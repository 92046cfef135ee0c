// CPU model of the device LSD sort's index arithmetic (turbosort_amd/csrc/radix.hip):
// the segmented digit tables of k_hist, the divert decision and per-pass
// tables of k_scan, the XCD-affine tile claims of k_sweep, each key's output
// position, and the run finish of k_finish (same tile, halo and per-lane
// chunk arithmetic, tiles visited in a random order so that in-place
// finishing sees other tiles' writes).  For every pass it checks that every
// tile is claimed once after its predecessor, that every write lands inside
// the array exactly once, and that the final order equals a stable sort by
// key with positions as tie-break (the order() the reference defines); after
// a divert, that the keys come out sorted and that the second sort is asked
// for exactly when a long run was out of order.
//
// Built with g++ by tests/test_sweep_model.py; no GPU involved.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "../../turbosort_amd/csrc/keys.h"

using namespace ts;

static int failures = 0;
static int divert_cases = 0, fallback_cases = 0;
#define CHECK(c, ...)                                                   \
  do {                                                                  \
    if (!(c)) {                                                         \
      std::fprintf(stderr, "FAIL %s:%d %s: ", __FILE__, __LINE__, #c);  \
      std::fprintf(stderr, __VA_ARGS__);                                \
      std::fprintf(stderr, "\n");                                       \
      ++failures;                                                       \
      return false;                                                     \
    }                                                                   \
  } while (0)

// The reference's divert decision (divert.rs:38-54, 95-104), as k_scan does it.
static uint32_t log2_eighths(uint64_t x) {
  const uint32_t e = 63u - uint32_t(__builtin_clzll(x));
  const uint32_t frac = e >= 3 ? uint32_t(x >> (e - 3)) & 7u : uint32_t(x << (3 - e)) & 7u;
  return 8u * e + frac;
}

// k_finish_local then k_finish, position by position, with the same tiles,
// rounds, marks, run-start bit words and bounded word searches; tiles are
// visited in a random order.  Returns false on a failed check; *fallback
// tells whether the second sort was asked for.
constexpr int kFT = 4096, kFH = 256, kFW = kFT + 2 * kFH, kFWords = kFW / 32 + 1, kFReach = kFH / 32 + 1;
constexpr int kFRounds = 4;
template <class U>
static bool finish_model(std::vector<U>& a, const std::vector<U>& b, bool in_place, unsigned shift,
                         std::vector<int>& writes, std::mt19937_64& g, bool* fallback) {
  const size_t n = a.size();
  auto pre = [&](U k) { return uint64_t(k) >> shift; };
  const size_t tiles = (n + kFT - 1) / kFT;
  std::vector<size_t> order(tiles);
  std::iota(order.begin(), order.end(), size_t(0));
  *fallback = false;
  // k_finish_local
  std::vector<uint32_t> marks(tiles, 0);
  std::shuffle(order.begin(), order.end(), g);
  for (size_t tile : order) {
    const size_t base = tile * kFT;
    const int valid = int(std::min(n - base, size_t(kFT)));
    const std::vector<U>& src = in_place ? a : b;
    std::vector<U> w(src.begin() + base, src.begin() + base + valid), orig(w);
    for (int round = 0; round < kFRounds; ++round)
      for (int p = round & 1; p + 1 < valid; p += 2)
        if (w[p] > w[p + 1]) std::swap(w[p], w[p + 1]);
    bool left = false;
    for (int q = 0; q + 1 < valid; ++q) left |= w[q] > w[q + 1];
    const bool cross = base > 0 && pre(src[base - 1]) == pre(w[0]);
    marks[tile] = (left ? 1u : 0u) | (cross ? 2u : 0u);
    for (int q = 0; q < valid; ++q)
      if (!in_place || w[q] != orig[q]) {
        a[base + q] = w[q];
        ++writes[base + q];
      }
  }
  // k_finish on the marked tiles, in place
  std::shuffle(order.begin(), order.end(), g);
  std::vector<U> w(kFW);
  std::vector<uint32_t> starts(kFWords);
  for (size_t tile : order) {
    if (!(marks[tile] & 3u) && !(tile + 1 < tiles && (marks[tile + 1] & 2u))) continue;
    const size_t base = tile * kFT;
    const size_t lo = base >= size_t(kFH) ? base - kFH : 0;
    const size_t hi = std::min(n, base + kFT + kFH);
    const int wn = int(hi - lo);
    const int own0 = int(base - lo), own1 = int(std::min(n, base + kFT) - lo);
    for (int q = 0; q < wn; ++q) w[q] = a[lo + q];
    std::fill(starts.begin(), starts.end(), 0u);
    for (int q = 0; q <= kFW; ++q) {
      bool bit;
      if (q < wn)
        bit = q == 0 ? lo == 0 : pre(w[q]) != pre(w[q - 1]);
      else
        bit = q == wn && hi == n;
      if (bit) starts[q >> 5] |= 1u << (q & 31);
    }
    for (int q = own0; q < wn; ++q) {
      const bool is_start = (starts[q >> 5] >> (q & 31)) & 1u;
      int rs = -1;
      {
        int wi = q >> 5;
        uint32_t bits = starts[wi] & (0xFFFFFFFFu >> (31 - (q & 31)));
        for (int k = 0; !bits && wi > 0 && k < kFReach; ++k) bits = starts[--wi];
        if (bits) rs = wi * 32 + 31 - __builtin_clz(bits);
      }
      int re = -1;
      if (rs >= 0 && q - rs < kFH) {
        const int q1 = q + 1;
        int wi = q1 >> 5;
        uint32_t bits = starts[wi] & (0xFFFFFFFFu << (q1 & 31));
        for (int k = 0; !bits && wi < kFWords - 1 && k < kFReach; ++k) bits = starts[++wi];
        if (bits) re = wi * 32 + __builtin_ctz(bits);
      }
      const bool short_run = rs >= 0 && re >= 0 && re - rs <= kFH;
      if (short_run) {
        if (rs < own0 || rs >= own1 || re - rs == 1) continue;
        const U k = w[q];
        int rank = 0;
        for (int j = rs; j < re; ++j) rank += (w[j] < k) || (w[j] == k && j < q);
        const int o = rs + rank;
        CHECK(o >= rs && o < re, "finish rank %d outside [%d, %d)", o, rs, re);
        if (o != q) a[lo + o] = k;
      } else if (q < own1 && !is_start && w[q - 1] > w[q]) {
        *fallback = true;
      }
    }
  }
  return true;
}

template <class U>
bool model(const std::vector<U>& keys, size_t tile, const char* name, bool divert, std::mt19937_64& g) {
  const size_t n = keys.size();
  constexpr int P = int(sizeof(U));
  const int S = kSegments;
  const size_t per = (n + S - 1) / S;
  const size_t seg0 = ((per + tile - 1) / tile) * tile;
  auto digit = [](U k, int p) { return unsigned((uint64_t(k) >> (8 * p)) & 0xFF); };
  auto at = [&](int p, int s, int x) { return (size_t(p) * S + s) * 256 + x; };

  // k_hist: value-segment tables, position-segment tables.
  std::vector<uint64_t> pair(size_t(P) * S * 256, 0), pair_pos(size_t(P) * S * 256, 0);
  for (size_t i = 0; i < n; ++i) {
    for (int p = 0; p < P; ++p) {
      const unsigned seg = p == 0 ? unsigned(i / seg0) : unsigned((uint64_t(keys[i]) >> (8 * p - 3)) & 7);
      CHECK(seg < unsigned(S), "%s: segment %u", name, seg);
      pair[at(p, seg, digit(keys[i], p))]++;
      pair_pos[at(p, unsigned(i / seg0), digit(keys[i], p))]++;
    }
  }
  // k_scan
  std::vector<uint64_t> hist(size_t(P) * 256), base(size_t(P) * 256), maxb(P);
  std::vector<int> live;
  for (int p = 0; p < P; ++p) {
    uint64_t run = 0, mx = 0;
    bool dead = false;
    for (int x = 0; x < 256; ++x) {
      uint64_t t = 0;
      for (int s = 0; s < S; ++s) t += pair[at(p, s, x)];
      hist[size_t(p) * 256 + x] = t;
      base[size_t(p) * 256 + x] = run;
      run += t;
      mx = std::max(mx, t);
      dead |= t == n;
    }
    CHECK(run == n, "%s: histogram total", name);
    maxb[p] = mx;
    if (!dead) live.push_back(p);
  }
  uint32_t skip = 0;
  const uint32_t n_live = uint32_t(live.size());
  if (divert && n_live > 3) {
    const uint32_t log_len = log2_eighths(n), need = log_len + 16;
    uint32_t have = 0;
    for (int i = int(n_live) - 1; i >= 3; --i) {
      have += log_len - log2_eighths(maxb[live[i]]);
      if (have >= need) {
        skip = uint32_t(i);
        break;
      }
    }
  }
  const uint32_t n_exec = n_live - skip;
  divert_cases += skip != 0;

  // passes
  std::vector<std::pair<U, uint32_t>> cur(n), out(n);
  for (size_t i = 0; i < n; ++i) cur[i] = {keys[i], uint32_t(i)};
  for (uint32_t pi = 0; pi < n_exec; ++pi) {
    const int d = live[skip + pi];
    uint64_t st[kSegments], ct[kSegments];
    uint32_t nt[kSegments], slot[kSegments];
    std::vector<uint64_t> pb(size_t(S) * 256);
    uint32_t total = 0;
    for (int s = 0; s < S; ++s) {
      if (pi == 0) {
        st[s] = std::min<uint64_t>(n, uint64_t(s) * seg0);
        ct[s] = std::min<uint64_t>(n, st[s] + seg0) - st[s];
      } else {
        ct[s] = 0;
        for (int x = 0; x < 32; ++x) ct[s] += hist[size_t(d - 1) * 256 + s * 32 + x];
        st[s] = base[size_t(d - 1) * 256 + s * 32];
      }
      nt[s] = uint32_t((ct[s] + tile - 1) / tile);
      slot[s] = total;
      total += nt[s];
    }
    CHECK(total <= (n + tile - 1) / tile + S, "%s: more tiles than the launch bound", name);
    for (int x = 0; x < 256; ++x) {
      uint64_t run = base[size_t(d) * 256 + x];
      for (int s = 0; s < S; ++s) {
        pb[size_t(s) * 256 + x] = run;
        run += (pi == 0 && d > 0) ? pair_pos[at(d, s, x)] : pair[at(d, s, x)];
      }
    }
    // Claims: segments' rounds in order, segments interleaved at random (the
    // order XCDs happen to claim them in).
    std::vector<uint32_t> claimed(S, 0);
    std::vector<std::vector<uint64_t>> before(S, std::vector<uint64_t>(256, 0));
    std::vector<char> written(n, 0);
    uint32_t done = 0;
    while (done < total) {
      int s = int(g() % S);
      while (claimed[s] >= nt[s]) s = (s + 1) % S;
      const uint32_t round = claimed[s]++;
      ++done;
      const uint64_t lo = st[s] + uint64_t(round) * tile;
      const uint64_t hi = std::min<uint64_t>(st[s] + ct[s], lo + tile);
      CHECK(lo < hi && hi <= n, "%s: tile range [%llu, %llu)", name, (unsigned long long)lo, (unsigned long long)hi);
      CHECK(slot[s] + round < total, "%s: slot", name);
      std::vector<uint64_t> cnt(256, 0);
      for (uint64_t i = lo; i < hi; ++i) {
        const unsigned x = digit(cur[i].first, d);
        const uint64_t o = pb[size_t(s) * 256 + x] + before[s][x] + cnt[x];
        cnt[x]++;
        CHECK(o < n, "%s: pass %u writes %llu", name, pi, (unsigned long long)o);
        CHECK(!written[o], "%s: pass %u writes %llu twice", name, pi, (unsigned long long)o);
        written[o] = 1;
        out[o] = cur[i];
      }
      for (int x = 0; x < 256; ++x) before[s][x] += cnt[x];
    }
    for (size_t i = 0; i < n; ++i) CHECK(written[i], "%s: pass %u leaves %zu unwritten", name, pi, i);
    std::swap(cur, out);
  }
  std::vector<std::pair<U, uint32_t>> want(n);
  for (size_t i = 0; i < n; ++i) want[i] = {keys[i], uint32_t(i)};
  std::stable_sort(want.begin(), want.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
  if (skip == 0) {
    for (size_t i = 0; i < n; ++i)
      CHECK(cur[i] == want[i], "%s: position %zu holds (%llu, %u), want (%llu, %u)", name, i,
            (unsigned long long)cur[i].first, cur[i].second, (unsigned long long)want[i].first, want[i].second);
    return true;
  }
  // After the top passes the keys are ordered by their bits from `shift` up.
  const unsigned shift = 8u * unsigned(live[skip]);
  for (size_t i = 1; i < n; ++i)
    CHECK((uint64_t(cur[i - 1].first) >> shift) <= (uint64_t(cur[i].first) >> shift), "%s: prefix order at %zu",
          name, i);
  // k_finish: in place after an even number of passes, else from the second buffer.
  const bool in_place = (n_exec & 1u) == 0;
  std::vector<U> a(n), b(n);
  for (size_t i = 0; i < n; ++i) (in_place ? a : b)[i] = cur[i].first;
  if (!in_place)
    for (size_t i = 0; i < n; ++i) a[i] = U(~U(0));
  std::vector<int> writes(n, 0);
  bool fallback = false;
  if (!finish_model(a, b, in_place, shift, writes, g, &fallback)) return false;
  fallback_cases += fallback;
  if (!in_place)
    for (size_t i = 0; i < n; ++i) CHECK(writes[i] == 1, "%s: local finish writes %zu %d times", name, i, writes[i]);
  // Expected: the second sort is needed exactly when a run longer than the
  // halo is out of order.
  bool long_unsorted = false;
  for (size_t i = 0; i < n;) {
    size_t e = i + 1;
    while (e < n && (uint64_t(cur[e].first) >> shift) == (uint64_t(cur[i].first) >> shift)) ++e;
    if (e - i > size_t(kFH))
      for (size_t j = i + 1; j < e; ++j) long_unsorted |= cur[j - 1].first > cur[j].first;
    i = e;
  }
  CHECK(fallback == long_unsorted, "%s: fallback %d, long unsorted run %d", name, int(fallback), int(long_unsorted));
  std::vector<U> multiset(a);
  std::sort(multiset.begin(), multiset.end());
  for (size_t i = 0; i < n; ++i) CHECK(multiset[i] == want[i].first, "%s: finish lost key at %zu", name, i);
  if (!fallback)
    for (size_t i = 0; i < n; ++i)
      CHECK(a[i] == want[i].first, "%s: finished position %zu holds %llu, want %llu", name, i,
            (unsigned long long)a[i], (unsigned long long)want[i].first);
  return true;
}

template <class U> void shapes(size_t n, size_t tile, uint64_t seed) {
  std::mt19937_64 g(seed);
  std::vector<U> v(n);
  auto go = [&](const char* nm) {
    for (int dv = 0; dv < 2; ++dv) {
      if (dv && sizeof(U) != 8) continue;
      char buf[160];
      std::snprintf(buf, sizeof buf, "%s w=%zu n=%zu tile=%zu divert=%d", nm, sizeof(U), n, tile, dv);
      model<U>(v, tile, buf, dv != 0, g);
    }
  };
  for (auto& x : v) x = U(g());
  go("uniform");
  for (size_t i = 0; i < n; ++i) v[i] = U(i % 7);
  go("few unique");
  for (size_t i = 0; i < n; ++i) v[i] = U(i);
  go("sorted");
  for (size_t i = 0; i < n; ++i) v[i] = U(n - i);
  go("reversed");
  for (auto& x : v) x = U(42);
  go("constant");
  for (auto& x : v) x = U(g() & ~(uint64_t(0xFF) << 8));  // digit 1 dead
  go("dead middle digit");
  for (auto& x : v) x = U(g() & ~uint64_t(0xFF));  // digit 0 dead
  go("dead low digit");
  constexpr int hi_shift = sizeof(U) >= 2 ? 8 * (int(sizeof(U)) - 2) : 0;
  for (auto& x : v) x = U(uint64_t(g() % 5000) << hi_shift);
  go("high digits only");
  for (auto& x : v) {
    const uint64_t r = g();
    x = U((r & 1) ? r : (r & 0xE0E0E0E0E0E0E0E0ull));
  }
  go("skewed segments");
}

// 8-byte keys whose top half takes `k` random values (each byte of it looks
// uniform, so the divert skips the low half) and whose low half is random
// (runs of about n / k keys, out of order) or constant (sorted runs).
void divert_shapes(size_t n, uint64_t seed) {
  std::mt19937_64 g(seed);
  std::vector<uint64_t> v(n);
  for (size_t k : {size_t(n / 3), size_t(n / 100), size_t(n / 700), size_t(n / 5000 + 1)}) {
    std::vector<uint64_t> tops(k);
    for (auto& t : tops) t = g() >> 32 << 32;
    for (int low_random = 0; low_random < 2; ++low_random) {
      for (auto& x : v) x = tops[g() % k] | (low_random ? (g() & 0xFFFFFFFFull) : 7ull);
      char buf[160];
      std::snprintf(buf, sizeof buf, "runs n=%zu k=%zu low=%s", n, k, low_random ? "random" : "constant");
      model<uint64_t>(v, 4096, buf, true, g);
    }
  }
}

int main() {
  const size_t tiles[] = {64, 256};
  const size_t sizes[] = {1, 2, 63, 64, 65, 100, 513, 1000, 4097, 20000, 100003};
  uint64_t seed = 1;
  for (size_t t : tiles)
    for (size_t n : sizes) {
      shapes<uint8_t>(n, t, seed++);
      shapes<uint16_t>(n, t, seed++);
      shapes<uint32_t>(n, t, seed++);
      shapes<uint64_t>(n, t, seed++);
    }
  // The device tile sizes at a few larger lengths.
  for (size_t n : {size_t(8191), size_t(8193), size_t(70000), size_t(300001)}) {
    shapes<uint32_t>(n, 8192, seed++);
    shapes<uint64_t>(n, 4096, seed++);
  }
  for (size_t n : {size_t(50000), size_t(200003)}) divert_shapes(n, seed++);
  std::printf("sweep model: %d diverted sorts, %d asked for the second sort\n", divert_cases, fallback_cases);
  if (divert_cases < 20 || fallback_cases < 2) {
    std::printf("sweep model: too few diverted cases\n");
    return 1;
  }
  if (failures) {
    std::printf("sweep model: %d failures\n", failures);
    return 1;
  }
  std::printf("sweep model: all checks passed\n");
  return 0;
}
