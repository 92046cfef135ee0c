// CPU model of the device LSD sort's index arithmetic (turbosort_amd/csrc/radix.hip):
// the segmented pair histogram of k_hist, the tables of k_scan, the ticket
// order of k_sweep (the shared ticket_to_tile helper) and each key's output
// position.  For every pass it checks that the tickets enumerate every tile
// once with each tile's predecessor earlier, that every write lands inside the
// array exactly once, and that the final order equals a stable sort by key
// with positions as tie-break (the order() the reference defines).
//
// Built with g++ by tests/test_sweep_model.py; no GPU involved.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "../../turbosort_amd/csrc/keys.h"

using namespace ts;

static int failures = 0;
#define CHECK(c, ...)                                              \
  do {                                                             \
    if (!(c)) {                                                    \
      std::fprintf(stderr, "FAIL %s:%d %s: ", __FILE__, __LINE__, #c); \
      std::fprintf(stderr, __VA_ARGS__);                           \
      std::fprintf(stderr, "\n");                                  \
      ++failures;                                                  \
      return false;                                                \
    }                                                              \
  } while (0)

template <class U> bool model(const std::vector<U>& keys, size_t tile, const char* name) {
  const size_t n = keys.size();
  constexpr int P = int(sizeof(U));
  const int S = kSegments;
  const size_t per = (n + S - 1) / S;
  const size_t seg0 = ((per + tile - 1) / tile) * tile;
  auto digit = [](U k, int p) { return unsigned((uint64_t(k) >> (8 * p)) & 0xFF); };

  // k_hist
  std::vector<uint64_t> pair(size_t(P) * S * 256, 0);
  for (size_t i = 0; i < n; ++i) {
    for (int p = 0; p < P; ++p) {
      const unsigned seg = p == 0 ? unsigned(i / seg0) : unsigned((uint64_t(keys[i]) >> (8 * p - 3)) & 7);
      CHECK(seg < unsigned(S), "%s: segment %u", name, seg);
      pair[(size_t(p) * S + seg) * 256 + digit(keys[i], p)]++;
    }
  }
  // k_scan
  std::vector<uint64_t> base(size_t(P) * 256), segpre(size_t(P) * S * 256), tot(size_t(P) * 256);
  std::vector<uint64_t> seg_start(size_t(P) * S), seg_count(size_t(P) * S);
  std::vector<uint32_t> seg_tiles(size_t(P) * S), seg_slot(size_t(P) * S), total(P);
  std::vector<int> live;
  for (int p = 0; p < P; ++p) {
    bool dead = false;
    uint64_t run = 0;
    for (int x = 0; x < 256; ++x) {
      uint64_t t = 0;
      for (int s = 0; s < S; ++s) {
        segpre[(size_t(p) * S + s) * 256 + x] = t;
        t += pair[(size_t(p) * S + s) * 256 + x];
      }
      tot[size_t(p) * 256 + x] = t;
      base[size_t(p) * 256 + x] = run;
      run += t;
      dead |= t == n;
    }
    CHECK(run == n, "%s: histogram total", name);
    if (!dead) live.push_back(p);
    uint32_t slot = 0;
    for (int s = 0; s < S; ++s) {
      uint64_t st, ct;
      if (p == 0) {
        st = std::min<uint64_t>(n, uint64_t(s) * seg0);
        ct = std::min<uint64_t>(n, st + seg0) - st;
      } else {
        ct = 0;
        for (int x = 0; x < 32; ++x) ct += tot[size_t(p - 1) * 256 + s * 32 + x];
        st = base[size_t(p - 1) * 256 + s * 32];
      }
      seg_start[size_t(p) * S + s] = st;
      seg_count[size_t(p) * S + s] = ct;
      seg_tiles[size_t(p) * S + s] = uint32_t((ct + tile - 1) / tile);
      seg_slot[size_t(p) * S + s] = slot;
      slot += seg_tiles[size_t(p) * S + s];
    }
    total[p] = slot;
    CHECK(slot <= (n + tile - 1) / tile + S, "%s: more tiles than the launch bound", name);
  }

  // passes
  std::vector<std::pair<U, uint32_t>> cur(n), out(n);
  for (size_t i = 0; i < n; ++i) cur[i] = {keys[i], uint32_t(i)};
  for (int p : live) {
    uint32_t nt[kSegments];
    for (int s = 0; s < S; ++s) nt[s] = seg_tiles[size_t(p) * S + s];
    std::vector<std::vector<int64_t>> ticket_of(S);
    for (int s = 0; s < S; ++s) ticket_of[s].assign(nt[s], -1);
    std::vector<std::vector<uint64_t>> before(S, std::vector<uint64_t>(256, 0));
    std::vector<char> written(n, 0);
    for (uint32_t k = 0; k < total[p]; ++k) {
      uint32_t seg, round;
      ticket_to_tile(k, nt, &seg, &round);
      CHECK(seg < unsigned(S) && round < nt[seg], "%s: ticket %u -> (%u, %u)", name, k, seg, round);
      CHECK(ticket_of[seg][round] < 0, "%s: tile (%u, %u) claimed twice", name, seg, round);
      ticket_of[seg][round] = k;
      CHECK(round == 0 || (ticket_of[seg][round - 1] >= 0 && ticket_of[seg][round - 1] < int64_t(k)),
            "%s: predecessor of (%u, %u) not claimed earlier", name, seg, round);
      const uint64_t lo = seg_start[size_t(p) * S + seg] + uint64_t(round) * tile;
      const uint64_t hi = std::min<uint64_t>(seg_start[size_t(p) * S + seg] + seg_count[size_t(p) * S + seg],
                                             lo + tile);
      CHECK(lo < hi && hi <= n, "%s: tile range [%llu, %llu)", name, (unsigned long long)lo,
            (unsigned long long)hi);
      std::vector<uint64_t> cnt(256, 0);
      for (uint64_t i = lo; i < hi; ++i) {
        const unsigned d = digit(cur[i].first, p);
        const uint64_t o = base[size_t(p) * 256 + d] + segpre[(size_t(p) * S + seg) * 256 + d] + before[seg][d] + cnt[d];
        cnt[d]++;
        CHECK(o < n, "%s: pass %d writes %llu", name, p, (unsigned long long)o);
        CHECK(!written[o], "%s: pass %d writes %llu twice", name, p, (unsigned long long)o);
        written[o] = 1;
        out[o] = cur[i];
      }
      for (int x = 0; x < 256; ++x) before[seg][x] += cnt[x];
    }
    for (size_t i = 0; i < n; ++i) CHECK(written[i], "%s: pass %d leaves %zu unwritten", name, p, i);
    std::swap(cur, out);
  }
  std::vector<std::pair<U, uint32_t>> want(n);
  for (size_t i = 0; i < n; ++i) want[i] = {keys[i], uint32_t(i)};
  std::stable_sort(want.begin(), want.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
  for (size_t i = 0; i < n; ++i)
    CHECK(cur[i] == want[i], "%s: position %zu holds (%llu, %u), want (%llu, %u)", name, i,
          (unsigned long long)cur[i].first, cur[i].second, (unsigned long long)want[i].first, want[i].second);
  return true;
}

template <class U> void shapes(size_t n, size_t tile, uint64_t seed) {
  std::mt19937_64 g(seed);
  std::vector<U> v(n);
  auto go = [&](const char* nm) {
    char buf[128];
    std::snprintf(buf, sizeof buf, "%s w=%zu n=%zu tile=%zu", nm, sizeof(U), n, tile);
    model<U>(v, tile, buf);
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
  if (failures) {
    std::printf("sweep model: %d failures\n", failures);
    return 1;
  }
  std::printf("sweep model: all checks passed\n");
  return 0;
}
