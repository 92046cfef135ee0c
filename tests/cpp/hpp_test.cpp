// Exercises include/turbosort.hpp on the GPU: the C++ mirror of the
// reference API, checked against expected outputs the reference asserts
// (lib.rs:18-20, 82-84, 92-96, 116-119; by_key.rs:102-106, 127-134) and the
// panic messages of dispatch.rs:78-83 and by_key.rs:137-141.
// Built by __graft_entry__.build(); run by tests/test_gpu_cpp_api.py.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "turbosort.hpp"

#define REQUIRE(c)                                                        \
  do {                                                                    \
    if (!(c)) {                                                           \
      std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
      return 1;                                                           \
    }                                                                     \
  } while (0)

template <class T> T* to_dev(const std::vector<T>& v) {
  T* p = nullptr;
  if (hipMalloc(&p, (v.size() ? v.size() : 1) * sizeof(T)) != hipSuccess) return nullptr;
  if (!v.empty()) hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
  return p;
}
template <class T> std::vector<T> to_host(const T* p, size_t n) {
  std::vector<T> v(n);
  if (n) hipMemcpy(v.data(), p, n * sizeof(T), hipMemcpyDeviceToHost);
  return v;
}

struct Person {
  uint32_t name;
  uint32_t age;
};

int main() {
  using namespace turbosort;
  {  // lib.rs:18-20
    std::vector<uint32_t> v{5, 3, 8, 1, 9, 2, 7, 4, 6};
    uint32_t* d = to_dev(v);
    sort(DeviceSlice<uint32_t>{d, v.size()});
    REQUIRE((to_host(d, v.size()) == std::vector<uint32_t>{1, 2, 3, 4, 5, 6, 7, 8, 9}));
    hipFree(d);
  }
  {  // lib.rs:82-84
    std::vector<int32_t> v{3, -1, 4, -1, 5, -9};
    int32_t* d = to_dev(v);
    sort(DeviceSlice<int32_t>{d, v.size()});
    REQUIRE((to_host(d, v.size()) == std::vector<int32_t>{-9, -1, -1, 3, 4, 5}));
    hipFree(d);
  }
  {  // lib.rs:92-96: -inf, -0.0, +0.0, 1.0, NaN
    std::vector<uint32_t> bits{0x3F800000u, 0x7FC00000u, 0x80000000u, 0x00000000u, 0xFF800000u};
    std::vector<float> v(bits.size());
    std::memcpy(v.data(), bits.data(), bits.size() * 4);
    float* d = to_dev(v);
    sort(DeviceSlice<float>{d, v.size()});
    auto out = to_host(d, v.size());
    std::vector<uint32_t> ob(out.size());
    std::memcpy(ob.data(), out.data(), out.size() * 4);
    REQUIRE((ob == std::vector<uint32_t>{0xFF800000u, 0x80000000u, 0x00000000u, 0x3F800000u, 0x7FC00000u}));
    hipFree(d);
  }
  {  // lib.rs:116-119 and the panic of dispatch.rs:78-83
    std::vector<uint64_t> v{42, 7, 99, 1, 0};
    uint64_t* d = to_dev(v);
    uint64_t* b = to_dev(std::vector<uint64_t>(5, 0));
    sort_with_buffer(DeviceSlice<uint64_t>{d, 5}, DeviceSlice<uint64_t>{b, 5});
    REQUIRE((to_host(d, 5) == std::vector<uint64_t>{0, 1, 7, 42, 99}));
    bool panicked = false;
    try {
      sort_with_buffer(DeviceSlice<uint64_t>{d, 5}, DeviceSlice<uint64_t>{b, 4});
    } catch (const Panic& p) {
      panicked = std::strcmp(p.what(), "buffer too small: need 5, got 4") == 0;
    }
    REQUIRE(panicked);
    hipFree(d);
    hipFree(b);
  }
  {  // lib.rs:145-147 / by_key.rs:127-134: order and apply_order
    std::vector<Person> people{{0, 31}, {1, 25}, {2, 31}, {3, 25}};
    Person* d = to_dev(people);
    SortByKeyBuffer buffer;
    REQUIRE(buffer.order().len == 0);
    sort_by_key_with_buffer(DeviceSlice<Person>{d, 4}, buffer, KeyField<uint32_t>{offsetof(Person, age)});
    auto sorted = to_host(d, 4);
    REQUIRE(sorted[0].name == 1 && sorted[1].name == 3 && sorted[2].name == 0 && sorted[3].name == 2);
    auto ord = buffer.order();
    REQUIRE((to_host(ord.ptr, ord.len) == std::vector<uint32_t>{1, 3, 0, 2}));
    std::vector<uint64_t> tags{10, 11, 12, 13};
    uint64_t* t = to_dev(tags);
    buffer.apply_order(DeviceSlice<uint64_t>{t, 4});
    REQUIRE((to_host(t, 4) == std::vector<uint64_t>{11, 13, 10, 12}));
    bool panicked = false;
    try {
      buffer.apply_order(DeviceSlice<uint64_t>{t, 3});
    } catch (const Panic& p) {
      panicked = std::strstr(p.what(), "apply_order: the slice's length differs") != nullptr;
    }
    REQUIRE(panicked);
    REQUIRE(buffer.capacity() >= 4);
    hipFree(d);
    hipFree(t);
  }
  {  // a larger slice through the radix tier
    const size_t n = 1 << 20;
    std::vector<uint32_t> v(n);
    uint64_t s = 0x9E3779B97F4A7C15ull;
    for (auto& x : v) {
      s ^= s << 13;
      s ^= s >> 7;
      s ^= s << 17;
      x = uint32_t(s);
    }
    uint32_t* d = to_dev(v);
    sort(DeviceSlice<uint32_t>{d, n});
    auto out = to_host(d, n);
    std::vector<uint32_t> want = v;
    std::sort(want.begin(), want.end());
    REQUIRE(out == want);
    hipFree(d);
  }
  std::printf("turbosort.hpp: all checks passed\n");
  return 0;
}
