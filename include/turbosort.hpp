// turbosort.hpp -- C++ host API over the C ABI (tsort.h), with the names,
// argument meaning and error behaviour of the reference crate's public API
// (turbosort 0.2.2, lib.rs:99-229 and by_key.rs:51-159), for device-resident
// slices on one AMD Instinct MI355X:
//
//   turbosort::sort(slice)                                  lib.rs:99
//   turbosort::sort_with_buffer(slice, buffer)              lib.rs:122
//   turbosort::sort_by_key(records, KeyField<K>{offset})    lib.rs:152
//   turbosort::sort_by_key_with_buffer(records, buf, key)   lib.rs:196
//   turbosort::SortByKeyBuffer: capacity, order, apply_order  by_key.rs:51-159
//
// A DeviceSlice<T> is a (pointer, length) pair into device memory; `stream` is
// a hipStream_t passed as void* (nullptr: the default stream).  Where the
// reference panics, these functions throw turbosort::Panic with the
// reference's message; other failures throw std::runtime_error or
// std::invalid_argument.  Only tsort.h is needed to compile against it.
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "tsort.h"

namespace turbosort {

struct Panic : std::logic_error {
  using std::logic_error::logic_error;
};

template <class T> struct DeviceSlice {
  T* ptr = nullptr;
  size_t len = 0;
  size_t size() const { return len; }
};

// The sealed SortableKey set (key.rs:64-77).
template <class T> struct SortableKey;
template <> struct SortableKey<uint8_t> { static constexpr ts_dtype dtype = TS_U8; };
template <> struct SortableKey<uint16_t> { static constexpr ts_dtype dtype = TS_U16; };
template <> struct SortableKey<uint32_t> { static constexpr ts_dtype dtype = TS_U32; };
template <> struct SortableKey<uint64_t> { static constexpr ts_dtype dtype = TS_U64; };
template <> struct SortableKey<int8_t> { static constexpr ts_dtype dtype = TS_I8; };
template <> struct SortableKey<int16_t> { static constexpr ts_dtype dtype = TS_I16; };
template <> struct SortableKey<int32_t> { static constexpr ts_dtype dtype = TS_I32; };
template <> struct SortableKey<int64_t> { static constexpr ts_dtype dtype = TS_I64; };
template <> struct SortableKey<float> { static constexpr ts_dtype dtype = TS_F32; };
template <> struct SortableKey<double> { static constexpr ts_dtype dtype = TS_F64; };

inline void throw_for(ts_status s, const char* msg) {
  if (s == TS_OK) return;
  if (s == TS_ERR_BUFFER_TOO_SMALL || s == TS_ERR_LENGTH_MISMATCH) throw Panic(msg);
  if (s == TS_ERR_INVALID_ARG) throw std::invalid_argument(msg);
  throw std::runtime_error(std::string(ts_status_string(s)) + ": " + msg);
}

// A device workspace (ts_ctx); grow-only, one per thread and device by default.
class Context {
 public:
  explicit Context(int device = 0) {
    ts_status s = ts_ctx_create(&ctx_, device);
    if (s != TS_OK) throw std::runtime_error(std::string("ts_ctx_create: ") + ts_status_string(s));
  }
  ~Context() { ts_ctx_destroy(ctx_); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  ts_ctx* get() const { return ctx_; }
  void check(ts_status s) const { throw_for(s, ts_ctx_last_error(ctx_)); }
  static Context& default_for(int device = 0) {
    static thread_local Context* ctx[16] = {};
    if (device < 0 || device >= 16) throw std::invalid_argument("device index out of range");
    if (!ctx[device]) ctx[device] = new Context(device);
    return *ctx[device];
  }

 private:
  ts_ctx* ctx_ = nullptr;
};

// lib.rs:99-101.
template <class T> void sort(DeviceSlice<T> slice, void* stream = nullptr, Context& ctx = Context::default_for()) {
  ctx.check(ts_sort(ctx.get(), SortableKey<T>::dtype, slice.ptr, slice.len, stream));
}

// lib.rs:122-124: throws Panic("buffer too small: need N, got M").
template <class T>
void sort_with_buffer(DeviceSlice<T> slice, DeviceSlice<T> buffer, void* stream = nullptr,
                      Context& ctx = Context::default_for()) {
  ctx.check(ts_sort_with_buffer(ctx.get(), SortableKey<T>::dtype, slice.ptr, slice.len, buffer.ptr, buffer.len,
                                stream));
}

// A key stored inside a POD record at byte `offset`: the device form of a key
// closure `|r| r.field`, read once per record in index order (by_key.rs:197-200).
template <class K> struct KeyField {
  size_t offset = 0;
};

// by_key.rs:51-159.
class SortByKeyBuffer {
 public:
  explicit SortByKeyBuffer(int device = 0) {
    ts_status s = ts_sbk_create(&b_, device);
    if (s != TS_OK) throw std::runtime_error(std::string("ts_sbk_create: ") + ts_status_string(s));
  }
  ~SortByKeyBuffer() { ts_sbk_destroy(b_); }
  SortByKeyBuffer(const SortByKeyBuffer&) = delete;
  SortByKeyBuffer& operator=(const SortByKeyBuffer&) = delete;

  size_t capacity() const { return ts_sbk_capacity(b_); }
  // order()[i] is the original position of the element now at i; empty
  // before the first sort.
  DeviceSlice<const uint32_t> order() const {
    size_t len = 0;
    const uint32_t* p = ts_sbk_order(b_, &len);
    return {p, len};
  }
  // by_key.rs:136-144: throws Panic when the lengths differ.
  template <class U> void apply_order(DeviceSlice<U> other, void* stream = nullptr) {
    throw_for(ts_sbk_apply_order(b_, other.ptr, other.len, sizeof(U), stream), ts_sbk_last_error(b_));
  }
  // Sorts a key array by itself (key = the element).
  template <class K> void sort_keys(DeviceSlice<K> keys, void* stream) {
    throw_for(ts_sbk_sort(b_, SortableKey<K>::dtype, keys.ptr, keys.len, sizeof(K), 0, 1, stream),
              ts_sbk_last_error(b_));
  }
  template <class R, class K> void sort_records(DeviceSlice<R> records, KeyField<K> key, void* stream) {
    static_assert(std::is_trivially_copyable<R>::value, "records move bitwise on the device");
    throw_for(ts_sbk_sort(b_, SortableKey<K>::dtype, records.ptr, records.len, sizeof(R), key.offset, 1, stream),
              ts_sbk_last_error(b_));
  }

 private:
  ts_sbk* b_ = nullptr;
};

// lib.rs:196-202 for POD records keyed by a field.
template <class R, class K>
void sort_by_key_with_buffer(DeviceSlice<R> records, SortByKeyBuffer& buffer, KeyField<K> key,
                             void* stream = nullptr) {
  buffer.sort_records(records, key, stream);
}

// lib.rs:152-158: a fresh buffer per call.  Its device memory is released when
// it goes out of scope (releasing device memory waits for the sort).
template <class R, class K> void sort_by_key(DeviceSlice<R> records, KeyField<K> key, void* stream = nullptr) {
  SortByKeyBuffer buffer;
  buffer.sort_records(records, key, stream);
}

}  // namespace turbosort
