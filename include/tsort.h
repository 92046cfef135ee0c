/*
 * tsort.h -- C ABI of the MI355X-native (gfx950) radix sort with turbosort's API.
 *
 * Every pointer named d_* is device memory on the context's GPU.  `stream` is a
 * hipStream_t passed as an opaque pointer (NULL = the default stream); all work
 * of a call is enqueued on it and the call returns without waiting, except
 * where noted.  No C++ or torch types cross this boundary.
 *
 * The reference crate (turbosort 0.2.2) is a Rust library with no FFI; each
 * entry point below names the reference function it replaces.  Rust panics in
 * the reference become status codes here, and ts_ctx_last_error() returns the
 * reference's panic message for them, word for word.
 */
#ifndef TSORT_H
#define TSORT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ts_status {
  TS_OK = 0,
  TS_ERR_BUFFER_TOO_SMALL = 1, /* dispatch.rs:78-83, radix/mod.rs:92-97 (panic) */
  TS_ERR_LENGTH_MISMATCH = 2,  /* by_key.rs:137-141 (panic) */
  TS_ERR_TOO_LONG = 3,         /* more than u32::MAX elements where an order was asked for (by_key.rs:190-194) */
  TS_ERR_INVALID_ARG = 4,
  TS_ERR_HIP = 5,              /* a HIP runtime call failed; see ts_ctx_last_error */
  TS_ERR_VERIFY_FAILED = 6,    /* an output failed sortedness or the multiset checksum */
  TS_ERR_NO_DEVICE = 7
} ts_status;

/* The ten sealed SortableKey types (key.rs:64-77), in this order. */
typedef enum ts_dtype {
  TS_U8 = 0, TS_U16 = 1, TS_U32 = 2, TS_U64 = 3,
  TS_I8 = 4, TS_I16 = 5, TS_I32 = 6, TS_I64 = 7,
  TS_F32 = 8, TS_F64 = 9
} ts_dtype;

/* A device workspace: look-back state, digit histograms and grow-only scratch.
 * One context per stream at a time; like &mut SortByKeyBuffer it is not
 * thread-safe.  Separate contexts may run concurrently on separate streams. */
typedef struct ts_ctx ts_ctx;

ts_status ts_ctx_create(ts_ctx** out, int device);
void ts_ctx_destroy(ts_ctx* ctx);
/* Grow the workspace so that sorts of n keys of `dtype` (and by-key sorts of n
 * records of `record_bytes`) allocate nothing.  Mirrors
 * SortByKeyBuffer::with_capacity (by_key.rs:74-81). */
ts_status ts_ctx_reserve(ts_ctx* ctx, ts_dtype dtype, size_t n, size_t record_bytes);
/* Bytes of device memory the context holds. */
size_t ts_ctx_workspace_bytes(const ts_ctx* ctx);
/* Message for the last failed call on this context (reference panic text). */
const char* ts_ctx_last_error(const ts_ctx* ctx);
const char* ts_status_string(ts_status s);

/* turbosort::sort (lib.rs:99-101): sort n keys of `dtype` in place, ascending
 * in the type's total order (f32/f64: IEEE-754 totalOrder, key.rs:162-198). */
ts_status ts_sort(ts_ctx* ctx, ts_dtype dtype, void* d_data, size_t n, void* stream);

/* turbosort::sort_with_buffer (lib.rs:122-124): as ts_sort with the caller's
 * scratch d_buf of buf_len elements.  buf_len < n is TS_ERR_BUFFER_TOO_SMALL
 * with "buffer too small: need {n}, got {buf_len}". */
ts_status ts_sort_with_buffer(ts_ctx* ctx, ts_dtype dtype, void* d_data, size_t n, void* d_buf,
                              size_t buf_len, void* stream);

/* No counterpart in the reference: ts_sort_with_buffer for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, checks, messages
 * and output as ts_sort_with_buffer; n is at most 2^28 keys (larger n is
 * TS_ERR_INVALID_ARG).  Nothing is allocated per call: the sort's
 * bookkeeping and look-back status live in one file-scope device array of
 * about 30 MB that each GPU the library runs on holds until the process
 * ends (its address is looked up in ts_ctx_create).  Every call on one GPU
 * shares that array, whatever the context: calls that overlap on one GPU
 * race and give wrong results; serialize them (one stream, or events). */
ts_status ts_sort_with_buffer_static(ts_ctx* ctx, ts_dtype dtype, void* d_data, size_t n, void* d_buf,
                                     size_t buf_len, void* stream);

/* turbosort::sort (lib.rs:99-101) on a slice in host memory: copies it to the device
 * (in 64 MiB pieces; a pinned slice streams at the link rate, a pageable one
 * through the runtime's staging), sorts it there and copies it back, and
 * returns once the host slice is sorted.  The device copy is a grow-only
 * buffer of the context, so repeated calls allocate nothing once warm. */
ts_status ts_sort_host(ts_ctx* ctx, ts_dtype dtype, void* host_data, size_t n, void* stream);

/* turbosort::sort_by_key_with_buffer (lib.rs:196-202, by_key.rs:175-235) for
 * POD records: record i starts at d_records + i*stride and its key, of type
 * key_type, at byte key_offset.  Writes d_order_out[i] = original position of
 * the record that sorts to i (SortByKeyBuffer::order, by_key.rs:108-110), the
 * stable order.  If gather_records != 0 the records are also moved into that
 * order in place (move_into_order, by_key.rs:315-334); 8- and 16-byte records
 * aligned to their size move through the sort's passes instead of being
 * gathered afterwards.  stride == key width and key_offset == 0 sorts a plain
 * key array by itself and yields its stable argsort.
 * n > u32::MAX (by_key.rs:190-194): no order is written (d_order_out is not
 * used and may be NULL) and the records are sorted stably in place, as
 * ts_sort_by_key_unordered does; with gather_records == 0 that is
 * TS_ERR_TOO_LONG. */
ts_status ts_sort_by_key(ts_ctx* ctx, ts_dtype key_type, void* d_records, size_t n, size_t stride,
                         size_t key_offset, uint32_t* d_order_out, int gather_records, void* stream);

/* The by-key sort beyond u32::MAX elements (by_key.rs:190-194: order()
 * cleared, the slice sorted stably by key), for any n: the n records of
 * record_bytes bytes at d_records are put in stable order of the keys of
 * type key_type at d_keys + i*key_stride + key_offset, which may lie inside
 * the records (d_keys == d_records, key_stride == record_bytes) or in an
 * array of their own (the keys a key closure gave, evaluated once).  No order
 * is produced.  The device memory it takes grows with n: 16-byte (key,
 * position) records twice over plus one copy of the records, less when the
 * records are the keys themselves or 8- or 16-byte records that hold them. */
ts_status ts_sort_by_key_unordered(ts_ctx* ctx, ts_dtype key_type, const void* d_keys, size_t key_stride,
                                   size_t key_offset, void* d_records, size_t n, size_t record_bytes,
                                   void* stream);

/* SortByKeyBuffer::apply_order (by_key.rs:136-144): d_other[i] <- d_other[order[i]]
 * for n elements of elem_bytes bytes, in place.  order_len != n is
 * TS_ERR_LENGTH_MISMATCH with the reference's message. */
ts_status ts_apply_order(ts_ctx* ctx, const uint32_t* d_order, size_t order_len, void* d_other, size_t n,
                         size_t elem_bytes, void* stream);

/* Batched segmented sort.  No single reference function: per segment it is
 * sort_by_key_with_buffer keyed by the element itself (by_key.rs:175-235)
 * followed by apply_order on the values (by_key.rs:136-144), the GPU form of
 * chunks_exact_mut(size).for_each(turbosort::sort) (benches/sort_bench.rs:229-250).
 * Every segment [d_offsets[s], d_offsets[s+1]) of the n keys of `dtype` is
 * sorted stably in place; the u32 d_vals (optional, may be NULL) move with
 * their keys.  Elements outside every segment are left alone.
 * max_segment_len bounds every segment's length, or is 0 when unknown.  When
 * it is 0 or above ts_segmented_lds_max_len, the call waits for the
 * device: segments too long for one workgroup go through the device radix
 * sort (with values: by key, stably; keys only: after three-way splits
 * around pivot keys), and invalid offsets (decreasing, or past n) return
 * TS_ERR_INVALID_ARG.  Otherwise the call is asynchronous and can be captured
 * in a graph; the caller guarantees the bound and valid offsets. */
ts_status ts_segmented_sort_pairs(ts_ctx* ctx, ts_dtype dtype, void* d_keys, uint32_t* d_vals, size_t n,
                                  const uint64_t* d_offsets, size_t n_segments, size_t max_segment_len,
                                  void* stream);

/* The longest segment sorted inside one workgroup for this key type (with or
 * without values); longer segments take the device radix sort. */
size_t ts_segmented_lds_max_len(ts_dtype dtype, int with_values);

/* Autotuning hooks.  Size class c (0 .. 8: segments of up to 16, 32, 64, 256,
 * 1024, 2048, 4096, 8192 and 16384 keys; the last class holds 17408 8-byte
 * keys without values and 11264 with values) can
 * run one of ts_segmented_variants(...) kernel shapes, named by
 * ts_segmented_variant_name(...).  A new context starts from the shapes
 * measured fastest on MI355X (variant 0 except: for 4-byte keys with values
 * in classes 3 .. 8, one MSD pass into one bin per key slot with stable ranks
 * inside the bins; the MSD shape for 4-byte keys without values in class 8
 * and its pipelined form for 8-byte keys without values there).
 * The choice is kept per context and per key width
 * (with / without values); the learned sort's bucket sorts use the
 * keys-only choice. */
int ts_segmented_variants(ts_dtype dtype, int with_values, int size_class);
const char* ts_segmented_variant_name(ts_dtype dtype, int with_values, int size_class, int variant);
ts_status ts_segmented_set_variant(ts_ctx* ctx, ts_dtype dtype, int with_values, int size_class, int variant);
/* The variant a context runs for that size class (-1: no such class). */
int ts_segmented_get_variant(const ts_ctx* ctx, ts_dtype dtype, int with_values, int size_class);

/* Learned partitioner.  No reference counterpart: the output equals
 * ts_sort's (turbosort::sort, lib.rs:99-101) for 4- and 8-byte key types,
 * reached by bucketing the keys with a monotone model fitted on the device
 * to a sample the device radix sort has ranked, then sorting each bucket.
 * Buckets too long for one workgroup are split three ways around a pivot
 * key until the parts fit (equal keys are equal bits), with the device radix
 * sort as the last resort.  Waits for the device (bucket overflow).  cfg may
 * be NULL (defaults); report may be NULL.  Faster than ts_sort for 8-byte
 * keys (two partition passes instead of eight); for 4-byte keys ts_sort's
 * four passes are faster. */
typedef struct ts_learned_cfg {
  uint32_t buckets_log2;     /* number of buckets = 2^buckets_log2 (1..24); default 16 */
  uint32_t sample_per_mille; /* sample size as a fraction of n; default 5 (0.5 %) */
  uint32_t train;            /* 1: fit the model (start from the sample's interval histogram, then Adam on the
                                mean squared CDF error); 0: uniform knots (the ablation); default 1 */
  uint32_t iters;            /* optimiser steps; default 40 */
  float lr;                  /* Adam step size; default 0.02 */
  uint32_t points;           /* training points taken from the sorted sample; default 16384 */
  uint32_t verify;           /* 1: judge the output (see below); default 1 */
} ts_learned_cfg;

typedef struct ts_learned_report {
  uint64_t sample;           /* sample size */
  uint32_t buckets;
  uint32_t overflow_buckets; /* buckets longer than one workgroup sorts */
  uint32_t split_rounds;     /* three-way split rounds those buckets took */
  uint64_t overflow_keys;    /* keys in those buckets */
  uint64_t largest_bucket;   /* keys in the largest bucket */
  float loss_before;         /* mean squared CDF error on the training points: the histogram start */
  float loss_after;          /* ... after the optimiser steps (the model used) */
  float loss_uniform;        /* ... of uniform knots (the untrained model) */
  uint32_t verified;         /* 1: the judge ran and passed */
} ts_learned_report;

/* With cfg->verify (the default) ts_learned_sort judges its own output before
 * returning: the input's multiset fingerprint is taken on the device before
 * the partition, and the output must hold the same fingerprint with no
 * adjacent pair out of order (ts_verify_sorted); otherwise the call returns
 * TS_ERR_VERIFY_FAILED.  Device errors recorded during the sort (see
 * ts_ctx_check) also fail it. */
void ts_learned_cfg_default(ts_learned_cfg* cfg);
ts_status ts_learned_sort(ts_ctx* ctx, ts_dtype dtype, void* d_data, size_t n, const ts_learned_cfg* cfg,
                          ts_learned_report* report, void* stream);

/* One fit of the learned model's piecewise-linear CDF (the training kernel of
 * ts_learned_sort): T points (u in knot units [0, 1024), target t), 1024
 * parameters theta (in/out), the gradient of the mean squared error at the
 * initial theta, the loss before and after `iters` Adam steps. */
ts_status ts_pwl_fit(ts_ctx* ctx, const float* d_u, const float* d_t, size_t T, float* d_theta, float* d_grad,
                     float* d_loss2, int iters, float lr, void* stream);

/* ---- SortByKeyBuffer (by_key.rs:51-159) as an object ----
 * A context plus a grow-only device order array.  ts_sbk_sort is
 * sort_by_key_with_buffer (lib.rs:196-202); ts_sbk_order is order()
 * (by_key.rs:108-110), empty (len 0) before the first sort; ts_sbk_apply_order
 * is apply_order (by_key.rs:136-144); ts_sbk_capacity is capacity()
 * (by_key.rs:84-90). */
typedef struct ts_sbk ts_sbk;
ts_status ts_sbk_create(ts_sbk** out, int device);
void ts_sbk_destroy(ts_sbk* b);
size_t ts_sbk_capacity(const ts_sbk* b);
const uint32_t* ts_sbk_order(const ts_sbk* b, size_t* len);
ts_status ts_sbk_sort(ts_sbk* b, ts_dtype key_type, void* d_records, size_t n, size_t stride, size_t key_offset,
                      int gather_records, void* stream);
ts_status ts_sbk_apply_order(ts_sbk* b, void* d_other, size_t n, size_t elem_bytes, void* stream);
const char* ts_sbk_last_error(const ts_sbk* b);

/* ---- Utilities for tests and benchmarks (no reference counterpart) ---- */

/* Fill d_out with n synthetic keys from SplitMix64 (Vigna), element i taking
 * output i of the stream with state `seed`:
 *   kind 0: u32 = high 32 bits;  kind 1: u64 = the output;
 *   kind 2: u64 = zero-extended high 32 bits. */
ts_status ts_fill_splitmix64(ts_ctx* ctx, void* d_out, size_t n, uint64_t seed, int kind, void* stream);

/* The output judge: TS_OK when no adjacent pair of the n keys at d_data is
 * out of order in the type's total order and, if expect_sum_xor is not NULL,
 * their multiset fingerprint (as ts_checksum computes it) equals
 * expect_sum_xor; else TS_ERR_VERIFY_FAILED, with ts_ctx_last_error saying
 * which.  Waits for the device.  The closest reference check is the
 * adjacent-pair order scan of presorted.rs:35-65 (all_pairs); the
 * fingerprint has no counterpart there. */
ts_status ts_verify_sorted(ts_ctx* ctx, ts_dtype dtype, const void* d_data, size_t n,
                           const uint64_t* expect_sum_xor, void* stream);

/* Order-independent multiset checksum of n keys of `dtype` (sum and xor of a
 * 64-bit mix of each key's bits) and the count of adjacent pairs out of order
 * in the type's total order.  Synchronous: results land in host memory. */
ts_status ts_checksum(ts_ctx* ctx, ts_dtype dtype, const void* d_data, size_t n, uint64_t out_sum_xor[2],
                      uint64_t* out_unsorted_pairs, void* stream);

/* Per-kernel timing of the last sort on this context (enabled with
 * ts_ctx_set_profiling).  Fills up to `cap` (name, milliseconds, bytes moved
 * algorithmically) triples; returns how many kernels were recorded.  The
 * timings come from HIP events on the sort's own stream. */
ts_status ts_ctx_set_profiling(ts_ctx* ctx, int enable);
int ts_ctx_kernel_times(const ts_ctx* ctx, const char** names, float* ms, double* bytes, int cap);

/* How the scatter passes rank a wave's keys by digit (no counterpart in the
 * reference, whose scatter is sequential: scatter.rs:26-40).  Both give the
 * same stable order.  TS_RANK_ATOMIC uses LDS atomic adds and relies on the
 * adds of one wave instruction to one address taking effect in lane order;
 * ts_ctx_create checks that once per device and selects it when it holds,
 * else TS_RANK_BALLOT (eight wave ballots per key).  Setting TS_RANK_ATOMIC
 * on a device that failed the check returns TS_ERR_INVALID_ARG. */
enum { TS_RANK_BALLOT = 0, TS_RANK_ATOMIC = 1 };
ts_status ts_ctx_set_ranking(ts_ctx* ctx, int method);
int ts_ctx_ranking(const ts_ctx* ctx);

/* Keys-only segments longer than one workgroup holds (segmented sort without
 * values, the learned sort's overflowing buckets) go through rounds of
 * three-way splits; what is still too long after this many rounds (default
 * 12, 0..64) takes the device radix sort.  No counterpart in the reference.
 * Any value gives the same output. */
ts_status ts_ctx_set_split_rounds(ts_ctx* ctx, int rounds);
int ts_ctx_split_rounds(const ts_ctx* ctx);

/* Sorts of 8-byte keys without values (ts_sort, ts_sort_with_buffer,
 * ts_sort_host) of at least this many keys (default 2^23; SIZE_MAX: never)
 * divert as the reference's byte-digit core does (radix/mod.rs:240-242,
 * 268-270; divert.rs:38-90): only the top live digits that carry about
 * log2(n) + 2 bits are sorted, decided on the device from the histograms,
 * and one finishing read sorts each run of keys that share them.  A run
 * longer than the finish sorts that is out of order makes the device sort
 * every digit after all.  Any value gives the same output; sizes at or
 * below 2^21 keys never divert. */
ts_status ts_ctx_set_divert_min(ts_ctx* ctx, size_t n);
size_t ts_ctx_divert_min(const ts_ctx* ctx);

/* Device-side checks.  Kernels record what they cannot return in a sticky
 * error word of the context: a look-back wait that hit its bound (the sort's
 * output is wrong; never expected), a segment longer than the max_segment_len
 * given to an asynchronous ts_segmented_sort_pairs (left unsorted), and
 * invalid segment offsets in such a call.  ts_ctx_check waits for `stream`,
 * reads and clears the word, and returns TS_ERR_VERIFY_FAILED (with
 * ts_ctx_last_error naming each cause) when it was set.  ts_sort_host,
 * ts_learned_sort and the waiting forms of ts_segmented_sort_pairs check it
 * themselves before returning. */
ts_status ts_ctx_check(ts_ctx* ctx, void* stream);

/* What the last radix sort on this context did (waits for `stream`): live
 * digits, scatter passes executed (by the second sort, when it ran), digits
 * the divert left to its finish and the bit the finish started from (0 after
 * a second sort), whether the second sort ran, and whether non-increasing
 * input was reversed instead of sorted. */
typedef struct ts_sort_info {
  uint32_t live_digits;
  uint32_t executed_passes;
  uint32_t skipped_digits;
  uint32_t finish_shift;
  uint32_t fallback;
  uint32_t reversed;
} ts_sort_info;
ts_status ts_ctx_last_sort_info(ts_ctx* ctx, ts_sort_info* info, void* stream);

/* Workspace planning, host arithmetic only (touches no device): bytes of
 * look-back status one radix sort of n keys of `dtype` uses with scatter
 * tiles of `kind` (0 keys only, 1 keys with u32 order values, 2 learned
 * bucket digits, 3 / 4 carried 8- / 16-byte records with the order, 5 / 6
 * carried 8- / 16-byte records without it), or, for kind -1, the bytes a
 * context reserves so that every radix sort of at most n keys of that width
 * and of the kinds 0-2 fits. */
size_t ts_plan_status_bytes(ts_dtype dtype, size_t n, int kind);

#ifdef __cplusplus
}
#endif

#endif /* TSORT_H */
