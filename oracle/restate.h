/*
 * CPU restatement of turbosort's sort path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it.  The product path (libtsort.so) never
 * links or calls it.
 *
 * Every function follows a file of the reference crate (turbosort 0.2.2); the
 * citation is given beside each declaration.  Outputs of the sorting functions
 * are additionally pinned by a closed form (ascending order of the radix key,
 * stable argsort for the order), which oracle/closed_form.py computes
 * independently with numpy.
 */
#ifndef TS_ORACLE_RESTATE_H
#define TS_ORACLE_RESTATE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same numbering as ts_dtype in include/tsort.h. */
enum { TSO_U8 = 0, TSO_U16, TSO_U32, TSO_U64, TSO_I8, TSO_I16, TSO_I32, TSO_I64, TSO_F32, TSO_F64 };

enum { TSO_OK = 0, TSO_ERR_BUFFER_TOO_SMALL = 1, TSO_ERR_LENGTH_MISMATCH = 2, TSO_ERR_TOO_LONG = 3,
       TSO_ERR_INVALID_ARG = 4 };

/* turbosort::sort (lib.rs:99 -> dispatch.rs:31-67 -> radix/mod.rs:61-74). */
int tso_sort(int dtype, void* data, size_t n);

/* turbosort::sort_with_buffer (lib.rs:122 -> dispatch.rs:77-111, radix/mod.rs:91-108).
 * Returns TSO_ERR_BUFFER_TOO_SMALL where the reference panics (dispatch.rs:78-83). */
int tso_sort_with_buffer(int dtype, void* data, size_t n, void* buf, size_t buf_len);

/* turbosort::sort_parallel (lib.rs:224 -> parallel.rs:46-291), with `threads`
 * std::threads in place of rayon's pool. */
int tso_sort_parallel(int dtype, void* data, size_t n, int threads);

/* The order() of sort_by_key_with_buffer (by_key.rs:175-235, 239-310) for keys
 * of type `key_dtype` given in their original type: order[i] is the original
 * position of the element that lands at i.  Returns TSO_ERR_TOO_LONG (order
 * left untouched) for n > u32::MAX, where the reference leaves order() empty
 * (by_key.rs:190-194). */
int tso_sort_by_key_order(int key_dtype, const void* keys, size_t n, uint32_t* order);

/* move_into_order / apply_order (by_key.rs:136-144, 315-334):
 * data[i] <- data[order[i]] for elements of elem_bytes bytes. */
int tso_apply_order(const uint32_t* order, size_t order_len, void* data, size_t n, size_t elem_bytes);

/* Every segment [offsets[s], offsets[s+1]) sorted stably by key with its
 * u32 values (vals may be NULL): per segment the order of
 * sort_by_key_with_buffer and the move into it (by_key.rs:175-235, 315-334),
 * the reference's chunks-and-sort usage (benches/sort_bench.rs:229-250);
 * `threads` std::threads take the segments in turn. */
int tso_segmented_sort_pairs(int dtype, void* keys, uint32_t* vals, const uint64_t* offsets, size_t n_segments,
                             int threads);

/* Internal decisions, exported so the reference's own KATs can pin them. */
uint32_t tso_log2_eighths(uint64_t x);                                       /* divert.rs:95-104 */
size_t tso_skippable_digits(size_t n_live, size_t len, const uint64_t* max_bucket); /* divert.rs:38-54 */
/* histogram.rs:55-68 on radix keys widened to u64; returns 1 and writes *out for
 * Some(bits), 0 for None. key_bytes is the key width (1, 2, 4 or 8). */
int tso_varying_bits(const uint64_t* keys, size_t n, int key_bytes, uint64_t* out);
/* divert.rs:60-81 on u16 keys (the KAT at divert.rs:151-164). */
void tso_finish_runs_u16(uint16_t* keys, size_t n, uint32_t shift);

/* SplitMix64 (Vigna): out[i] = mix(seed + (i+1) * 0x9E3779B97F4A7C15). */
void tso_splitmix64_fill(uint64_t seed, size_t n, uint64_t* out);

#ifdef __cplusplus
}
#endif

#endif
