/*
 * bolt_ops.h — operator C ABI of the MI355X execution core.
 *
 * Entry points replace the reference's executor functions for the GPU
 * relational-operator path (the dispatch arms of Engine::execute_leaf_whole,
 * src/exec/engine.rs:3898-4130).  Data crosses the boundary as device
 * column descriptors (plain pointers + lengths + a dtype code), never as
 * framework types.
 *
 * Conventions (kept from the reference, src/error.rs and
 * src/exec/engine.rs:3915-3925):
 *   - every function returns an int status; BOLT_OK = 0;
 *   - BOLT_DECLINE is the GpuCapacity analogue: the operator declines the
 *     input (e.g. a build side beyond the table cap) and the caller reruns it
 *     on its host path;
 *   - BOLT_ERR_TYPE carries type errors, including the integer-SUM overflow
 *     error ("SUM(integer) overflow: accumulator exceeds i64 range",
 *     src/exec/aggregate.rs:2120);
 *   - the message of the last failure on the calling thread is returned by
 *     bolt_last_error() (declared in bolt_hip.h);
 *   - when no GPU is usable every entry point returns BOLT_ERR_NO_DEVICE
 *     (100000), the stub contract of tests/cuda_stub_contract_test.rs:40-60.
 *
 * Unlike the reference (src/exec/launch.rs:353,409), entry points do not
 * synchronise after each launch: work is queued on the caller's stream and
 * only the calls that must return a host-visible count or scalar wait for
 * that stream.  Device outputs are caller-allocated unless a function says it
 * returns a library-owned table (freed with the matching *_release).
 */
#ifndef BOLT_OPS_H
#define BOLT_OPS_H

#include <stdint.h>

#include "bolt_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (src/error.rs BoltError variants) ---- */
#define BOLT_OK 0
#define BOLT_ERR_PLAN 1        /* BoltError::Plan */
#define BOLT_ERR_TYPE 2        /* BoltError::Type */
#define BOLT_ERR_MEMORY 3      /* BoltError::Memory */
#define BOLT_DECLINE 4         /* BoltError::GpuCapacity: rerun on the host */
#define BOLT_ERR_UNSUPPORTED 5 /* BoltError::Unsupported */
#define BOLT_ERR_OTHER 6       /* BoltError::Other */
#define BOLT_ERR_DEVICE 7      /* BoltError::CudaWithCode: a HIP call failed */
#define BOLT_ERR_NO_DEVICE 100000

/* ---- column dtypes (src/plan/logical_plan.rs:57 DataType) ---- */
enum bolt_dtype {
  BOLT_BOOL = 1,   /* one byte per value */
  BOLT_I32 = 2,
  BOLT_I64 = 3,
  BOLT_F32 = 4,
  BOLT_F64 = 5,
  BOLT_DEC128 = 6, /* 16-byte little-endian two's complement, scale in `scale` */
  BOLT_DATE32 = 7, /* days since the epoch, int32 */
  BOLT_TIMESTAMP = 8 /* ticks since the epoch, int64; `scale` is the unit as digits of a
                        second: 0 s, 3 ms, 6 us, 9 ns (TimeUnit, src/plan/logical_plan.rs:7) */
};

/* A device column: Arrow layout (values + optional LSB-first validity
 * bitmap, NULL = no nulls).  `data` and `validity` are device pointers. */
typedef struct bolt_column {
  void* data;
  const uint8_t* validity;
  int64_t length;
  int32_t dtype;
  int32_t scale; /* Decimal128 scale; precision is not needed by any kernel */
} bolt_column_t;

/* ---- aggregates (src/plan/logical_plan.rs:2056 AggregateExpr) ---- */
enum bolt_agg_op {
  BOLT_AGG_SUM = 1,
  BOLT_AGG_COUNT = 2,      /* COUNT(col): non-NULL values */
  BOLT_AGG_COUNT_STAR = 3, /* COUNT(*): rows */
  BOLT_AGG_MIN = 4,
  BOLT_AGG_MAX = 5,
  BOLT_AGG_AVG = 6,
  BOLT_AGG_VAR_POP = 7,
  BOLT_AGG_VAR_SAMP = 8,
  BOLT_AGG_STDDEV_POP = 9,
  BOLT_AGG_STDDEV_SAMP = 10
};

typedef struct bolt_agg {
  int32_t op;
  int32_t col; /* index into the operator's input columns; ignored for COUNT(*) */
} bolt_agg_t;

/* A scalar result (src/exec/aggregate.rs Scalar).  Output dtypes follow
 * sum_output_dtype / AggregateExpr::output_dtype
 * (src/plan/logical_plan.rs:2126,2191): SUM(Int32) -> Int64,
 * SUM(Decimal128(p,s)) -> Decimal128(38,s), COUNT -> Int64,
 * AVG/VAR/STDDEV -> Float64, MIN/MAX keep the input type. */
typedef struct bolt_scalar {
  int32_t dtype;
  int32_t is_null; /* SQL NULL: SUM/MIN/MAX/AVG over no values */
  int64_t i64;     /* integer / date result, decimal low word */
  double f64;      /* float result (Float32 results are widened) */
  int64_t hi;      /* decimal high word */
} bolt_scalar_t;

/* ---- predicates: the filter form of a KernelSpec (src/plan/physical_plan.rs:537)
 *      restricted to an AND of column-vs-literal comparisons.  A comparison
 *      with a NULL value is false (the row is dropped), as in the reference's
 *      predicate kernel (src/jit/scan_kernel.rs:265). ---- */
enum bolt_cmp_op {
  BOLT_CMP_LT = 1, BOLT_CMP_LE = 2, BOLT_CMP_GT = 3, BOLT_CMP_GE = 4,
  BOLT_CMP_EQ = 5, BOLT_CMP_NE = 6, BOLT_CMP_IS_NULL = 7, BOLT_CMP_IS_NOT_NULL = 8
};

typedef struct bolt_pred_term {
  int32_t col;     /* index into the operator's input columns */
  int32_t op;      /* bolt_cmp_op */
  int64_t lit_i64; /* literal for integer / date columns; decimal low word */
  double lit_f64;  /* literal for float columns */
  int64_t lit_hi;  /* decimal high word */
} bolt_pred_term_t;

#define BOLT_MAX_PRED_TERMS 8

typedef struct bolt_filter {
  const bolt_pred_term_t* terms; /* host array */
  int32_t n_terms;               /* 0 = keep every row */
} bolt_filter_t;

/* ------------------------------------------------------------------------
 * Scalar aggregates — replaces execute_aggregate (src/exec/aggregate.rs:154)
 * and agg_with_pre::try_execute_resident (src/exec/agg_with_pre.rs:157).
 * Computes n_aggs aggregates over the rows of `inputs` that pass `filter`
 * (NULL filter = all rows) in one pass, and writes `results` (host array of
 * n_aggs).  Synchronises the stream.  Integer SUM overflow -> BOLT_ERR_TYPE,
 * decided as the reference's checked_add fold decides it: as soon as a
 * running total of the passing non-NULL values, in row order, leaves i64
 * (src/exec/aggregate.rs:2120-2140), even when the total fits.
 * ------------------------------------------------------------------------ */
int bolt_aggregate(bolt_stream_t stream, const bolt_column_t* inputs, int32_t n_inputs,
                   const bolt_agg_t* aggs, int32_t n_aggs, const bolt_filter_t* filter,
                   bolt_scalar_t* results);

/* Allocation-free form of bolt_aggregate.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and errors as bolt_aggregate, computed with the
 * precompiled kernels (no hipRTC module is loaded) in a fixed 1 MiB device
 * scratch array per GPU, so it calls no allocator.  That array holds every
 * call's intermediates: calls that overlap on one GPU race and must be
 * serialised by the caller.  An input needing more scratch than the array
 * holds returns BOLT_ERR_MEMORY (no input does with at most 16 aggregates). */
int bolt_aggregate_noalloc(bolt_stream_t stream, const bolt_column_t* inputs, int32_t n_inputs,
                           const bolt_agg_t* aggs, int32_t n_aggs, const bolt_filter_t* filter,
                           bolt_scalar_t* results);

/* ------------------------------------------------------------------------
 * Filter + stream compaction — replaces launch_predicate_kernel
 * (src/exec/compact.rs:183) + compact_columns_on_gpu
 * (src/exec/gpu_compact.rs:1005).  Keeps the rows passing `filter` AND
 * (if `mask` is non-NULL) with mask[i] != 0, in input order.  `out` holds
 * n_cols caller-allocated device columns whose `length` is their capacity
 * (their `validity` must be a writable, 4-byte aligned bitmap when the input
 * column has one); n_rows always suffices.  Writes the survivor count to
 * *n_out (host) and synchronises the stream.  If more rows survive than the
 * smallest capacity, nothing beyond it is written, *n_out is the full count
 * and BOLT_DECLINE is returned (the capacity contract of bolt_join_probe).
 * Single pass: predicate, decoupled look-back scan and gather in one kernel.
 * ------------------------------------------------------------------------ */
int bolt_filter_compact(bolt_stream_t stream, const bolt_column_t* cols, int32_t n_cols,
                        const bolt_filter_t* filter, const uint8_t* mask, bolt_column_t* out,
                        int64_t* n_out);

/* Exclusive prefix sum of a u8 mask into u32 offsets, plus the total
 * (prefix_scan_mask, src/exec/gpu_compact.rs:415).  n < 2^32. */
int bolt_prefix_scan_mask(bolt_stream_t stream, const uint8_t* mask, int64_t n, uint32_t* offsets,
                          int64_t* total);

/* Gather by row ids — the take that follows a join (the host arrow-select
 * take of execute_inner_join's callers, src/exec/join.rs:169, and the device
 * gather of src/jit/prefix_scan.rs:1227): dst[i] = src[rows[i]] for i < n,
 * validity bits likewise.  dst is caller-allocated with capacity n (and a
 * writable bitmap of ceil(n / 8) bytes when src has one or a row id is
 * BOLT_NULL_ROW); it takes src's
 * dtype and scale and length n.  A row id of BOLT_NULL_ROW gathers a NULL
 * (the NULL index of an outer join's take, src/exec/join.rs:754; dst then
 * needs a bitmap).  Any other row id >= src->length is an error
 * (BOLT_ERR_PLAN; that row gathers zero).  Synchronises. */
int bolt_take(bolt_stream_t stream, const bolt_column_t* src, const uint32_t* rows, int64_t n,
              bolt_column_t* dst);

/* ------------------------------------------------------------------------
 * Hash GROUP BY — replaces execute_groupby (src/exec/groupby.rs:295) and its
 * Tier-1 / Tier-2 / classic cascade.  One or two key columns (Int32 / Int64 /
 * Float32 / Float64 / Date32; two only when both are 32-bit, the packing of
 * src/exec/groupby_common.rs:352), or one Decimal128 key, grouped by its exact
 * value (key_bit_width rejects it, src/exec/groupby_common.rs:117; here values
 * that fit 64 bits group on the GPU and a wider value returns BOLT_DECLINE),
 * any number of aggregates.  A single-column
 * key with NULLs forms one NULL group (src/exec/groupby.rs:528-606).
 * Rows that fail `filter` are skipped.  The result is a library-owned table:
 * n_keys key columns then one column per aggregate, groups in ascending
 * packed-key order (the reference's output order, groupby.rs:800) when
 * flags has BOLT_GROUPBY_SORTED, otherwise in hash order.
 * ------------------------------------------------------------------------ */
#define BOLT_GROUPBY_SORTED 1u
#define BOLT_GROUPBY_FORCE_GLOBAL 2u /* skip the LDS tier (tests / autotuner) */
#define BOLT_GROUPBY_FORCE_PARTITIONED 4u /* force the partitioned HBM tier */

typedef struct bolt_table bolt_table_t;

int bolt_groupby(bolt_stream_t stream, const bolt_column_t* keys, int32_t n_keys,
                 const bolt_column_t* inputs, int32_t n_inputs, const bolt_agg_t* aggs,
                 int32_t n_aggs, const bolt_filter_t* filter, uint32_t flags, bolt_table_t** out);
/* How the calling thread's last bolt_groupby ran (no reference counterpart;
 * the reference logs its tier choice, src/exec/groupby.rs:355-454): tier 0 =
 * the LDS tier's partial records merged into the table, 1 = rows also spilled
 * to the global table with atomics, 2 = the radix-partitioned tier; the rows
 * the LDS tier spilled, its partial records, and the partitions (tier 2).
 * Any pointer may be NULL. */
int bolt_groupby_last_plan(int32_t* tier, int64_t* spilled_rows, int64_t* partial_records, int32_t* partitions);
/* Autotuner of the radix-partitioned tier's fan-out, the partition count
 * the reference fixes at 4096 (TIER2 partitions,
 * src/exec/groupby_tier2_orchestrator.rs:214).  Runs bolt_groupby on these
 * inputs with the partitioned tier forced, once untimed and `reps` times
 * timed, for each partition count from a quarter to four times the default
 * (partitions of about half an LDS table of groups, at most 4096), and keeps
 * the fastest for the rest of the process for inputs of this shape (the
 * log2 of the estimated group count and the accumulator words), so later
 * bolt_groupby calls use it.  Returns the chosen partition count and its
 * mean time per call (milliseconds, host clock around the synchronising
 * call).  Errors as bolt_groupby. */
int bolt_groupby_autotune_partitions(bolt_stream_t stream, const bolt_column_t* keys, int32_t n_keys,
                                     const bolt_column_t* inputs, int32_t n_inputs, const bolt_agg_t* aggs,
                                     int32_t n_aggs, int32_t reps, int32_t* best_partitions, float* best_ms);
int bolt_table_num_rows(const bolt_table_t* t, int64_t* n_rows);
int bolt_table_num_columns(const bolt_table_t* t, int32_t* n_cols);
/* Column i of a result table (device pointers owned by the table). */
int bolt_table_column(const bolt_table_t* t, int32_t i, bolt_column_t* col);
void bolt_table_release(bolt_table_t* t);

/* ------------------------------------------------------------------------
 * Hash join — replaces hash_join_indices_on_gpu_resident
 * (src/exec/gpu_join.rs:1583) with its build (src/jit/hash_join_kernel.rs:484)
 * and probe (:711) kernels, and execute_inner_join (src/exec/join.rs:169).
 * Keys are Int32 / Int64 / Date32; NULL keys never match.  The table keeps
 * every build row (duplicate keys allowed).  `carry` (optional, 4-byte
 * Int32 / Float32 / Date32 column of the build side) is stored beside each
 * key so probe-side aggregates read it without a gather.
 * A build side beyond the table cap returns BOLT_DECLINE (gpu_join.rs:378-393).
 * Streams: bolt_join_build returns with its kernels still queued on `stream`
 * and records an event after them; a probe on any other stream waits for
 * that event on the device before reading the table.  Release frees the
 * table in `stream` order: the probes below synchronise their own stream
 * before returning, so a released table is never read.
 * ------------------------------------------------------------------------ */
typedef struct bolt_join_table bolt_join_table_t;

int bolt_join_build(bolt_stream_t stream, const bolt_column_t* build_key, const bolt_column_t* carry,
                    bolt_join_table_t** out);
/* 1 if every non-NULL build key is distinct.  Synchronises. */
int bolt_join_table_unique(const bolt_join_table_t* t, int32_t* unique);
void bolt_join_table_release(bolt_join_table_t* t);

/* INNER probe emitting (probe_row, build_row) pairs for the probe rows that
 * pass `filter` (its columns index `probe_cols`, whose column `key_col` is the
 * probe key).  Pairs come out in ascending probe-row order (the build rows of
 * one probe row in hash-chain order); the row set equals the reference's,
 * whose order is unspecified (compare after canonical ordering).  If more than
 * `capacity` pairs match, nothing beyond capacity is written, *n_out is the
 * full count and BOLT_DECLINE is returned.  Synchronises. */
int bolt_join_probe(bolt_stream_t stream, const bolt_join_table_t* t, const bolt_column_t* probe_cols,
                    int32_t n_probe_cols, int32_t key_col, const bolt_filter_t* filter,
                    uint32_t* out_probe_rows, uint32_t* out_build_rows, int64_t capacity,
                    int64_t* n_out);

/* Fused filter + INNER join + aggregates: aggregates over the joined rows
 * without materialising them (the reference composes this as a join query
 * plus a scalar-aggregate query, src/plan/physical_plan.rs:3691-3712).
 * An aggregate whose `col` is BOLT_JOIN_CARRY reads the build side's carried
 * column; otherwise it indexes `probe_cols`.  Synchronises. */
#define BOLT_JOIN_CARRY (-2)
int bolt_join_probe_aggregate(bolt_stream_t stream, const bolt_join_table_t* t,
                              const bolt_column_t* probe_cols, int32_t n_probe_cols, int32_t key_col,
                              const bolt_filter_t* filter, const bolt_agg_t* aggs, int32_t n_aggs,
                              bolt_scalar_t* results);

/* execute_inner_join (src/exec/join.rs:169): the smaller side builds, the
 * other probes; pairs are returned as (left_row, right_row).  Same capacity
 * contract as bolt_join_probe. */
int bolt_inner_join(bolt_stream_t stream, const bolt_column_t* left_key, const bolt_column_t* right_key,
                    uint32_t* out_left_rows, uint32_t* out_right_rows, int64_t capacity,
                    int64_t* n_out);

/* Equi-join of any kind — replaces execute_join's equi-join paths
 * (src/exec/join.rs:132: execute_inner_join :169, execute_outer_join :283)
 * and their GPU form execute_outer_join_indices_on_gpu
 * (src/exec/gpu_join.rs:1907) with the key encodings of
 * encode_keys_for_shape (:719).  n_keys key columns per side, of equal dtypes
 * left and right: one column of Int32 / Int64 / Date32 / Float32 / Float64 /
 * Bool (floats compare -0.0 equal to +0.0 and every NaN equal to every NaN,
 * canonical_f64_bits :682), or 2 to 4 Int32 / Int64 / Date32 columns: two
 * 32-bit columns pack into one key (TwoI32, :788); other shapes (TwoI64,
 * MultiI32) are hashed and every hash match is checked against the key
 * columns on the device, so the result is exact without the reference's host
 * verify (verify_pairs_on_host).  A row with a NULL
 * in any key column matches nothing (src/exec/join.rs:1045); on a preserved
 * side it is still emitted.
 * Output: (left_row, right_row) pairs, BOLT_NULL_ROW on the padded side.
 * Build side: the smaller for INNER (join.rs:192), else the side that is not
 * preserved (the right for LEFT and FULL, the left for RIGHT, join.rs:305).
 * Order: the probe side's rows ascending, each with its matches (build rows
 * in hash-chain order) or, when preserved and unmatched, one NULL-padded
 * pair in its place; for FULL then the unmatched build rows ascending
 * (join.rs:396-449: the reference's host order up to the order of one probe
 * row's matches).  Capacity contract of bolt_join_probe.  Synchronises. */
enum bolt_join_type {
  BOLT_JOIN_INNER = 0,
  BOLT_JOIN_LEFT = 1,  /* LEFT [OUTER] */
  BOLT_JOIN_RIGHT = 2, /* RIGHT [OUTER] */
  BOLT_JOIN_FULL = 3   /* FULL [OUTER] */
};
#define BOLT_NULL_ROW 0xFFFFFFFFu /* the NULL-padded side of an outer-join pair */
int bolt_join(bolt_stream_t stream, const bolt_column_t* left_keys, const bolt_column_t* right_keys,
              int32_t n_keys, int32_t join_type, uint32_t* out_left_rows, uint32_t* out_right_rows,
              int64_t capacity, int64_t* n_out);

/* ------------------------------------------------------------------------
 * Radix sort — replaces sort_indices_on_gpu_radix{,_multi}
 * (src/exec/gpu_sort.rs:944,995) and run_radix_pipeline_i64 (:1493).
 * Stable LSD radix sort, one up-front histogram pass for all digits, then per
 * digit one onesweep pass (LDS ranking + decoupled look-back), no host round
 * trip between passes.  n < 2^30.
 * ------------------------------------------------------------------------ */

/* Sort (key, value) pairs by int64 key (ascending unless `descending`).
 * keys_out / vals_out receive the sorted pairs; inputs are not modified. */
int bolt_sort_pairs_i64(bolt_stream_t stream, const int64_t* keys_in, const uint32_t* vals_in,
                        int64_t n, int32_t descending, int64_t* keys_out, uint32_t* vals_out);

/* ORDER BY: stable row permutation for a lexicographic multi-column key
 * (execute_sort, src/exec/sort.rs:161; key transforms of
 * src/exec/gpu_sort.rs:1070-1180: signed ints by sign flip, floats by the
 * IEEE-monotonic flip with NaN canonicalised last and -0.0 == +0.0;
 * Decimal128 keys as signed 128-bit integers, which the reference's host
 * lexsort orders and its GPU transform does not).
 * descending / nulls_first are host arrays of n_keys flags. */
int bolt_sort_indices(bolt_stream_t stream, const bolt_column_t* keys, int32_t n_keys,
                      const int32_t* descending, const int32_t* nulls_first, uint32_t* out_rows);

/* ORDER BY ... LIMIT limit OFFSET offset (execute_sort + execute_limit,
 * src/exec/sort.rs:161 and src/exec/limit.rs:18, which sort every row on
 * the host and then slice): writes the row ids at positions
 * [offset, offset + limit) of bolt_sort_indices's permutation for the same
 * keys, *n_out = min(limit, max(n - offset, 0)) of them, to out_rows
 * (capacity limit).  A radix select on the leading key picks the rows that
 * can reach those positions, and only they are sorted by every key, then by
 * row id, so the result equals the slice of the stable full sort.
 * n < 2^31.  Synchronises. */
int bolt_sort_limit(bolt_stream_t stream, const bolt_column_t* keys, int32_t n_keys, const int32_t* descending,
                    const int32_t* nulls_first, int64_t offset, int64_t limit, uint32_t* out_rows, int64_t* n_out);

/* ------------------------------------------------------------------------
 * Per-row expression programs — replaces the projection kernel of
 * Engine::execute_projection (src/exec/engine.rs:4267) generated from a
 * KernelSpec's Op list (src/plan/physical_plan.rs:81, lowered by
 * src/jit/ptx_gen.rs).  A program is a list of ops over registers (each
 * register written once, before use).  Semantics of ptx_gen.rs:
 *   integer Add / Sub / Mul wrap (two's complement); integer Div and Mod by
 *   zero give 0, INT_MIN / -1 gives INT_MIN and INT_MIN % -1 gives 0
 *   (emit_int_div_guarded / emit_int_rem_guarded, :2231,2332); float ops are
 *   IEEE, one rounding per op (no contraction); float comparisons are ordered
 *   (NaN compares false, != included); Float -> Int casts truncate and
 *   saturate (NaN -> 0); numeric -> Bool is "!= 0"; Date32 - Date32 is Int32
 *   days; Decimal128 Add / Sub / Mul act on the unscaled 128-bit integers
 *   (wrapping) and Decimal128 values compare as signed 128-bit integers.
 * Casts to and from Decimal128 follow emit_decimal_cast
 * (src/plan/physical_plan.rs:2046-2060): the op's `op` field is the target
 * scale when the target is Decimal128 and its `c` field the source scale when
 * the source is; integers scale up by 10^s (wrapping); a decimal rescales up
 * by multiplying and down by dividing, truncating toward zero; decimal ->
 * integer divides out the scale (truncating) and keeps the low 64 / 32 bits;
 * float -> decimal scales by 10^s in f64 and rounds half away from zero
 * (NaN -> 0, magnitudes of 2^127 and more saturate); decimal -> float is
 * (double)hi * 2^64 + (double)lo, divided by 10^s (and rounded to f32 for a
 * Float32 target).  Float conversions take scales up to 22 (exact powers of
 * ten in f64).  A Date32 casts to Decimal128 as its day number; Bool /
 * Timestamp <-> Decimal128 and Decimal128 -> Date32 casts are type errors.
 * Date32 / Timestamp fields (date_scalar.rs:459,539) use the proleptic
 * Gregorian calendar (civil_from_days, date_scalar.rs:37-60) with floor
 * division for instants before the epoch: EXTRACT gives an Int64 (YEAR,
 * MONTH 1-12, DAY 1-31, HOUR, MINUTE, SECOND; the last three need a
 * Timestamp), DATE_TRUNC gives the value's own dtype truncated to the start of
 * its unit.  A Timestamp register keeps its unit (a LOAD takes the column's
 * `scale`, a CONST its `op` field); comparisons, SELECT and STORE need equal
 * units.
 * Each output with a validity bitmap receives the AND of the validity of
 * every input column its value loads (the AND-of-inputs fold of
 * ptx_gen.rs:474); values of NULL rows are computed from the stored bits.
 * ------------------------------------------------------------------------ */
enum bolt_expr_kind {
  BOLT_X_LOAD = 1,    /* dst = inputs[col][row]                      (LoadColumn) */
  BOLT_X_CONST = 2,   /* dst = literal of `dtype`                      (Const) */
  BOLT_X_CAST = 3,    /* dst = (dtype) a                               (Cast) */
  BOLT_X_BINARY = 4,  /* dst = a op b, operands of `dtype`             (Binary) */
  BOLT_X_SELECT = 5,  /* dst = a ? b : c, a Bool                       (Select) */
  BOLT_X_NOT = 6,     /* dst = !a, a Bool                              (Not) */
  BOLT_X_IS_NULL = 7, /* dst = inputs[col] is NULL (op 1) / not (op 0) (IsNullCheck) */
  BOLT_X_STORE = 8,   /* outputs[col][row] = a                         (Store) */
  BOLT_X_EXTRACT = 9, /* dst = EXTRACT(field `op` FROM a), Int64      (compile_extract_kernel) */
  BOLT_X_DATE_TRUNC = 10 /* dst = DATE_TRUNC(unit `op`, a), a's dtype  (compile_date_trunc_kernel) */
};
/* EXTRACT fields and DATE_TRUNC units (DateField / DateTruncUnit,
 * src/plan/logical_plan.rs:568,613) */
enum bolt_date_field {
  BOLT_DF_YEAR = 1, BOLT_DF_MONTH = 2, BOLT_DF_DAY = 3, BOLT_DF_HOUR = 4, BOLT_DF_MINUTE = 5, BOLT_DF_SECOND = 6
};
enum bolt_bin_op {
  BOLT_BIN_ADD = 1, BOLT_BIN_SUB = 2, BOLT_BIN_MUL = 3, BOLT_BIN_DIV = 4, BOLT_BIN_MOD = 5,
  BOLT_BIN_EQ = 6, BOLT_BIN_NE = 7, BOLT_BIN_LT = 8, BOLT_BIN_LE = 9, BOLT_BIN_GT = 10, BOLT_BIN_GE = 11,
  BOLT_BIN_AND = 12, BOLT_BIN_OR = 13, BOLT_BIN_BITAND = 14, BOLT_BIN_BITOR = 15
};
typedef struct bolt_expr_op {
  int32_t kind;    /* bolt_expr_kind */
  int32_t dst;     /* destination register, 0 .. BOLT_MAX_REGS - 1 */
  int32_t a, b, c; /* operand registers (CAST from Decimal128: c is the source scale) */
  int32_t col;     /* LOAD / IS_NULL: input column; STORE: output column */
  int32_t op;      /* BINARY: bolt_bin_op; IS_NULL: 1 = IS NULL, 0 = IS NOT NULL;
                      CAST to Decimal128: target scale; CONST Timestamp: unit digits;
                      EXTRACT / DATE_TRUNC: bolt_date_field */
  int32_t dtype;   /* LOAD / CONST: value dtype; CAST: target; BINARY: operand dtype */
  int64_t lit_i64; /* CONST: integer / date / Bool value, Decimal128 low word */
  double lit_f64;  /* CONST: float value */
  int64_t lit_hi;  /* CONST: Decimal128 high word */
} bolt_expr_op_t;
#define BOLT_MAX_REGS 256
#define BOLT_MAX_PROJECT_INPUTS 16
#define BOLT_MAX_PROJECT_OUTPUTS 16

/* Runs `prog` for every row: inputs and outputs (caller-allocated, dtypes
 * set, validity bitmaps optional) all have outputs[0].length rows.  The
 * program is compiled once per distinct program and column shape (hipRTC)
 * and cached.  A type error in the program returns BOLT_ERR_TYPE, a
 * malformed program BOLT_ERR_PLAN.  Synchronises. */
int bolt_project(bolt_stream_t stream, const bolt_column_t* inputs, int32_t n_inputs, const bolt_expr_op_t* prog,
                 int32_t n_ops, bolt_column_t* outputs, int32_t n_outputs);

/* Fused WHERE + aggregates over an expression program — replaces
 * agg_with_pre::try_execute_resident (src/exec/agg_with_pre.rs:157), whose
 * pre kernel materialises the expression inputs (and expr_agg.rs's host
 * evaluator) before the aggregate runs.  Here one generated kernel runs the
 * program, the predicate and the aggregates per row and materialises
 * nothing.  The program's STOREs define n_values values (STORE col =
 * value index, each once); `filter_col` (or -1) names a Bool value as the
 * WHERE predicate (a row is kept where it is true and not NULL); each
 * aggregate's `col` indexes a value (ignored for COUNT(*)).  A value is NULL
 * where an input column it loads is NULL, and NULL values are skipped as in
 * bolt_aggregate, whose result rules (output dtypes, SUM overflow ->
 * BOLT_ERR_TYPE) apply.  Synchronises. */
int bolt_aggregate_expr(bolt_stream_t stream, const bolt_column_t* inputs, int32_t n_inputs,
                        const bolt_expr_op_t* prog, int32_t n_ops, int32_t n_values, int32_t filter_col,
                        const bolt_agg_t* aggs, int32_t n_aggs, bolt_scalar_t* results);

/* GROUP BY over an expression program — replaces execute_groupby_with_pre
 * (src/exec/groupby_with_pre.rs:222), whose pre kernel materialises the
 * program's outputs, compacts them on the host by the predicate and then
 * groups bare columns.  Here the program's n_values values (its STOREs, as
 * in bolt_aggregate_expr) are materialised once into pooled device columns
 * (a value is NULL where an input column it loads is NULL), the Bool value
 * `filter_col` (or -1) becomes the GROUP BY's row filter (a row is kept
 * where it is true and not NULL: nothing is compacted), and bolt_groupby
 * runs with key_values (1 or 2 value indexes) as its keys and the values as
 * its inputs (each aggregate's `col` is a value index).  value_scales
 * (optional, n_values entries) gives each Decimal128 value's scale for the
 * result; Timestamp values keep the program's unit.  Result, flags, DECLINE
 * and errors as bolt_groupby. */
int bolt_groupby_expr(bolt_stream_t stream, const bolt_column_t* inputs, int32_t n_inputs,
                      const bolt_expr_op_t* prog, int32_t n_ops, int32_t n_values, const int32_t* value_scales,
                      int32_t filter_col, const int32_t* key_values, int32_t n_keys, const bolt_agg_t* aggs,
                      int32_t n_aggs, uint32_t flags, bolt_table_t** out);

/* The fused join probe over an expression program: the probe side's key,
 * predicate and aggregate inputs are values of `prog` over the probe
 * columns `inputs` (the pre stage of agg_with_pre.rs:157 feeding the
 * aggregate-over-join composition, src/plan/physical_plan.rs:3691-3712).
 * The values are materialised once as in bolt_groupby_expr; key_value is the
 * probe key value, filter_col (or -1) a Bool value, and each aggregate's
 * `col` a value index or BOLT_JOIN_CARRY.  Results and errors as
 * bolt_join_probe_aggregate. */
int bolt_join_probe_aggregate_expr(bolt_stream_t stream, const bolt_join_table_t* t, const bolt_column_t* inputs,
                                   int32_t n_inputs, const bolt_expr_op_t* prog, int32_t n_ops, int32_t n_values,
                                   const int32_t* value_scales, int32_t key_value, int32_t filter_col,
                                   const bolt_agg_t* aggs, int32_t n_aggs, bolt_scalar_t* results);

/* ------------------------------------------------------------------------
 * Library information
 * ------------------------------------------------------------------------ */
/* Build identifier string (target arch, version). */
const char* bolt_build_info(void);

/* Generated-kernel cache counters of this process (src/jit/disk_cache.rs
 * keeps its persistent layer behind BOLT_PTX_CACHE_DIR; here
 * BOLT_JIT_CACHE_DIR holds gfx950 code objects): kernels compiled with
 * hipRTC, kernels loaded from the directory instead, and files written to it.
 * Any pointer may be NULL. */
int bolt_jit_cache_stats(int64_t* compiled, int64_t* disk_hits, int64_t* disk_stores);

/* Measurement hook (no reference counterpart; for benchmarks).  While set on
 * the calling thread, the operators record `start` immediately before and
 * `end` immediately after their dominant kernel, on the stream that kernel
 * runs on (bolt_join_probe_aggregate: the fused probe; bolt_aggregate: the
 * row pass; bolt_sort_pairs_i64: the whole sort; bolt_groupby: the pass over
 * the input rows, i.e. the LDS tier, or the global-table pass when the LDS
 * tier is off), so that bolt_hip_event_elapsed_ms gives the kernel's own time.
 * Pass NULL, NULL to clear. */
int bolt_profile_events(bolt_event_t start, bolt_event_t end);

#ifdef __cplusplus
}
#endif
#endif /* BOLT_OPS_H */
