/*
 * bolt_oracle.h — CPU oracle for the GPU relational-operator path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates the reference's host
 * algorithms (single-threaded, as the reference runs them) so that the HIP
 * operators can be checked against them.  Only tests/, the smoke() check of
 * __graft_entry__.py and the cpu_baseline leg of bench.py may load it; the
 * product library never links or calls it.
 *
 * Each function cites the reference file:line whose behaviour it restates.
 * Pinning (what checks each function against something other than itself) is
 * listed in tests/test_oracle_golden.py; every function here is pinned by at
 * least one reference known-answer fixture, closed form or brute-force check.
 */
#ifndef BOLT_ORACLE_H
#define BOLT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* dtype codes (same values as the product ABI's bolt_dtype) */
enum { BO_BOOL = 1, BO_I32 = 2, BO_I64 = 3, BO_F32 = 4, BO_F64 = 5, BO_DEC128 = 6, BO_DATE32 = 7, BO_TIMESTAMP = 8 };
/* aggregate codes (same values as the product ABI's bolt_agg_op) */
enum {
  BO_SUM = 1, BO_COUNT = 2, BO_COUNT_STAR = 3, BO_MIN = 4, BO_MAX = 5, BO_AVG = 6,
  BO_VAR_POP = 7, BO_VAR_SAMP = 8, BO_STDDEV_POP = 9, BO_STDDEV_SAMP = 10
};
/* comparison codes for predicate terms (same values as bolt_cmp_op) */
enum { BO_LT = 1, BO_LE = 2, BO_GT = 3, BO_GE = 4, BO_EQ = 5, BO_NE = 6, BO_IS_NULL = 7, BO_IS_NOT_NULL = 8 };
/* status codes (same values as the product ABI) */
enum { BO_OK = 0, BO_ERR_PLAN = 1, BO_ERR_TYPE = 2, BO_ERR_CAP = 3 };

/* A host column: Arrow-style values buffer plus optional LSB-first validity
 * bitmap (NULL = all valid).  Decimal128 values are 16-byte little-endian. */
typedef struct bo_col {
  const void* data;
  const uint8_t* validity;
  int64_t n;
  int32_t dtype;
  int32_t scale;
} bo_col;

typedef struct bo_scalar {
  int32_t status;  /* BO_OK or an error code */
  int32_t is_null; /* SQL NULL result */
  int32_t dtype;   /* output dtype */
  int32_t pad;
  int64_t i64;     /* integer / date / decimal low word */
  double f64;      /* float result */
  int64_t hi;      /* decimal high word */
} bo_scalar;

typedef struct bo_term {
  int32_t col;
  int32_t op;
  int64_t lit_i64;
  double lit_f64;
  int64_t lit_hi;
} bo_term;

/* ---- Xorshift64Star (tests/common/mod.rs:46-90) ---- */
uint64_t bo_rng_seed(uint64_t seed);                /* state after `new` */
uint64_t bo_rng_next(uint64_t* state);              /* next_u64 */
double bo_rng_unit(uint64_t* state);                /* next_unit_f64 */
void bo_rng_fill(uint64_t seed, uint64_t* out, int64_t n); /* n successive draws */
/* Knuth-LCG Fisher-Yates shuffle of u32 (tests/common/mod.rs:102-110). */
void bo_shuffle_u32(uint32_t* xs, int64_t n, uint64_t seed);

/* ---- Neumaier compensated sum (src/exec/aggregate.rs:2001) ---- */
double bo_neumaier_sum(const double* x, int64_t n);

/* ---- Scalar aggregate (src/exec/aggregate.rs:154,1596,2032-2262;
 *      src/exec/welford.rs:47-130) ---- */
bo_scalar bo_scalar_agg(const bo_col* col, int32_t op);

/* ---- Filter (src/exec/compact.rs:100-146 + predicate semantics of
 *      src/jit/scan_kernel.rs:265): AND of terms, NULL compares false.
 *      Writes the surviving row indices in order; returns their count. ---- */
int64_t bo_filter_rows(const bo_col* cols, int32_t n_cols, const bo_term* terms, int32_t n_terms,
                       int64_t n_rows, uint32_t* out_rows);

/* ---- GROUP BY (src/exec/groupby.rs:295-833 classic semantics,
 *      src/exec/groupby_common.rs:204-420 key packing,
 *      src/exec/groupby_wide.rs:85 accumulators,
 *      src/exec/groupby.rs:1775 checked SUM) ----
 * keys: 1 or 2 key columns (two only when both are 32-bit).
 * Outputs are written for every group in ascending packed-key order (the
 * reference's `sort_unstable_by_key` order, groupby.rs:800):
 *   out_keys[g]       packed i64 key, out_key_null[g] 1 for the NULL group;
 *   out_i64[a][g] / out_f64[a][g] / out_hi[a][g] per aggregate a (integer,
 *   float and decimal-high results), out_null[a][g] 1 for an SQL NULL.
 * Every output array holds out_cap groups.
 * Returns BO_OK and sets *n_groups, BO_ERR_TYPE on integer SUM overflow, or
 * BO_ERR_CAP (nothing written, *n_groups = the group count) when the groups
 * exceed out_cap. */
int32_t bo_groupby(const bo_col* keys, int32_t n_keys, const bo_col* inputs, const int32_t* agg_ops,
                   const int32_t* agg_cols, int32_t n_aggs, int64_t* n_groups, int64_t* out_keys,
                   uint8_t* out_key_null, int64_t** out_i64, double** out_f64, int64_t** out_hi,
                   uint8_t** out_null, int64_t out_cap);

/* ---- INNER hash join (src/exec/join.rs:169-270, build_hash_map :844) ----
 * Smaller side builds (left when n_left <= n_right); NULL keys never match.
 * Pairs are emitted in probe-row order, build matches in build-row order, then
 * oriented to (left, right).  Returns the pair count, or -1 if `cap` is too
 * small (the count is then written to *needed). */
int64_t bo_inner_join(const bo_col* left_key, const bo_col* right_key, uint32_t* out_left,
                      uint32_t* out_right, int64_t cap, int64_t* needed);

/* ---- Equi-join of any kind (execute_join, src/exec/join.rs:132: INNER
 *      :169-270, LEFT / RIGHT / FULL OUTER :283-460; key values of
 *      extract_key :1045-1130 with the float / multi-key encodings of
 *      src/exec/gpu_join.rs:682-790) ----
 * join_type 0 INNER, 1 LEFT, 2 RIGHT, 3 FULL.  n_keys 1 (any of Bool / Int32
 * / Int64 / Date32 / Float32 / Float64) or 2 (Int32 / Date32).  A NULL in any
 * key column matches nothing.  Pairs are (left_row, right_row) with
 * 0xFFFFFFFF on the NULL-padded side, in the reference's host order: probe
 * rows ascending, each with its build matches in build-row order or one
 * padded pair, then (FULL) the unmatched build rows ascending.  Returns the
 * pair count, or -1 if `cap` is too small (*needed holds the count). */
int64_t bo_join(const bo_col* left_keys, const bo_col* right_keys, int32_t n_keys, int32_t join_type,
                uint32_t* out_left, uint32_t* out_right, int64_t cap, int64_t* needed);

/* ---- ORDER BY (src/exec/sort.rs:161-230 lexsort + take; key transforms of
 *      src/exec/gpu_sort.rs:1070-1180) ----
 * Stable: equal keys keep input order (the canonical order used to compare
 * against the reference's unstable lexsort). */
void bo_sort_indices(const bo_col* keys, int32_t n_keys, const int32_t* descending,
                     const int32_t* nulls_first, int64_t n_rows, uint32_t* out_idx);

/* ---- Fused filter + INNER join + aggregate of the probe-side star query,
 *      as the reference composes it: filter (compact.rs:146) -> host inner join
 *      (join.rs:169) -> scalar SUM/SUM/COUNT (aggregate.rs:154,2001). ----
 * Returns SUM(build payload) (checked i64), SUM(probe amount) as the plain
 * left fold in join-output order and as Neumaier, and COUNT(*). */
int32_t bo_filter_join_agg(const int64_t* build_key, const int32_t* build_payload, int64_t n_build,
                           const int64_t* probe_key, const double* probe_amount, int64_t n_probe,
                           double amount_gt, int64_t* sum_payload, double* sum_amount_naive,
                           double* sum_amount_neumaier, int64_t* count);

/* ---- Per-row expression programs (the projection kernel of
 *      Engine::execute_projection, src/exec/engine.rs:4267, generated from
 *      the Op list of src/plan/physical_plan.rs:81 by src/jit/ptx_gen.rs) ----
 * Same op layout and kinds as the product ABI's bolt_expr_op_t.  Interprets
 * the program row by row with ptx_gen.rs's semantics: wrapping integer
 * arithmetic, guarded Div / Mod (emit_int_div_guarded :2231,
 * emit_int_rem_guarded :2332), IEEE float ops, ordered float comparisons,
 * saturating truncating Float -> Int casts (cvt.rzi), numeric -> Bool as
 * != 0, Date32 - Date32 as Int32, Decimal128 Add / Sub / Mul on the
 * unscaled integers.  Outputs with a validity bitmap get the AND of the
 * validity of the input columns their value loads (:474).  Returns BO_OK,
 * BO_ERR_TYPE or BO_ERR_PLAN. */
typedef struct bo_expr_op {
  int32_t kind, dst, a, b, c, col, op, dtype;
  int64_t lit_i64;
  double lit_f64;
  int64_t lit_hi;
} bo_expr_op;
int32_t bo_project(const bo_col* inputs, int32_t n_inputs, const bo_expr_op* prog, int32_t n_ops, bo_col* outputs,
                   int32_t n_outputs, int64_t n_rows);

#ifdef __cplusplus
}
#endif
#endif
