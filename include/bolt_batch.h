/*
 * bolt_batch.h — batched GROUP BY: many independent queries of uneven row
 * counts and cardinalities in one persistent work-queue pipeline, and the
 * autotuner that picks its launch parameters.
 *
 * No reference counterpart: the reference runs each query through its own
 * Engine::sql -> execute_groupby call (src/exec/groupby.rs:295), i.e. one
 * launch sequence with host round trips per query.  The closest pieces are
 * the per-query launch sequence, the KernelSpec module cache
 * (src/exec/module_cache.rs:291) and the CUDA-graph capture of the sort
 * (src/exec/gpu_sort.rs:2298-2513).
 *
 * Semantics per query are those of bolt_groupby (bolt_ops.h) without a
 * filter: the same key encoding, NULL group, accumulators (exact 128-bit
 * integer SUM with the i64 overflow error, exact COUNT, float SUM in f64) and
 * output dtypes.  Groups come out in hash order.
 *
 * Life cycle:
 *   bolt_gb_batch_create   plans the batch (one sketch pass), sizes one HBM
 *                          arena for all queries' hash tables and outputs and
 *                          captures the whole run as a hipGraph;
 *   bolt_gb_batch_run      replays the graph: sketch -> device-side plan ->
 *                          aggregate -> extract, no host round trip inside;
 *                          synchronises and checks the error flags;
 *   bolt_gb_batch_num_groups / bolt_gb_batch_column  read query q's result
 *                          (valid until the next run or release);
 *   bolt_gb_batch_release.
 * Input columns must stay allocated and unchanged while the batch exists
 * (their contents may change between runs: every run re-plans on the device).
 */
#ifndef BOLT_BATCH_H
#define BOLT_BATCH_H

#include "bolt_hip.h"
#include "bolt_ops.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Launch parameters of the batched pipeline (0 = take the tuned / default
 * value).  The autotuner searches these. */
typedef struct bolt_gb_tune {
  int32_t block;      /* threads per workgroup of the aggregation pass: 256, 512 or 1024 */
  int32_t tile_rows;  /* rows per work item (a power of two, 4096 .. 1048576) */
  int32_t lds_slots;  /* maximum LDS hash-table slots per workgroup (power of two) */
  int32_t grid;       /* persistent workgroups (0 = 2 per CU) */
} bolt_gb_tune_t;

typedef struct bolt_gb_batch bolt_gb_batch_t;

/* Query q (0 <= q < n_queries) groups keys[q * n_keys .. + n_keys) and
 * aggregates inputs[q * n_inputs .. + n_inputs) with the shared aggregate list
 * `aggs` (column indices are per query).  Every query must use the same key /
 * input dtypes and nullability.  `tune` may be NULL.  A non-NULL `stream`
 * must outlive the batch: every run is queued on it and release frees the
 * batch's device buffers in its order (the NULL stream gets a private one). */
int bolt_gb_batch_create(bolt_stream_t stream, const bolt_column_t* keys, int32_t n_keys,
                         const bolt_column_t* inputs, int32_t n_inputs, int32_t n_queries,
                         const bolt_agg_t* aggs, int32_t n_aggs, const bolt_gb_tune_t* tune,
                         bolt_gb_batch_t** out);
int bolt_gb_batch_run(bolt_gb_batch_t* b);
int bolt_gb_batch_num_groups(const bolt_gb_batch_t* b, int32_t q, int64_t* n_groups);
/* Column i (n_keys key columns, then one per aggregate) of query q's result. */
int bolt_gb_batch_column(const bolt_gb_batch_t* b, int32_t q, int32_t i, bolt_column_t* col);
/* The parameters in use, and the HBM bytes the batch holds. */
int bolt_gb_batch_get_tune(const bolt_gb_batch_t* b, bolt_gb_tune_t* tune, int64_t* arena_bytes);
void bolt_gb_batch_release(bolt_gb_batch_t* b);

/* Autotuner: times `reps` runs of each candidate parameter set on this
 * batch's own data (after one warm-up run each), keeps the fastest, and
 * records it under the batch's shape class so that later batches of the same
 * class start from it.  With BOLT_TUNE_FILE set, the table of tuned classes
 * is loaded from and saved to that file.  *best_ms receives the winner's
 * mean run time. */
int bolt_gb_batch_autotune(bolt_gb_batch_t* b, int32_t reps, bolt_gb_tune_t* best, float* best_ms);

#ifdef __cplusplus
}
#endif
#endif /* BOLT_BATCH_H */
