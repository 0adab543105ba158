/*
 * bolt_serve.h — served form: a request loop over a device-resident star
 * schema with HBM-budget admission and continuous batching.
 *
 * No reference counterpart.  The closest pieces are the Mutex<Engine> Flight
 * server (src/flight/mod.rs:86,129-168), which runs one query at a time, the
 * pinned-memory budget (src/exec/streaming.rs:392), memory_budget
 * (src/exec/engine.rs:1159) and GpuTable residency (src/exec/gpu_table.rs:775).
 *
 * Query shape (the star-query mix): over the fact table,
 *   WHERE <fact predicates>                    (AND of column-vs-literal terms)
 *   INNER JOIN dim_j ON fact.fk_j = dim_j.key  (1..3 dimensions; a fact row
 *                                               whose key is absent is dropped)
 *   GROUP BY <one dimension attribute, or two 32-bit attributes packed>
 *   aggregates over fact columns (SUM / COUNT / COUNT(*) / MIN / MAX / AVG)
 *   ORDER BY one aggregate DESC (ties: ascending group key) LIMIT k (k <= 1024)
 * with the semantics of the per-query operators in bolt_ops.h (exact integer
 * and Decimal128 SUM with overflow errors, NULL handling, output dtypes).
 *
 * Execution: one server thread owns the device.  The fact table is cut into
 * chunks; every step, one launch scans the next chunk once for every active
 * query (a shared circular scan): each row is loaded once, then filtered,
 * joined (probing the dimension hash tables built at registration) and
 * aggregated into each query's own HBM hash table.  A request that arrives
 * while others run is admitted at the next chunk boundary when a slot is free
 * and its HBM need (hash table + top-k scratch) fits the budget next to the
 * queries already admitted; it starts at the current chunk, wraps around the
 * table, and completes after one full lap (extraction + ORDER BY/LIMIT),
 * freeing its slot and memory.  Up to `slots` queries therefore overlap on
 * the device; slots = 1 is the plain one-query-at-a-time path.
 */
#ifndef BOLT_SERVE_H
#define BOLT_SERVE_H

#include "bolt_hip.h"
#include "bolt_ops.h"

#ifdef __cplusplus
extern "C" {
#endif

#define BOLT_SERVE_MAX_DIMS 3
#define BOLT_SERVE_MAX_AGGS 4
#define BOLT_SERVE_MAX_TERMS 4
#define BOLT_SERVE_MAX_LIMIT 1024
#define BOLT_SERVE_MAX_SLOTS 64

typedef struct bolt_server bolt_server_t;

typedef struct bolt_serve_config {
  int32_t slots;          /* queries overlapping on the device (1 .. BOLT_SERVE_MAX_SLOTS) */
  int64_t hbm_budget;     /* bytes of hash tables + scratch the admitted queries may hold */
  int64_t chunk_rows;     /* fact rows per step (0 = default) */
} bolt_serve_config_t;

/* A dimension: unique key column plus attribute columns (all device
 * resident, owned by the caller). */
typedef struct bolt_serve_dim {
  bolt_column_t key;                  /* Int32 / Int64, unique, no NULLs */
  const bolt_column_t* attrs;         /* attribute columns (Int32 / Int64 / Date32) */
  int32_t n_attrs;
} bolt_serve_dim_t;

/* The query template.  Column indices: fact columns index `fact` of
 * bolt_server_create; a group attribute is (dim index, attr index). */
typedef struct bolt_serve_query {
  int32_t n_terms;
  bolt_pred_term_t terms[BOLT_SERVE_MAX_TERMS]; /* on fact columns */
  int32_t n_joins;
  int32_t join_fk[BOLT_SERVE_MAX_DIMS];         /* fact column holding the foreign key */
  int32_t join_dim[BOLT_SERVE_MAX_DIMS];        /* dimension joined */
  int32_t n_group;                              /* 1 or 2 (two 32-bit attributes) */
  int32_t group_dim[2];                         /* which join (0..n_joins-1) supplies the attribute */
  int32_t group_attr[2];
  int32_t n_aggs;
  bolt_agg_t aggs[BOLT_SERVE_MAX_AGGS];         /* col indexes the fact columns */
  int32_t order_agg;                            /* ORDER BY aggs[order_agg] DESC */
  int32_t limit;                                /* 1 .. BOLT_SERVE_MAX_LIMIT */
  int64_t est_groups;                           /* expected group count (sizes the hash table) */
} bolt_serve_query_t;

/* One result row set: `n_rows` rows of n_group key columns then n_aggs
 * aggregate columns, host memory owned by the server until released. */
typedef struct bolt_serve_result {
  int32_t status;            /* BOLT_OK or an error code (message in `error`) */
  int32_t n_rows;
  int64_t n_groups;          /* groups before LIMIT */
  int64_t hbm_bytes;         /* device memory the query held while admitted */
  int32_t n_cols;
  const bolt_scalar_t* cells;/* row-major n_rows x n_cols */
  double submit_ms;          /* server clock: submission, admission, completion */
  double admit_ms;
  double done_ms;
  const char* error;
} bolt_serve_result_t;

/* Registers the fact table and dimensions (hash tables of the dimension keys
 * are built here, once) and starts the server thread. */
int bolt_server_create(bolt_stream_t stream, const bolt_column_t* fact, int32_t n_fact,
                       const bolt_serve_dim_t* dims, int32_t n_dims, const bolt_serve_config_t* cfg,
                       bolt_server_t** out);
/* Queues a request; returns its ticket.  Thread-safe.  A query whose HBM need
 * exceeds the whole budget fails with BOLT_DECLINE. */
int bolt_server_submit(bolt_server_t* srv, const bolt_serve_query_t* q, int64_t* ticket);
/* Waits for a ticket (timeout_ms < 0: forever); BOLT_HIP_ERROR_NOT_READY on
 * timeout.  The result stays valid until bolt_server_result_release. */
int bolt_server_wait(bolt_server_t* srv, int64_t ticket, double timeout_ms, bolt_serve_result_t* res);
void bolt_server_result_release(bolt_server_t* srv, int64_t ticket);
/* Milliseconds on the server clock (for arrival schedules). */
double bolt_server_now_ms(const bolt_server_t* srv);
/* Counters: steps run, queries completed, peak concurrent queries, peak
 * admitted HBM bytes, and the milliseconds spent in scan steps (launch to
 * completion) and in finishing queries (extraction, ORDER BY, copy-out; done
 * on a second stream, overlapping the scan). */
int bolt_server_stats(const bolt_server_t* srv, int64_t* steps, int64_t* completed, int32_t* peak_active,
                      int64_t* peak_bytes, double* step_ms, double* finish_ms);
/* Drains the queue, stops the thread and frees the device state. */
void bolt_server_destroy(bolt_server_t* srv);

#ifdef __cplusplus
}
#endif
#endif /* BOLT_SERVE_H */
