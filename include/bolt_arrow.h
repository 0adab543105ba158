/* bolt_arrow.h — Arrow C Data Interface ingestion into device-resident tables,
 * and device columns back out as Arrow arrays.
 *
 * The operator entry points take device columns (bolt_column_t, Arrow layout:
 * values + LSB-first validity bitmap).  This header is the format boundary on
 * either side of them: record batches arriving as Arrow C Data Interface
 * structs (exported by pyarrow after it decodes an IPC file / stream, or by
 * any Arrow implementation) are uploaded into a table the library owns, whose
 * columns then feed any operator; results go back out as Arrow arrays in host
 * memory with the Arrow release-callback convention.
 *
 * Thread safety: a table is used by one thread at a time. */
#pragma once

#include <stdint.h>

#include "bolt_hip.h"
#include "bolt_ops.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The Arrow C Data Interface structs, as the Arrow specification defines
 * them (an ABI shared by every Arrow implementation; the guard is the one the
 * specification prescribes, so the header coexists with Arrow's own). */
#ifndef ARROW_C_DATA_INTERFACE
#define ARROW_C_DATA_INTERFACE
#define ARROW_FLAG_DICTIONARY_ORDERED 1
#define ARROW_FLAG_NULLABLE 2
#define ARROW_FLAG_MAP_KEYS_SORTED 4
struct ArrowSchema {
  const char* format;
  const char* name;
  const char* metadata;
  int64_t flags;
  int64_t n_children;
  struct ArrowSchema** children;
  struct ArrowSchema* dictionary;
  void (*release)(struct ArrowSchema*);
  void* private_data;
};
struct ArrowArray {
  int64_t length;
  int64_t null_count;
  int64_t offset;
  int64_t n_buffers;
  int64_t n_children;
  const void** buffers;
  struct ArrowArray** children;
  struct ArrowArray* dictionary;
  void (*release)(struct ArrowArray*);
  void* private_data;
};
#endif

typedef struct bolt_gpu_table* bolt_gpu_table_t;

/* Arrow format string -> column dtype, the mapping of
 * GpuTable::upload_column_from_batch (src/exec/gpu_table.rs:805-892):
 *   "i" Int32, "l" Int64, "f" Float32, "g" Float64, "b" Bool,
 *   "tdD" Date32, "ts{s,m,u,n}:<tz>" Timestamp (scale = unit digits 0/3/6/9;
 *   the time zone is kept as the column's metadata only),
 *   "d:<p>,<s>" / "d:<p>,<s>,128" Decimal128 (scale = s, 0 <= s <= p <= 38).
 * Everything else (Utf8, dictionaries, Date64, 256-bit decimals, nested
 * types) is BOLT_ERR_TYPE "unsupported Arrow dtype".  Host only. */
int bolt_arrow_format_dtype(const char* format, int32_t* dtype, int32_t* scale);

/* GpuTable::from_record_batch (src/exec/gpu_table.rs:898): `schema` is a
 * struct ("+s") whose children are the columns, `batch` the matching struct
 * array (a record batch as Arrow exports it; the struct's own offset and
 * every child's offset are honoured, a NULL struct row is an error).  Every
 * column is copied host -> device on `stream` into storage the table owns:
 * values in the operators' layout (Bool unpacked to one byte per row),
 * validity re-based to bit 0 (no bitmap when the column has no NULLs).
 * The caller keeps ownership of `schema` and `batch` (nothing is released).
 * Returns once the copies are complete (the host buffers may then go). */
int bolt_gpu_table_from_arrow(bolt_stream_t stream, const struct ArrowSchema* schema, const struct ArrowArray* batch,
                          bolt_gpu_table_t* out);

/* Appends the next record batch of the same schema (the next batch of an IPC
 * stream or of a Flight DoPut, src/flight/mod.rs:129): column count, dtypes
 * and scales must match.  Device storage grows geometrically; column
 * pointers returned before the append are invalid after it. */
int bolt_gpu_table_append_arrow(bolt_gpu_table_t table, bolt_stream_t stream, const struct ArrowSchema* schema,
                            const struct ArrowArray* batch);

/* Grows the table's storage to hold `rows` rows in total without further
 * reallocation (a reader that knows the stream's length calls it first). */
int bolt_gpu_table_reserve(bolt_gpu_table_t table, int64_t rows, bolt_stream_t stream);

int bolt_gpu_table_num_rows(bolt_gpu_table_t table, int64_t* n_rows);
int bolt_gpu_table_num_columns(bolt_gpu_table_t table, int32_t* n_columns);
/* Column i as an operator input (device pointers owned by the table). */
int bolt_gpu_table_column(bolt_gpu_table_t table, int32_t i, bolt_column_t* out);
/* The column's Arrow field name, and its Timestamp time zone ("" if none);
 * the strings live as long as the table. */
int bolt_gpu_table_column_name(bolt_gpu_table_t table, int32_t i, const char** name, const char** timezone);
/* Frees the table's device storage (after the null stream's pending work). */
int bolt_gpu_table_release(bolt_gpu_table_t table);

/* A device column -> an Arrow array in host memory (the result direction):
 * fills `out_schema` / `out_array` with library-allocated buffers that the
 * caller frees by calling their `release` callbacks (Arrow convention).
 * Decimal128 columns export as "d:38,<scale>", Timestamps as "ts?:" of the
 * column's unit.  Synchronises `stream`. */
int bolt_column_to_arrow(bolt_stream_t stream, const bolt_column_t* column, const char* name,
                         struct ArrowSchema* out_schema, struct ArrowArray* out_array);

#ifdef __cplusplus
}
#endif
