// arrow.hip — Arrow C Data Interface record batches -> device-resident tables,
// device columns -> Arrow arrays (bolt_arrow.h).
//
// Ingestion is a host -> device copy per buffer plus two small kernels: a
// bit re-basing kernel for validity bitmaps (an Arrow slice starts at any bit
// offset, a table append lands at any row) and a bit -> byte unpack for Bool
// values (Arrow packs Bool to bits, the operators read one byte per row).
// Both kernels write each destination byte from exactly one thread.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "../../include/bolt_arrow.h"
#include "bolt_internal.h"

namespace bolt {
namespace {

// dst bits [dst_bit0, dst_bit0 + n) = src bits [src_bit0, src_bit0 + n) (or
// all ones when src is null); the other bits of the first and last
// destination bytes keep their values.  One thread per destination byte.
__global__ void rebase_bits_kernel(uint8_t* __restrict__ dst, int64_t dst_bit0, const uint8_t* __restrict__ src,
                                   int64_t src_bit0, int64_t n) {
  const int64_t first = dst_bit0 >> 3, last = (dst_bit0 + n - 1) >> 3;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t byte = first + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; byte <= last; byte += stride) {
    uint8_t v = (byte == first || byte == last) ? dst[byte] : 0;
    for (int b = 0; b < 8; ++b) {
      const int64_t g = byte * 8 + b - dst_bit0;  // row within the appended range
      if (g < 0 || g >= n) continue;
      const int64_t s = src_bit0 + g;
      const int bit = src ? (src[s >> 3] >> (s & 7)) & 1 : 1;
      v = (uint8_t)((v & ~(1u << b)) | ((unsigned)bit << b));
    }
    dst[byte] = v;
  }
}

// dst[i] = bit (src_bit0 + i) of src, for i < n.
__global__ void unpack_bits_kernel(uint8_t* __restrict__ dst, const uint8_t* __restrict__ src, int64_t src_bit0,
                                   int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int64_t s = src_bit0 + i;
    dst[i] = (src[s >> 3] >> (s & 7)) & 1;
  }
}

struct TableCol {
  std::string name, tz;
  int32_t dtype = 0, scale = 0;
  void* data = nullptr;        // cap rows
  uint8_t* validity = nullptr;  // cap bits (+ slack), nullptr while no NULL has arrived
};

int parse_int(const char*& p, int* out) {
  if (*p < '0' || *p > '9') return 0;
  long v = 0;
  while (*p >= '0' && *p <= '9' && v < 1000) v = v * 10 + (*p++ - '0');
  *out = (int)v;
  return 1;
}

int unsupported(const char* format) {
  return set_error(BOLT_ERR_TYPE, "unsupported Arrow dtype (format \"%s\")", format ? format : "(null)");
}

size_t bitmap_bytes(int64_t rows) { return (size_t)((rows + 7) / 8) + 8; }

}  // namespace
}  // namespace bolt

struct bolt_gpu_table {
  int device = 0;
  int64_t rows = 0, cap = 0;
  std::vector<bolt::TableCol> cols;
};

using namespace bolt;

namespace {

int format_dtype(const char* f, int32_t* dtype, int32_t* scale, std::string* tz) {
  *scale = 0;
  if (!f) return unsupported(f);
  if (!strcmp(f, "i")) { *dtype = BOLT_I32; return BOLT_OK; }
  if (!strcmp(f, "l")) { *dtype = BOLT_I64; return BOLT_OK; }
  if (!strcmp(f, "f")) { *dtype = BOLT_F32; return BOLT_OK; }
  if (!strcmp(f, "g")) { *dtype = BOLT_F64; return BOLT_OK; }
  if (!strcmp(f, "b")) { *dtype = BOLT_BOOL; return BOLT_OK; }
  if (!strcmp(f, "tdD")) { *dtype = BOLT_DATE32; return BOLT_OK; }
  if (f[0] == 't' && f[1] == 's' && f[2] && f[3] == ':') {
    const char* units = "smun";
    const char* u = strchr(units, f[2]);
    if (!u) return unsupported(f);
    *dtype = BOLT_TIMESTAMP;
    *scale = 3 * (int)(u - units);
    if (tz) *tz = f + 4;
    return BOLT_OK;
  }
  if (f[0] == 'd' && f[1] == ':') {
    const char* p = f + 2;
    int prec = 0, sc = 0, bits = 128;
    if (!parse_int(p, &prec) || *p++ != ',') return unsupported(f);
    const bool neg = *p == '-';
    if (neg) ++p;
    if (!parse_int(p, &sc)) return unsupported(f);
    if (*p == ',') {
      ++p;
      if (!parse_int(p, &bits)) return unsupported(f);
    }
    if (*p || bits != 128 || neg || prec < 1 || prec > 38 || sc > prec) return unsupported(f);
    *dtype = BOLT_DEC128;
    *scale = sc;
    return BOLT_OK;
  }
  return unsupported(f);
}

// Schema of a record batch: a struct of supported primitive columns.
int check_schema(const ArrowSchema* s) {
  if (!s || !s->format || strcmp(s->format, "+s") != 0)
    return set_error(BOLT_ERR_TYPE, "a record batch is an Arrow struct (\"+s\"), got \"%s\"",
                     s && s->format ? s->format : "(null)");
  if (s->n_children < 0 || (s->n_children > 0 && !s->children))
    return set_error(BOLT_ERR_PLAN, "malformed Arrow schema: %lld children", (long long)s->n_children);
  for (int64_t i = 0; i < s->n_children; ++i) {
    const ArrowSchema* c = s->children[i];
    if (!c) return set_error(BOLT_ERR_PLAN, "malformed Arrow schema: child %lld is null", (long long)i);
    if (c->dictionary)
      return set_error(BOLT_ERR_TYPE, "unsupported Arrow dtype: dictionary-encoded column '%s'",
                       c->name ? c->name : "");
    int32_t dt, sc;
    BOLT_TRY(format_dtype(c->format, &dt, &sc, nullptr));
  }
  return BOLT_OK;
}

int check_batch(const ArrowSchema* s, const ArrowArray* a) {
  if (!a || a->length < 0 || a->offset < 0 || a->n_children != s->n_children || (a->n_children > 0 && !a->children))
    return set_error(BOLT_ERR_PLAN, "Arrow batch does not match its schema (%lld columns)", (long long)s->n_children);
  if (a->null_count != 0 && a->n_buffers > 0 && a->buffers && a->buffers[0])
    return set_error(BOLT_ERR_TYPE, "a record batch has no NULL rows (the struct array has a validity bitmap)");
  for (int64_t i = 0; i < a->n_children; ++i) {
    const ArrowArray* c = a->children[i];
    int32_t dt, sc;
    BOLT_TRY(format_dtype(s->children[i]->format, &dt, &sc, nullptr));
    if (!c || c->offset < 0 || c->length < a->offset + a->length || c->n_buffers < 2 || !c->buffers)
      return set_error(BOLT_ERR_PLAN, "Arrow column %lld is malformed or shorter than its batch", (long long)i);
    if (a->length > 0 && !c->buffers[1])
      return set_error(BOLT_ERR_PLAN, "Arrow column %lld has no values buffer", (long long)i);
  }
  return BOLT_OK;
}

// Capacity for `need` rows: exactly `need` when `exact` (a first batch, a
// reservation), else doubling, so a stream of appends copies each row O(1) times.
int grow(bolt_gpu_table* t, int64_t need, hipStream_t s, bool exact) {
  if (need <= t->cap) return BOLT_OK;
  int64_t cap = need;
  if (!exact) {
    cap = t->cap ? t->cap : 1024;
    while (cap < need) cap *= 2;
  }
  for (auto& c : t->cols) {
    const size_t w = (size_t)dtype_size(c.dtype);
    void* d = nullptr;
    BOLT_TRY(dev_alloc(&d, (size_t)cap * w, s));
    if (t->rows) BOLT_HIP(hipMemcpyAsync(d, c.data, (size_t)t->rows * w, hipMemcpyDeviceToDevice, s));
    dev_free(c.data, s);
    c.data = d;
    if (c.validity) {
      void* v = nullptr;
      BOLT_TRY(dev_alloc(&v, bitmap_bytes(cap), s));
      if (t->rows) BOLT_HIP(hipMemcpyAsync(v, c.validity, (size_t)(t->rows + 7) / 8, hipMemcpyDeviceToDevice, s));
      dev_free(c.validity, s);
      c.validity = static_cast<uint8_t*>(v);
    }
  }
  t->cap = cap;
  return BOLT_OK;
}

// Device copy of host bits [bit0, bit0 + n) as bytes [bit0 / 8, ...): returns the bit offset within the copy.
int stage_bits(DevBuf<uint8_t>* tmp, const uint8_t* host, int64_t bit0, int64_t n, hipStream_t s, int64_t* off) {
  const int64_t b0 = bit0 >> 3, b1 = (bit0 + n + 7) >> 3;
  BOLT_TRY(tmp->alloc((size_t)(b1 - b0), s));
  BOLT_HIP(hipMemcpyAsync(tmp->ptr, host + b0, (size_t)(b1 - b0), hipMemcpyHostToDevice, s));
  *off = bit0 & 7;
  return BOLT_OK;
}

int append(bolt_gpu_table* t, hipStream_t s, const ArrowSchema* schema, const ArrowArray* a) {
  const int64_t n = a->length, r0 = t->rows;
  if (n == 0) return BOLT_OK;
  BOLT_TRY(grow(t, r0 + n, s, r0 == 0 && t->cap == 0));
  std::vector<DevBuf<uint8_t>> staged;  // freed (stream-ordered) after the kernels that read them
  staged.reserve(2 * t->cols.size());
  for (size_t i = 0; i < t->cols.size(); ++i) {
    TableCol& c = t->cols[i];
    const ArrowArray* ch = a->children[i];
    const int64_t row0 = ch->offset + a->offset;  // the struct's offset slices its children
    const uint8_t* vbits = static_cast<const uint8_t*>(ch->buffers[0]);
    const bool has_nulls = vbits && ch->null_count != 0;
    // values
    if (c.dtype == BOLT_BOOL) {
      staged.emplace_back();
      int64_t off = 0;
      BOLT_TRY(stage_bits(&staged.back(), static_cast<const uint8_t*>(ch->buffers[1]), row0, n, s, &off));
      unpack_bits_kernel<<<grid_for(n, 256, 4), 256, 0, s>>>(static_cast<uint8_t*>(c.data) + r0, staged.back().ptr,
                                                             off, n);
      BOLT_HIP(hipGetLastError());
    } else {
      const size_t w = (size_t)dtype_size(c.dtype);
      BOLT_HIP(hipMemcpyAsync(static_cast<uint8_t*>(c.data) + (size_t)r0 * w,
                              static_cast<const uint8_t*>(ch->buffers[1]) + (size_t)row0 * w, (size_t)n * w,
                              hipMemcpyHostToDevice, s));
    }
    // validity
    if (has_nulls && !c.validity) {  // the first NULL: every earlier row is valid
      void* v = nullptr;
      BOLT_TRY(dev_alloc(&v, bitmap_bytes(t->cap), s));
      c.validity = static_cast<uint8_t*>(v);
      BOLT_TRY(fill_u8(c.validity, 0xFF, bitmap_bytes(t->cap), s));
    }
    if (c.validity) {
      const uint8_t* src = nullptr;
      int64_t off = 0;
      if (has_nulls) {
        staged.emplace_back();
        BOLT_TRY(stage_bits(&staged.back(), vbits, row0, n, s, &off));
        src = staged.back().ptr;
      }
      const int64_t bytes = ((r0 + n - 1) >> 3) - (r0 >> 3) + 1;
      rebase_bits_kernel<<<grid_for(bytes, 256, 1), 256, 0, s>>>(c.validity, r0, src, off, n);
      BOLT_HIP(hipGetLastError());
    }
  }
  BOLT_HIP(hipStreamSynchronize(s));  // the host buffers may go once this returns
  t->rows = r0 + n;
  (void)schema;
  return BOLT_OK;
}

// What an exported array / schema owns; freed by its release callback.
struct Exported {
  std::vector<void*> blocks;
  std::string format, name;
  const void* buffers[2] = {nullptr, nullptr};
  ~Exported() {
    for (void* b : blocks) free(b);
  }
};

void release_array(ArrowArray* a) {
  if (!a || !a->release) return;
  delete static_cast<Exported*>(a->private_data);
  a->release = nullptr;
}

void release_schema(ArrowSchema* s) {
  if (!s || !s->release) return;
  delete static_cast<Exported*>(s->private_data);
  s->release = nullptr;
}

}  // namespace

extern "C" {

int bolt_arrow_format_dtype(const char* format, int32_t* dtype, int32_t* scale) {
  clear_error();
  if (!dtype || !scale) return set_error(BOLT_ERR_PLAN, "bolt_arrow_format_dtype: null output");
  return format_dtype(format, dtype, scale, nullptr);
}

int bolt_gpu_table_from_arrow(bolt_stream_t stream, const ArrowSchema* schema, const ArrowArray* batch,
                          bolt_gpu_table_t* out) {
  clear_error();
  if (!out) return set_error(BOLT_ERR_PLAN, "bolt_gpu_table_from_arrow: null output");
  *out = nullptr;
  BOLT_CHECK_DEVICE();
  BOLT_TRY(check_schema(schema));
  BOLT_TRY(check_batch(schema, batch));
  std::unique_ptr<bolt_gpu_table> t(new bolt_gpu_table());
  BOLT_HIP(hipGetDevice(&t->device));
  for (int64_t i = 0; i < schema->n_children; ++i) {
    TableCol c;
    const ArrowSchema* f = schema->children[i];
    c.name = f->name ? f->name : "";
    BOLT_TRY(format_dtype(f->format, &c.dtype, &c.scale, &c.tz));
    t->cols.push_back(c);
  }
  hipStream_t s = (hipStream_t)stream;
  const int st = append(t.get(), s, schema, batch);
  if (st != BOLT_OK) {
    const std::string msg = bolt_last_error();
    bolt_gpu_table_release(t.release());
    return set_error(st, "%s", msg.c_str());
  }
  *out = t.release();
  return BOLT_OK;
}

int bolt_gpu_table_append_arrow(bolt_gpu_table_t t, bolt_stream_t stream, const ArrowSchema* schema, const ArrowArray* batch) {
  clear_error();
  if (!t) return set_error(BOLT_ERR_PLAN, "bolt_gpu_table_append_arrow: null table");
  BOLT_CHECK_DEVICE();
  BOLT_TRY(check_schema(schema));
  if ((size_t)schema->n_children != t->cols.size())
    return set_error(BOLT_ERR_TYPE, "appended batch has %lld columns, the table %zu", (long long)schema->n_children,
                     t->cols.size());
  for (size_t i = 0; i < t->cols.size(); ++i) {
    int32_t dt, sc;
    BOLT_TRY(format_dtype(schema->children[i]->format, &dt, &sc, nullptr));
    if (dt != t->cols[i].dtype || sc != t->cols[i].scale)
      return set_error(BOLT_ERR_TYPE, "appended column %zu is %s(%d), the table's %s(%d)", i, dtype_name(dt), sc,
                       dtype_name(t->cols[i].dtype), t->cols[i].scale);
  }
  BOLT_TRY(check_batch(schema, batch));
  return append(t, (hipStream_t)stream, schema, batch);
}

int bolt_gpu_table_reserve(bolt_gpu_table_t t, int64_t rows, bolt_stream_t stream) {
  clear_error();
  if (!t || rows < 0) return set_error(BOLT_ERR_PLAN, "bolt_gpu_table_reserve: bad argument");
  BOLT_CHECK_DEVICE();
  hipStream_t s = (hipStream_t)stream;
  BOLT_TRY(grow(t, rows, s, true));
  BOLT_HIP(hipStreamSynchronize(s));
  return BOLT_OK;
}

int bolt_gpu_table_num_rows(bolt_gpu_table_t t, int64_t* n) {
  if (!t || !n) return set_error(BOLT_ERR_PLAN, "bolt_gpu_table_num_rows: null argument");
  *n = t->rows;
  return BOLT_OK;
}

int bolt_gpu_table_num_columns(bolt_gpu_table_t t, int32_t* n) {
  if (!t || !n) return set_error(BOLT_ERR_PLAN, "bolt_gpu_table_num_columns: null argument");
  *n = (int32_t)t->cols.size();
  return BOLT_OK;
}

int bolt_gpu_table_column(bolt_gpu_table_t t, int32_t i, bolt_column_t* out) {
  if (!t || !out) return set_error(BOLT_ERR_PLAN, "bolt_gpu_table_column: null argument");
  if (i < 0 || (size_t)i >= t->cols.size())
    return set_error(BOLT_ERR_PLAN, "bolt_gpu_table_column: column %d of %zu", i, t->cols.size());
  const TableCol& c = t->cols[i];
  out->data = c.data;
  out->validity = c.validity;
  out->length = t->rows;
  out->dtype = c.dtype;
  out->scale = c.scale;
  return BOLT_OK;
}

int bolt_gpu_table_column_name(bolt_gpu_table_t t, int32_t i, const char** name, const char** timezone) {
  if (!t || i < 0 || (size_t)i >= t->cols.size())
    return set_error(BOLT_ERR_PLAN, "bolt_gpu_table_column_name: no column %d", i);
  if (name) *name = t->cols[i].name.c_str();
  if (timezone) *timezone = t->cols[i].tz.c_str();
  return BOLT_OK;
}

int bolt_gpu_table_release(bolt_gpu_table_t t) {
  if (!t) return BOLT_OK;
  for (auto& c : t->cols) {
    dev_free(c.data, nullptr);
    dev_free(c.validity, nullptr);
  }
  delete t;
  return BOLT_OK;
}

int bolt_column_to_arrow(bolt_stream_t stream, const bolt_column_t* col, const char* name, ArrowSchema* out_schema,
                         ArrowArray* out_array) {
  clear_error();
  if (!col || !out_schema || !out_array) return set_error(BOLT_ERR_PLAN, "bolt_column_to_arrow: null argument");
  BOLT_CHECK_DEVICE();
  const int64_t n = col->length;
  const int w = dtype_size(col->dtype);
  if (w == 0 || n < 0) return set_error(BOLT_ERR_TYPE, "bolt_column_to_arrow: dtype %d", col->dtype);
  std::unique_ptr<Exported> arr(new Exported()), sch(new Exported());
  char fmt[32];
  switch (col->dtype) {
    case BOLT_I32: strcpy(fmt, "i"); break;
    case BOLT_I64: strcpy(fmt, "l"); break;
    case BOLT_F32: strcpy(fmt, "f"); break;
    case BOLT_F64: strcpy(fmt, "g"); break;
    case BOLT_BOOL: strcpy(fmt, "b"); break;
    case BOLT_DATE32: strcpy(fmt, "tdD"); break;
    case BOLT_DEC128: snprintf(fmt, sizeof fmt, "d:38,%d", col->scale); break;
    case BOLT_TIMESTAMP: {
      const int u = col->scale / 3;
      if (col->scale % 3 || u < 0 || u > 3) return set_error(BOLT_ERR_TYPE, "Timestamp unit %d", col->scale);
      snprintf(fmt, sizeof fmt, "ts%c:", "smun"[u]);
      break;
    }
    default: return set_error(BOLT_ERR_TYPE, "bolt_column_to_arrow: dtype %d", col->dtype);
  }
  hipStream_t s = (hipStream_t)stream;
  // values
  void* vals = malloc((size_t)(n ? n : 1) * w);
  if (!vals) return set_error(BOLT_ERR_MEMORY, "bolt_column_to_arrow: host allocation");
  arr->blocks.push_back(vals);
  if (n) BOLT_HIP(hipMemcpyAsync(vals, col->data, (size_t)n * w, hipMemcpyDeviceToHost, s));
  uint8_t* bits = nullptr;
  int64_t nulls = 0;
  if (col->validity && n) {
    bits = static_cast<uint8_t*>(calloc((size_t)(n + 7) / 8 + 1, 1));
    if (!bits) return set_error(BOLT_ERR_MEMORY, "bolt_column_to_arrow: host allocation");
    arr->blocks.push_back(bits);
    BOLT_HIP(hipMemcpyAsync(bits, col->validity, (size_t)(n + 7) / 8, hipMemcpyDeviceToHost, s));
  }
  BOLT_HIP(hipStreamSynchronize(s));
  if (bits) {
    if (n & 7) bits[n >> 3] &= (uint8_t)((1u << (n & 7)) - 1);  // bits past the end are zero
    for (int64_t i = 0; i < (n + 7) / 8; ++i) nulls += 8 - __builtin_popcount(bits[i]);
    nulls -= (8 - (n & 7)) & 7;
  }
  if (col->dtype == BOLT_BOOL) {  // one byte per row -> Arrow's packed bits
    uint8_t* packed = static_cast<uint8_t*>(calloc((size_t)(n + 7) / 8 + 1, 1));
    if (!packed) return set_error(BOLT_ERR_MEMORY, "bolt_column_to_arrow: host allocation");
    const uint8_t* b = static_cast<const uint8_t*>(vals);
    for (int64_t i = 0; i < n; ++i)
      if (b[i]) packed[i >> 3] |= (uint8_t)(1u << (i & 7));
    free(vals);
    arr->blocks[0] = vals = packed;
  }
  arr->buffers[0] = nulls ? bits : nullptr;
  arr->buffers[1] = vals;
  sch->format = fmt;
  sch->name = name ? name : "";
  memset(out_array, 0, sizeof *out_array);
  out_array->length = n;
  out_array->null_count = nulls;
  out_array->n_buffers = 2;
  out_array->buffers = arr->buffers;
  out_array->release = release_array;
  out_array->private_data = arr.release();
  memset(out_schema, 0, sizeof *out_schema);
  out_schema->format = sch->format.c_str();
  out_schema->name = sch->name.c_str();
  out_schema->flags = ARROW_FLAG_NULLABLE;
  out_schema->release = release_schema;
  out_schema->private_data = sch.release();
  return BOLT_OK;
}

}  // extern "C"
