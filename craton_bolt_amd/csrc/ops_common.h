// ops_common.h — device-side pieces shared by the operator kernels:
// predicate evaluation, exact aggregate accumulators, wave64 / block
// reductions and the host-side validation of operator arguments.
#pragma once

#include "bolt_internal.h"

namespace bolt {

constexpr int kMaxAggs = 16;

// ---------------------------------------------------------------------------
// Predicates: an AND of column-vs-literal terms.  A comparison with a NULL
// value is false (src/jit/scan_kernel.rs:265 gates stores on validity).
// Floats compare in IEEE order, so every ordered comparison with NaN is false.
// ---------------------------------------------------------------------------
struct DevTerm {
  const void* data;
  const uint8_t* validity;
  int32_t dtype;
  int32_t op;
  int64_t lit_i64;
  double lit_f64;
  int64_t lit_hi;
};
struct DevFilter {
  DevTerm t[BOLT_MAX_PRED_TERMS];
  int32_t n;
};

__device__ __forceinline__ bool cmp_apply(int op, int c) {
  switch (op) {
    case BOLT_CMP_LT: return c < 0;
    case BOLT_CMP_LE: return c <= 0;
    case BOLT_CMP_GT: return c > 0;
    case BOLT_CMP_GE: return c >= 0;
    case BOLT_CMP_EQ: return c == 0;
    case BOLT_CMP_NE: return c != 0;
  }
  return false;
}

__device__ __forceinline__ bool eval_term(const DevTerm& t, int64_t row) {
  bool valid = bit_valid(t.validity, row);
  if (t.op == BOLT_CMP_IS_NULL) return !valid;
  if (t.op == BOLT_CMP_IS_NOT_NULL) return valid;
  if (!valid) return false;
  switch (t.dtype) {
    case BOLT_F64:
    case BOLT_F32: {
      double v = t.dtype == BOLT_F64 ? static_cast<const double*>(t.data)[row]
                                     : (double)static_cast<const float*>(t.data)[row];
      double l = t.lit_f64;
      switch (t.op) {
        case BOLT_CMP_LT: return v < l;
        case BOLT_CMP_LE: return v <= l;
        case BOLT_CMP_GT: return v > l;
        case BOLT_CMP_GE: return v >= l;
        case BOLT_CMP_EQ: return v == l;
        case BOLT_CMP_NE: return v != l;
      }
      return false;
    }
    case BOLT_DEC128: {
      const int64_t* w = static_cast<const int64_t*>(t.data) + 2 * row;
      int64_t hi = w[1];
      uint64_t lo = (uint64_t)w[0];
      int c = hi < t.lit_hi ? -1 : hi > t.lit_hi ? 1 : (lo < (uint64_t)t.lit_i64 ? -1 : lo > (uint64_t)t.lit_i64 ? 1 : 0);
      return cmp_apply(t.op, c);
    }
    default: {
      int64_t v;
      if (t.dtype == BOLT_I64) v = static_cast<const int64_t*>(t.data)[row];
      else if (t.dtype == BOLT_BOOL) v = static_cast<const uint8_t*>(t.data)[row];
      else v = static_cast<const int32_t*>(t.data)[row];
      int c = v < t.lit_i64 ? -1 : (v > t.lit_i64 ? 1 : 0);
      return cmp_apply(t.op, c);
    }
  }
}

__device__ __forceinline__ bool eval_filter(const DevFilter& f, int64_t row) {
  for (int i = 0; i < f.n; i++)
    if (!eval_term(f.t[i], row)) return false;
  return true;
}

// Raw 8- or 16-byte image of one column value, loaded once per row so that
// the predicate and the aggregates read registers, not memory.
struct Raw {
  uint64_t lo;
  int64_t hi;
};

__device__ __forceinline__ Raw load_raw(const void* p, int dtype, int64_t row) {
  Raw r;
  r.hi = 0;
  switch (dtype) {
    case BOLT_I64: case BOLT_F64: r.lo = __builtin_nontemporal_load(static_cast<const uint64_t*>(p) + row); break;
    case BOLT_I32: case BOLT_F32: case BOLT_DATE32:
      r.lo = __builtin_nontemporal_load(static_cast<const uint32_t*>(p) + row);
      break;
    case BOLT_BOOL: r.lo = static_cast<const uint8_t*>(p)[row]; break;
    case BOLT_DEC128: {
      const uint64_t* w = static_cast<const uint64_t*>(p) + 2 * row;
      r.lo = w[0];
      r.hi = (int64_t)w[1];
      break;
    }
    default: r.lo = 0;
  }
  return r;
}

// Decode a raw image as (int, decimal-high, float).
__device__ __forceinline__ void raw_value(Raw r, int dtype, int64_t& vi, int64_t& vhi, double& vf) {
  vhi = 0;
  switch (dtype) {
    case BOLT_I32: case BOLT_DATE32: vi = (int64_t)(int32_t)(uint32_t)r.lo; vf = (double)vi; break;
    case BOLT_I64: vi = (int64_t)r.lo; vf = (double)vi; break;
    case BOLT_F32: vi = 0; vf = (double)__uint_as_float((uint32_t)r.lo); break;
    case BOLT_F64: vi = 0; vf = as_f64(r.lo); break;
    case BOLT_BOOL: vi = (int64_t)r.lo; vf = (double)vi; break;
    case BOLT_DEC128: vi = (int64_t)r.lo; vhi = r.hi; vf = 0; break;
    default: vi = 0; vf = 0;
  }
}

__device__ __forceinline__ bool eval_term_raw(const DevTerm& t, Raw r, bool valid) {
  if (t.op == BOLT_CMP_IS_NULL) return !valid;
  if (t.op == BOLT_CMP_IS_NOT_NULL) return valid;
  if (!valid) return false;
  int64_t vi, vhi;
  double vf;
  raw_value(r, t.dtype, vi, vhi, vf);
  if (t.dtype == BOLT_F32 || t.dtype == BOLT_F64) {
    double l = t.lit_f64;
    switch (t.op) {
      case BOLT_CMP_LT: return vf < l;
      case BOLT_CMP_LE: return vf <= l;
      case BOLT_CMP_GT: return vf > l;
      case BOLT_CMP_GE: return vf >= l;
      case BOLT_CMP_EQ: return vf == l;
      case BOLT_CMP_NE: return vf != l;
    }
    return false;
  }
  int c;
  if (t.dtype == BOLT_DEC128)
    c = vhi < t.lit_hi ? -1 : vhi > t.lit_hi ? 1 : ((uint64_t)vi < (uint64_t)t.lit_i64 ? -1 : (uint64_t)vi > (uint64_t)t.lit_i64 ? 1 : 0);
  else
    c = vi < t.lit_i64 ? -1 : (vi > t.lit_i64 ? 1 : 0);
  return cmp_apply(t.op, c);
}

// Evaluate the first NT terms (compile-time count) of `f` at `row`; with
// NT < 0 the runtime count f.n is used.
template <int NT>
__device__ __forceinline__ bool eval_filter_nt(const DevFilter& f, int64_t row) {
  if constexpr (NT < 0) {
    return eval_filter(f, row);
  } else {
    Raw r[NT > 0 ? NT : 1];
#pragma unroll
    for (int t = 0; t < NT; t++) r[t] = load_raw(f.t[t].data, f.t[t].dtype, row);
    bool ok = true;
#pragma unroll
    for (int t = 0; t < NT; t++) ok = ok && eval_term_raw(f.t[t], r[t], bit_valid(f.t[t].validity, row));
    return ok;
  }
}

// Host: validate a bolt_filter_t against `cols` and build the kernel form.
int make_dev_filter(const bolt_filter_t* f, const bolt_column_t* cols, int32_t n_cols, int64_t n_rows,
                    DevFilter* out);

// Accumulators (Acc, acc_update, acc_merge, wave_reduce_acc, ...) live in
// device_acc.h, shared with the kernels generated per query.
static_assert(AG_SUM == BOLT_AGG_SUM && AG_STDDEV_SAMP == BOLT_AGG_STDDEV_SAMP && AG_COUNT_STAR == BOLT_AGG_COUNT_STAR,
              "aggregate codes");
static_assert(DT_I32 == BOLT_I32 && DT_F64 == BOLT_F64 && DT_DEC128 == BOLT_DEC128 && DT_DATE32 == BOLT_DATE32,
              "dtype codes");

struct AggDesc {
  const void* data;
  const uint8_t* validity;
  int32_t op;
  int32_t dtype;
  int32_t src; // 0 = column, 1 = join carry
  int32_t pad;
};
struct AggList {
  AggDesc a[kMaxAggs];
  int32_t n;
};

// Host-side: turn a final Acc into a scalar result following the output
// dtype rules of src/plan/logical_plan.rs:2126,2191.
int acc_finalize(const Acc& a, int op, int in_dtype, bolt_scalar_t* out);

// Host: validate the aggregate list against the input columns and build the
// kernel form.  `carry_dtype` is the join carry's dtype (0 when none).
int make_agg_list(const bolt_agg_t* aggs, int32_t n_aggs, const bolt_column_t* cols, int32_t n_cols,
                  int32_t carry_dtype, AggList* out);

// Read column value `row` of an aggregate input as (int, hi, float).
__device__ __forceinline__ void load_value(const void* data, int dtype, int64_t row, int64_t& vi, int64_t& vhi,
                                           double& vf) {
  switch (dtype) {
    case BOLT_I32: case BOLT_DATE32: vi = static_cast<const int32_t*>(data)[row]; vf = (double)vi; vhi = 0; break;
    case BOLT_I64: vi = static_cast<const int64_t*>(data)[row]; vf = (double)vi; vhi = 0; break;
    case BOLT_F32: vf = (double)static_cast<const float*>(data)[row]; vi = 0; vhi = 0; break;
    case BOLT_F64: vf = static_cast<const double*>(data)[row]; vi = 0; vhi = 0; break;
    case BOLT_BOOL: vi = static_cast<const uint8_t*>(data)[row]; vf = (double)vi; vhi = 0; break;
    case BOLT_DEC128: {
      const int64_t* w = static_cast<const int64_t*>(data) + 2 * row;
      vi = w[0]; vhi = w[1];
      vf = 0;
      break;
    }
    default: vi = 0; vhi = 0; vf = 0;
  }
}

// Decode the 4-byte join carry word as (int, float) for its dtype.
__device__ __forceinline__ void carry_value(uint32_t w, int dtype, int64_t& vi, double& vf) {
  if (dtype == BOLT_F32) { vf = (double)__uint_as_float(w); vi = 0; }
  else { vi = (int64_t)(int32_t)w; vf = (double)vi; }
}

}  // namespace bolt
