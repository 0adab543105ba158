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

__host__ __device__ __forceinline__ double as_f64_raw(uint64_t u) {
  double d;
  memcpy(&d, &u, 8);
  return d;
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
    case BOLT_F64: vi = 0; vf = as_f64_raw(r.lo); break;
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

// ---------------------------------------------------------------------------
// Exact accumulators, three 64-bit words each:
//   integer / decimal SUM : a,b = 128-bit two's-complement (lo, hi); bit 63 of
//                           n flags a carry out of bit 127.  Integer SUM is
//                           range-checked against i64 at finalize, so the
//                           overflow error does not depend on row order;
//   float SUM / AVG       : a,b = Neumaier pair (s, c) in f64, merged with
//                           TwoSum;
//   MIN / MAX             : a = an order-preserving u64 image, so the merge
//                           is one unsigned min / max (ints: sign flip;
//                           floats: the NaN-largest total order of
//                           src/exec/aggregate.rs:2165);
//   VAR / STDDEV          : a,b = Welford (mean, M2), merged with Chan's
//                           formula (src/exec/welford.rs:84);
//   n                     : values seen (COUNT).
// ---------------------------------------------------------------------------
struct Acc {
  uint64_t a;
  uint64_t b;
  uint64_t n;  // count in bits 0..62, overflow flag in bit 63
};
constexpr uint64_t kOvfBit = 1ull << 63;

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

__host__ __device__ __forceinline__ uint64_t mm_identity(int op) { return op == BOLT_AGG_MIN ? ~0ull : 0ull; }
__host__ __device__ __forceinline__ double as_f64(uint64_t u) { double d; memcpy(&d, &u, 8); return d; }
__host__ __device__ __forceinline__ uint64_t as_u64(double d) { uint64_t u; memcpy(&u, &d, 8); return u; }
__host__ __device__ __forceinline__ uint64_t acc_count(const Acc& a) { return a.n & ~kOvfBit; }

__device__ __forceinline__ void acc_init(Acc& a, int op) {
  a.a = (op == BOLT_AGG_MIN || op == BOLT_AGG_MAX) ? mm_identity(op) : 0;
  a.b = 0;
  a.n = 0;
}

__device__ __forceinline__ void add128(Acc& a, uint64_t lo, int64_t hi) {
  uint64_t nlo = a.a + lo;
  uint64_t carry = nlo < a.a;
  int64_t t, h;
  // signed 128-bit overflow: exactly one of the two partial adds of the high
  // word overflows
  bool o1 = __builtin_add_overflow((int64_t)a.b, hi, &t);
  bool o2 = __builtin_add_overflow(t, (int64_t)carry, &h);
  a.n |= (uint64_t)(o1 ^ o2) << 63;
  a.a = nlo;
  a.b = (uint64_t)h;
}

__device__ __forceinline__ void neum_add(Acc& a, double v) {
  double s = as_f64(a.a), c = as_f64(a.b);
  double t = s + v;
  if (fabs(s) >= fabs(v)) c += (s - t) + v;
  else c += (v - t) + s;
  a.a = as_u64(t);
  a.b = as_u64(c);
}

__device__ __forceinline__ void welford_push(Acc& a, double x) {
  a.n += 1;
  double mean = as_f64(a.a), m2 = as_f64(a.b);
  double d = x - mean;
  mean += d / (double)acc_count(a);
  m2 += d * (x - mean);
  a.a = as_u64(mean);
  a.b = as_u64(m2);
}

// Fold one (non-NULL) value into `a` for aggregate `op`: `v_i` is the
// integer value (decimal low word), `v_hi` the decimal high word, `v_f` the
// value as f64.
__device__ __forceinline__ void acc_update(Acc& a, int op, int dtype, int64_t v_i, int64_t v_hi, double v_f) {
  switch (op) {
    case BOLT_AGG_SUM:
      if (dtype == BOLT_F32 || dtype == BOLT_F64) neum_add(a, v_f);
      else add128(a, (uint64_t)v_i, dtype == BOLT_DEC128 ? v_hi : (v_i >> 63));
      a.n++;
      break;
    case BOLT_AGG_AVG: neum_add(a, v_f); a.n++; break;
    case BOLT_AGG_COUNT: case BOLT_AGG_COUNT_STAR: a.n++; break;
    case BOLT_AGG_MIN: case BOLT_AGG_MAX: {
      uint64_t k;
      if (dtype == BOLT_F32 || dtype == BOLT_F64) k = f64_order_key(v_f);
      else k = (uint64_t)v_i ^ 0x8000000000000000ull;
      a.a = op == BOLT_AGG_MIN ? (k < a.a ? k : a.a) : (k > a.a ? k : a.a);
      a.n++;
      break;
    }
    default: welford_push(a, v_f); break;
  }
}

// Merge `b` into `a` (both for aggregate `op`).
__device__ __forceinline__ void acc_merge(Acc& a, const Acc& b, int op, int dtype) {
  switch (op) {
    case BOLT_AGG_SUM: case BOLT_AGG_AVG:
      if (op == BOLT_AGG_SUM && !(dtype == BOLT_F32 || dtype == BOLT_F64)) {
        add128(a, b.a, (int64_t)b.b);
        a.n = ((a.n | b.n) & kOvfBit) | (acc_count(a) + acc_count(b));
      } else {
        // TwoSum of the leading parts, compensations added
        double as = as_f64(a.a), bs = as_f64(b.a);
        double s = as + bs;
        double bp = s - as;
        double err = (as - (s - bp)) + (bs - bp);
        a.b = as_u64(as_f64(a.b) + as_f64(b.b) + err);
        a.a = as_u64(s);
        a.n += b.n;
      }
      break;
    case BOLT_AGG_COUNT: case BOLT_AGG_COUNT_STAR: a.n += b.n; break;
    case BOLT_AGG_MIN: a.a = b.a < a.a ? b.a : a.a; a.n += b.n; break;
    case BOLT_AGG_MAX: a.a = b.a > a.a ? b.a : a.a; a.n += b.n; break;
    default: {
      uint64_t nb_ = acc_count(b), na_ = acc_count(a);
      if (nb_ == 0) break;
      if (na_ == 0) { a = b; break; }
      double na = (double)na_, nb = (double)nb_, n = na + nb;
      double ma = as_f64(a.a), mb = as_f64(b.a);
      double d = mb - ma;
      a.a = as_u64(ma + d * (nb / n));
      a.b = as_u64(as_f64(a.b) + as_f64(b.b) + d * d * (na * nb / n));
      a.n += b.n;
      break;
    }
  }
}

__device__ __forceinline__ Acc shfl_xor_acc(const Acc& a, int m) {
  Acc r;
  r.a = __shfl_xor(a.a, m);
  r.b = __shfl_xor(a.b, m);
  r.n = __shfl_xor(a.n, m);
  return r;
}

// Reduce `a` across the 64 lanes of the wave (result valid in every lane).
__device__ __forceinline__ void wave_reduce_acc(Acc& a, int op, int dtype) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    Acc b = shfl_xor_acc(a, m);
    acc_merge(a, b, op, dtype);
  }
}

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
