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

// Host: validate a bolt_filter_t against `cols` and build the kernel form.
int make_dev_filter(const bolt_filter_t* f, const bolt_column_t* cols, int32_t n_cols, int64_t n_rows,
                    DevFilter* out);

// ---------------------------------------------------------------------------
// Exact accumulators.
//   integer / decimal SUM : 128-bit two's-complement (lo, hi) + an overflow
//                           flag for carries out of bit 127; integer SUM is
//                           range-checked against i64 at finalize, so the
//                           overflow error does not depend on row order;
//   float SUM / AVG       : Neumaier pair (s, c) in f64, merged with TwoSum;
//   MIN / MAX             : an order-preserving u64 image, so the merge is
//                           one unsigned min / max (ints: sign flip; floats:
//                           the NaN-largest total order of
//                           src/exec/aggregate.rs:2165);
//   VAR / STDDEV          : Welford (n, mean, M2), merged with Chan's formula
//                           (src/exec/welford.rs:84);
//   COUNT                 : u64.
// ---------------------------------------------------------------------------
struct Acc {
  uint64_t lo;
  int64_t hi;
  double s;   // float sum / Welford mean
  double c;   // Neumaier compensation / Welford M2
  uint64_t n; // values seen
  uint64_t mm;
  uint32_t ovf;
  uint32_t pad;
};

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

__device__ __forceinline__ void acc_init(Acc& a, int op) {
  a.lo = 0; a.hi = 0; a.s = 0; a.c = 0; a.n = 0; a.ovf = 0; a.pad = 0;
  a.mm = mm_identity(op);
}

__device__ __forceinline__ void add128(Acc& a, uint64_t lo, int64_t hi) {
  uint64_t nlo = a.lo + lo;
  uint64_t carry = nlo < a.lo;
  int64_t h;
  // overflow of the signed 128-bit sum: operands of equal sign, result of the other
  int64_t t;
  bool o1 = __builtin_add_overflow(a.hi, hi, &t);
  bool o2 = __builtin_add_overflow(t, (int64_t)carry, &h);
  a.ovf |= (uint32_t)(o1 ^ o2);
  a.lo = nlo;
  a.hi = h;
}

__device__ __forceinline__ void neum_add(Acc& a, double v) {
  double t = a.s + v;
  if (fabs(a.s) >= fabs(v)) a.c += (a.s - t) + v;
  else a.c += (v - t) + a.s;
  a.s = t;
}

__device__ __forceinline__ void welford_push(Acc& a, double x) {
  a.n += 1;
  double d = x - a.s;
  a.s += d / (double)a.n;
  a.c += d * (x - a.s);
}

// Fold one value (already known non-NULL) into `a` for aggregate `op`.
// `v_i` is the integer value (or decimal low word), `v_hi` the decimal high
// word, `v_f` the float value, `is_float` selects the float form.
__device__ __forceinline__ void acc_update(Acc& a, int op, int dtype, int64_t v_i, int64_t v_hi, double v_f) {
  switch (op) {
    case BOLT_AGG_SUM:
      if (dtype == BOLT_F32 || dtype == BOLT_F64) { neum_add(a, v_f); a.n++; }
      else { add128(a, (uint64_t)v_i, dtype == BOLT_DEC128 ? v_hi : (v_i >> 63)); a.n++; }
      break;
    case BOLT_AGG_AVG: neum_add(a, v_f); a.n++; break;
    case BOLT_AGG_COUNT: case BOLT_AGG_COUNT_STAR: a.n++; break;
    case BOLT_AGG_MIN: case BOLT_AGG_MAX: {
      uint64_t k;
      if (dtype == BOLT_F32 || dtype == BOLT_F64) k = f64_order_key(v_f);
      else k = (uint64_t)v_i ^ 0x8000000000000000ull;
      a.mm = op == BOLT_AGG_MIN ? (k < a.mm ? k : a.mm) : (k > a.mm ? k : a.mm);
      a.n++;
      break;
    }
    default: welford_push(a, v_f); break;
  }
}

// Merge `b` into `a` (both for aggregate `op`).
__device__ __forceinline__ void acc_merge(Acc& a, const Acc& b, int op) {
  switch (op) {
    case BOLT_AGG_SUM: case BOLT_AGG_AVG: {
      add128(a, b.lo, b.hi);
      a.ovf |= b.ovf;
      // TwoSum of the leading parts, compensations added
      double s = a.s + b.s;
      double bp = s - a.s;
      double err = (a.s - (s - bp)) + (b.s - bp);
      a.c = a.c + b.c + err;
      a.s = s;
      a.n += b.n;
      break;
    }
    case BOLT_AGG_COUNT: case BOLT_AGG_COUNT_STAR: a.n += b.n; break;
    case BOLT_AGG_MIN: a.mm = b.mm < a.mm ? b.mm : a.mm; a.n += b.n; break;
    case BOLT_AGG_MAX: a.mm = b.mm > a.mm ? b.mm : a.mm; a.n += b.n; break;
    default: {
      if (b.n == 0) break;
      if (a.n == 0) { a = b; break; }
      double na = (double)a.n, nb = (double)b.n, n = na + nb;
      double d = b.s - a.s;
      a.s = a.s + d * (nb / n);
      a.c = a.c + b.c + d * d * (na * nb / n);
      a.n += b.n;
      break;
    }
  }
}

__device__ __forceinline__ Acc shfl_xor_acc(const Acc& a, int m) {
  Acc r;
  r.lo = __shfl_xor(a.lo, m);
  r.hi = __shfl_xor(a.hi, m);
  r.s = __shfl_xor(a.s, m);
  r.c = __shfl_xor(a.c, m);
  r.n = __shfl_xor(a.n, m);
  r.mm = __shfl_xor(a.mm, m);
  r.ovf = __shfl_xor(a.ovf, m);
  r.pad = 0;
  return r;
}

// Reduce `a` across the 64 lanes of the wave (result valid in every lane).
__device__ __forceinline__ void wave_reduce_acc(Acc& a, int op) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    Acc b = shfl_xor_acc(a, m);
    acc_merge(a, b, op);
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
