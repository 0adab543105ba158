// device_acc.h — exact aggregate accumulators and small device helpers.
//
// Self-contained (no includes, no standard-library types) so that the same
// text is compiled into the library by hipcc and prepended to the kernels
// that the library generates per query and compiles with hipRTC.
#pragma once

namespace bolt {

// dtype and aggregate codes (equal to the C ABI's bolt_dtype / bolt_agg_op)
enum : int { DT_BOOL = 1, DT_I32 = 2, DT_I64 = 3, DT_F32 = 4, DT_F64 = 5, DT_DEC128 = 6, DT_DATE32 = 7 };
enum : int {
  AG_SUM = 1, AG_COUNT = 2, AG_COUNT_STAR = 3, AG_MIN = 4, AG_MAX = 5, AG_AVG = 6,
  AG_VAR_POP = 7, AG_VAR_SAMP = 8, AG_STDDEV_POP = 9, AG_STDDEV_SAMP = 10
};

typedef unsigned long long u64_t;
typedef long long i64_t;
typedef unsigned int u32_t;

__host__ __device__ __forceinline__ double as_f64(u64_t u) {
  double d;
  __builtin_memcpy(&d, &u, 8);
  return d;
}
__host__ __device__ __forceinline__ u64_t as_u64(double d) {
  u64_t u;
  __builtin_memcpy(&u, &d, 8);
  return u;
}

// Order-preserving map of a double to u64 under the NaN-largest total order
// of src/exec/aggregate.rs:2165 (all NaNs -> one top value, -0.0 < +0.0).
// Native unsigned 64-bit atomic min/max on the image implement float
// MIN/MAX without a CAS loop.
__host__ __device__ __forceinline__ u64_t f64_order_key(double d) {
  u64_t b = as_u64(d);
  if ((b & 0x7FFFFFFFFFFFFFFFull) > 0x7FF0000000000000ull) return ~0ull;  // any NaN -> top
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ __forceinline__ double f64_from_order_key(u64_t k) {
  u64_t b = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
  if (k == ~0ull) b = 0x7FF8000000000000ull;
  return as_f64(b);
}

__host__ __device__ __forceinline__ u64_t fmix64(u64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}

// Join-side Bloom filter block test: the 128-bit block must have one bit set
// in each 32-bit word at the positions given by four 5-bit fields of h.
__device__ __forceinline__ uint4 bloom_pattern(u64_t h) {
  return make_uint4(1u << ((h >> 20) & 31), 1u << ((h >> 25) & 31), 1u << ((h >> 30) & 31), 1u << ((h >> 35) & 31));
}
__device__ __forceinline__ bool bloom_hit(const uint4& b, u64_t h) {
  const uint4 p = bloom_pattern(h);
  return (b.x & p.x) && (b.y & p.y) && (b.z & p.z) && (b.w & p.w);
}

__device__ __forceinline__ bool vbit(const unsigned char* validity, i64_t i) {
  return validity == nullptr || ((validity[i >> 3] >> (i & 7)) & 1);
}

// ---------------------------------------------------------------------------
// Accumulators, three 64-bit words each:
//   integer / decimal SUM : a,b = 128-bit two's-complement (lo, hi); bit 63 of
//                           n flags a carry out of bit 127.  Integer SUM is
//                           range-checked against i64 at finalize, so the
//                           overflow error does not depend on row order;
//   float SUM / AVG       : a,b = Neumaier pair (s, c) in f64, merged with
//                           TwoSum;
//   MIN / MAX             : a = the order-preserving u64 image, merged with
//                           one unsigned min / max (ints: sign flip; floats:
//                           f64_order_key);
//   VAR / STDDEV          : a,b = Welford (mean, M2), merged with Chan's
//                           formula (src/exec/welford.rs:84);
//   n                     : values seen (COUNT).
// ---------------------------------------------------------------------------
struct Acc {
  u64_t a;
  u64_t b;
  u64_t n;  // count in bits 0..62, overflow flag in bit 63
};
constexpr u64_t kOvfBit = 1ull << 63;

__host__ __device__ __forceinline__ u64_t mm_identity(int op) { return op == AG_MIN ? ~0ull : 0ull; }
__host__ __device__ __forceinline__ u64_t acc_count(const Acc& a) { return a.n & ~kOvfBit; }
__host__ __device__ __forceinline__ bool is_float_dt(int dt) { return dt == DT_F32 || dt == DT_F64; }

__device__ __forceinline__ void acc_init(Acc& a, int op) {
  a.a = (op == AG_MIN || op == AG_MAX) ? mm_identity(op) : 0;
  a.b = 0;
  a.n = 0;
}

__device__ __forceinline__ void add128(Acc& a, u64_t lo, i64_t hi) {
  u64_t nlo = a.a + lo;
  u64_t carry = nlo < a.a;
  i64_t t, h;
  // signed 128-bit overflow: exactly one of the two partial adds of the high
  // word overflows
  bool o1 = __builtin_add_overflow((i64_t)a.b, hi, &t);
  bool o2 = __builtin_add_overflow(t, (i64_t)carry, &h);
  a.n |= (u64_t)(o1 ^ o2) << 63;
  a.a = nlo;
  a.b = (u64_t)h;
}

__device__ __forceinline__ void neum_add(Acc& a, double v) {
  double s = as_f64(a.a), c = as_f64(a.b);
  double t = s + v;
  if (__builtin_fabs(s) >= __builtin_fabs(v)) c += (s - t) + v;
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

__device__ __forceinline__ void minmax_push(Acc& a, bool is_min, u64_t k) {
  a.a = is_min ? (k < a.a ? k : a.a) : (k > a.a ? k : a.a);
  a.n++;
}

// Fold one (non-NULL) value into `a` for aggregate `op`: `v_i` is the integer
// value (decimal low word), `v_hi` the decimal high word, `v_f` the value as
// f64.
__device__ __forceinline__ void acc_update(Acc& a, int op, int dtype, i64_t v_i, i64_t v_hi, double v_f) {
  switch (op) {
    case AG_SUM:
      if (is_float_dt(dtype)) neum_add(a, v_f);
      else add128(a, (u64_t)v_i, dtype == DT_DEC128 ? v_hi : (v_i >> 63));
      a.n++;
      break;
    case AG_AVG: neum_add(a, v_f); a.n++; break;
    case AG_COUNT: case AG_COUNT_STAR: a.n++; break;
    case AG_MIN: case AG_MAX:
      minmax_push(a, op == AG_MIN, is_float_dt(dtype) ? f64_order_key(v_f) : ((u64_t)v_i ^ 0x8000000000000000ull));
      break;
    default: welford_push(a, v_f); break;
  }
}

// Merge `b` into `a` (both for aggregate `op` over `dtype`).
__device__ __forceinline__ void acc_merge(Acc& a, const Acc& b, int op, int dtype) {
  switch (op) {
    case AG_SUM: case AG_AVG:
      if (op == AG_SUM && !is_float_dt(dtype)) {
        add128(a, b.a, (i64_t)b.b);
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
    case AG_COUNT: case AG_COUNT_STAR: a.n += b.n; break;
    case AG_MIN: a.a = b.a < a.a ? b.a : a.a; a.n += b.n; break;
    case AG_MAX: a.a = b.a > a.a ? b.a : a.a; a.n += b.n; break;
    default: {
      u64_t nb_ = acc_count(b), na_ = acc_count(a);
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

// Reduce across a block of `waves` wave64s through `sh` (>= waves entries of
// LDS); thread 0 stores the block result to *out.  Must be called by every
// thread of the block.
__device__ __forceinline__ void block_reduce_acc(Acc a, int op, int dtype, Acc* sh, int waves, Acc* out) {
  wave_reduce_acc(a, op, dtype);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) sh[wave] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    Acc r = sh[0];
    for (int w = 1; w < waves; w++) acc_merge(r, sh[w], op, dtype);
    *out = r;
  }
  __syncthreads();
}

// Literal values of a generated kernel's predicate terms (passed by value, so
// the generated code does not change with the literals).
struct JitLits {
  i64_t i[8];
  double f[8];
  i64_t hi[8];
};

// Column pointers of a generated kernel (by value).
constexpr int kJitMaxCols = 16;
struct JitCols {
  const void* d[kJitMaxCols];
  const unsigned char* v[kJitMaxCols];
};

// 128-bit signed compare of (hi, lo) with (lhi, llo): -1, 0, 1
__device__ __forceinline__ int cmp128(i64_t hi, u64_t lo, i64_t lhi, u64_t llo) {
  return hi < lhi ? -1 : hi > lhi ? 1 : (lo < llo ? -1 : lo > llo ? 1 : 0);
}

}  // namespace bolt
