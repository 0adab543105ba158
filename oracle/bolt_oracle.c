/*
 * bolt_oracle.c — CPU oracle for the GPU relational-operator path.
 *
 * TEST INFRASTRUCTURE ONLY (see bolt_oracle.h).  Single-threaded restatements
 * of the reference's host algorithms; each function cites the reference
 * file:line it follows.
 */
#include "bolt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef __int128 i128;

/* ------------------------------------------------------------------------ */
/* Xorshift64Star — tests/common/mod.rs:46-90                               */
/* ------------------------------------------------------------------------ */

uint64_t bo_rng_seed(uint64_t seed) {
  /* tests/common/mod.rs:57-62: golden-ratio mix, zero guard */
  uint64_t s = seed + 0x9E3779B97F4A7C15ull;
  if (s == 0) s = 0xDEADBEEFCAFEBABEull;
  return s;
}

uint64_t bo_rng_next(uint64_t* state) {
  /* tests/common/mod.rs:67-74: shifts 12/25/27 then the odd multiplier */
  uint64_t x = *state;
  x ^= x >> 12;
  x ^= x << 25;
  x ^= x >> 27;
  *state = x;
  return x * 0x2545F4914F6CDD1Dull;
}

double bo_rng_unit(uint64_t* state) {
  /* tests/common/mod.rs:80-83: top 53 bits scaled by 2^-53 */
  uint64_t r = bo_rng_next(state);
  return (double)(r >> 11) * (1.0 / 9007199254740992.0);
}

void bo_rng_fill(uint64_t seed, uint64_t* out, int64_t n) {
  uint64_t s = bo_rng_seed(seed);
  for (int64_t i = 0; i < n; i++) out[i] = bo_rng_next(&s);
}

void bo_shuffle_u32(uint32_t* xs, int64_t n, uint64_t seed) {
  /* tests/common/mod.rs:102-110: Knuth LCG, j = s mod (i+1), i from n-1 down */
  uint64_t s = seed;
  for (int64_t i = n - 1; i >= 1; i--) {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    uint64_t j = s % (uint64_t)(i + 1);
    uint32_t t = xs[i];
    xs[i] = xs[j];
    xs[j] = t;
  }
}

/* ------------------------------------------------------------------------ */
/* Neumaier — src/exec/aggregate.rs:2001-2028                               */
/* ------------------------------------------------------------------------ */

typedef struct { double sum, c, naive; int nonfinite; } neum_t;

static void neum_add(neum_t* a, double v) {
  if (!isfinite(v)) a->nonfinite = 1;
  a->naive += v;
  double t = a->sum + v;
  if (fabs(a->sum) >= fabs(v)) a->c += (a->sum - t) + v;
  else a->c += (v - t) + a->sum;
  a->sum = t;
}

static double neum_result(const neum_t* a) { return a->nonfinite ? a->naive : a->sum + a->c; }

double bo_neumaier_sum(const double* x, int64_t n) {
  neum_t a = {0, 0, 0, 0};
  for (int64_t i = 0; i < n; i++) neum_add(&a, x[i]);
  return neum_result(&a);
}

/* ------------------------------------------------------------------------ */
/* column access helpers                                                    */
/* ------------------------------------------------------------------------ */

static int is_valid(const bo_col* c, int64_t i) {
  return c->validity == NULL || ((c->validity[i >> 3] >> (i & 7)) & 1);
}
static int is_int_dtype(int32_t dt) { return dt == BO_I32 || dt == BO_I64 || dt == BO_DATE32; }
static int is_float_dtype(int32_t dt) { return dt == BO_F32 || dt == BO_F64; }

static int64_t get_i64(const bo_col* c, int64_t i) {
  switch (c->dtype) {
    case BO_I32: case BO_DATE32: return ((const int32_t*)c->data)[i];
    case BO_I64: return ((const int64_t*)c->data)[i];
    case BO_BOOL: return ((const uint8_t*)c->data)[i];
    default: return 0;
  }
}
static double get_f64(const bo_col* c, int64_t i) {
  switch (c->dtype) {
    case BO_F32: return (double)((const float*)c->data)[i];
    case BO_F64: return ((const double*)c->data)[i];
    default: return (double)get_i64(c, i);
  }
}
static i128 get_i128(const bo_col* c, int64_t i) {
  if (c->dtype == BO_DEC128) {
    const uint64_t* w = (const uint64_t*)c->data + 2 * i;
    return (i128)(((unsigned __int128)w[1] << 64) | w[0]);
  }
  return (i128)get_i64(c, i);
}

/* float_total_cmp — src/exec/aggregate.rs:2165: NaN is the largest value and
 * all NaNs compare equal; otherwise IEEE total order (-0 < +0). */
static int total_cmp_f64(double a, double b) {
  int an = isnan(a), bn = isnan(b);
  if (an && bn) return 0;
  if (an) return 1;
  if (bn) return -1;
  int64_t ia, ib;
  memcpy(&ia, &a, 8);
  memcpy(&ib, &b, 8);
  ia ^= (int64_t)(((uint64_t)(ia >> 63)) >> 1);
  ib ^= (int64_t)(((uint64_t)(ib >> 63)) >> 1);
  return ia < ib ? -1 : (ia > ib ? 1 : 0);
}

/* ------------------------------------------------------------------------ */
/* Welford — src/exec/welford.rs:47-130                                     */
/* ------------------------------------------------------------------------ */

typedef struct { uint64_t count; double mean, m2; } welford_t;
static void welford_push(welford_t* w, double x) {
  w->count += 1;
  double delta = x - w->mean;
  w->mean += delta / (double)w->count;
  double delta2 = x - w->mean;
  w->m2 += delta * delta2;
}
/* returns 1 and writes *out when defined, 0 for SQL NULL */
static int welford_finalize(const welford_t* w, int32_t op, double* out) {
  switch (op) {
    case BO_VAR_POP: if (w->count == 0) return 0; *out = w->m2 / (double)w->count; return 1;
    case BO_VAR_SAMP: if (w->count <= 1) return 0; *out = w->m2 / (double)(w->count - 1); return 1;
    case BO_STDDEV_POP: if (w->count == 0) return 0; *out = sqrt(w->m2 / (double)w->count); return 1;
    case BO_STDDEV_SAMP: if (w->count <= 1) return 0; *out = sqrt(w->m2 / (double)(w->count - 1)); return 1;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Scalar aggregates — src/exec/aggregate.rs:154,1596,2032-2262             */
/* ------------------------------------------------------------------------ */

static const i128 I64_MAX_128 = (i128)INT64_MAX;
static const i128 I64_MIN_128 = (i128)INT64_MIN;

/* checked add in i128 range: returns 0 on overflow */
static int checked_add_i128(i128 a, i128 b, i128* out) { return !__builtin_add_overflow(a, b, out); }

bo_scalar bo_scalar_agg(const bo_col* col, int32_t op) {
  bo_scalar r;
  memset(&r, 0, sizeof r);
  int64_t n = col->n;
  if (op == BO_COUNT_STAR) {
    r.dtype = BO_I64;
    r.i64 = n;
    return r;
  }
  if (op == BO_COUNT) {
    int64_t c = 0;
    for (int64_t i = 0; i < n; i++) c += is_valid(col, i);
    r.dtype = BO_I64;
    r.i64 = c;
    return r;
  }
  /* value-NULLs are stripped first; an empty strip yields SQL NULL
   * (tests/aggregate_nulls_e2e.rs:424-470) */
  int64_t nn = 0;
  for (int64_t i = 0; i < n; i++) nn += is_valid(col, i);
  if (op == BO_SUM) {
    if (col->dtype == BO_DATE32) { r.status = BO_ERR_TYPE; return r; }
    r.dtype = col->dtype == BO_I32 ? BO_I64 : col->dtype; /* logical_plan.rs:2191 */
    if (nn == 0) { r.is_null = 1; return r; }
    if (is_int_dtype(col->dtype)) {
      /* aggregate.rs:2120-2140: checked_add in i64, error on overflow */
      int64_t s = 0;
      for (int64_t i = 0; i < n; i++) {
        if (!is_valid(col, i)) continue;
        if (__builtin_add_overflow(s, get_i64(col, i), &s)) { r.status = BO_ERR_TYPE; return r; }
      }
      r.i64 = s;
    } else if (col->dtype == BO_DEC128) {
      /* logical_plan.rs:2203-2214: i128 checked add, precision 38 */
      i128 s = 0;
      for (int64_t i = 0; i < n; i++) {
        if (!is_valid(col, i)) continue;
        if (!checked_add_i128(s, get_i128(col, i), &s)) { r.status = BO_ERR_TYPE; return r; }
      }
      r.i64 = (int64_t)(uint64_t)s;
      r.hi = (int64_t)(s >> 64);
    } else {
      neum_t a = {0, 0, 0, 0};
      for (int64_t i = 0; i < n; i++)
        if (is_valid(col, i)) neum_add(&a, get_f64(col, i));
      r.f64 = neum_result(&a);
      if (col->dtype == BO_F32) r.f64 = (double)(float)r.f64;
    }
    return r;
  }
  if (op == BO_MIN || op == BO_MAX) {
    r.dtype = col->dtype;
    if (nn == 0) { r.is_null = 1; return r; }
    int seen = 0;
    if (is_float_dtype(col->dtype)) {
      double best = 0;
      for (int64_t i = 0; i < n; i++) {
        if (!is_valid(col, i)) continue;
        double v = get_f64(col, i);
        if (!seen) { best = v; seen = 1; continue; }
        int c = total_cmp_f64(v, best);
        if ((op == BO_MIN && c < 0) || (op == BO_MAX && c > 0)) best = v;
      }
      r.f64 = best;
    } else if (col->dtype == BO_DEC128) {
      i128 best = 0;
      for (int64_t i = 0; i < n; i++) {
        if (!is_valid(col, i)) continue;
        i128 v = get_i128(col, i);
        if (!seen || (op == BO_MIN ? v < best : v > best)) best = v;
        seen = 1;
      }
      r.i64 = (int64_t)(uint64_t)best;
      r.hi = (int64_t)(best >> 64);
    } else {
      int64_t best = 0;
      for (int64_t i = 0; i < n; i++) {
        if (!is_valid(col, i)) continue;
        int64_t v = get_i64(col, i);
        if (!seen || (op == BO_MIN ? v < best : v > best)) best = v;
        seen = 1;
      }
      r.i64 = best;
    }
    return r;
  }
  if (op == BO_AVG) {
    r.dtype = BO_F64;
    if (nn == 0) { r.is_null = 1; return r; }
    neum_t a = {0, 0, 0, 0};
    for (int64_t i = 0; i < n; i++)
      if (is_valid(col, i)) neum_add(&a, get_f64(col, i));
    r.f64 = neum_result(&a) / (double)nn;
    return r;
  }
  if (op >= BO_VAR_POP && op <= BO_STDDEV_SAMP) {
    r.dtype = BO_F64;
    welford_t w = {0, 0, 0};
    for (int64_t i = 0; i < n; i++)
      if (is_valid(col, i)) welford_push(&w, get_f64(col, i));
    if (!welford_finalize(&w, op, &r.f64)) r.is_null = 1;
    return r;
  }
  r.status = BO_ERR_PLAN;
  return r;
}

/* ------------------------------------------------------------------------ */
/* Filter — src/exec/compact.rs:100-146                                     */
/* ------------------------------------------------------------------------ */

static int eval_term(const bo_col* c, const bo_term* t, int64_t i) {
  int valid = is_valid(c, i);
  if (t->op == BO_IS_NULL) return !valid;
  if (t->op == BO_IS_NOT_NULL) return valid;
  if (!valid) return 0; /* comparison with NULL is NULL -> row dropped */
  int cmp;
  if (is_float_dtype(c->dtype)) {
    double v = get_f64(c, i), l = t->lit_f64;
    switch (t->op) { /* IEEE: every ordered comparison with NaN is false */
      case BO_LT: return v < l;
      case BO_LE: return v <= l;
      case BO_GT: return v > l;
      case BO_GE: return v >= l;
      case BO_EQ: return v == l;
      case BO_NE: return v != l;
    }
    return 0;
  }
  if (c->dtype == BO_DEC128) {
    i128 v = get_i128(c, i);
    i128 l = (i128)(((unsigned __int128)(uint64_t)t->lit_hi << 64) | (uint64_t)t->lit_i64);
    cmp = v < l ? -1 : (v > l ? 1 : 0);
  } else {
    int64_t v = get_i64(c, i);
    cmp = v < t->lit_i64 ? -1 : (v > t->lit_i64 ? 1 : 0);
  }
  switch (t->op) {
    case BO_LT: return cmp < 0;
    case BO_LE: return cmp <= 0;
    case BO_GT: return cmp > 0;
    case BO_GE: return cmp >= 0;
    case BO_EQ: return cmp == 0;
    case BO_NE: return cmp != 0;
  }
  return 0;
}

int64_t bo_filter_rows(const bo_col* cols, int32_t n_cols, const bo_term* terms, int32_t n_terms,
                       int64_t n_rows, uint32_t* out_rows) {
  (void)n_cols;
  int64_t k = 0;
  for (int64_t i = 0; i < n_rows; i++) {
    int keep = 1;
    for (int32_t t = 0; t < n_terms && keep; t++) keep = eval_term(&cols[terms[t].col], &terms[t], i);
    if (keep) out_rows[k++] = (uint32_t)i;
  }
  return k;
}

/* ------------------------------------------------------------------------ */
/* int64 -> index hash map (open addressing, occupancy flags)              */
/* ------------------------------------------------------------------------ */

typedef struct { int64_t* keys; int64_t* vals; uint8_t* used; uint64_t mask; } hmap_t;

static uint64_t mix64(uint64_t x) {
  x ^= x >> 33; x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}
static void hmap_init(hmap_t* m, int64_t n) {
  uint64_t cap = 16;
  while (cap < (uint64_t)n * 2) cap <<= 1;
  m->keys = (int64_t*)malloc(cap * 8);
  m->vals = (int64_t*)malloc(cap * 8);
  m->used = (uint8_t*)calloc(cap, 1);
  m->mask = cap - 1;
}
static void hmap_free(hmap_t* m) { free(m->keys); free(m->vals); free(m->used); }
/* doubles the table and re-inserts every key (the groups' map grows with
 * the groups seen, not with the rows) */
static void hmap_grow(hmap_t* m) {
  hmap_t g;
  uint64_t cap = (m->mask + 1) * 2;
  g.keys = (int64_t*)malloc(cap * 8);
  g.vals = (int64_t*)malloc(cap * 8);
  g.used = (uint8_t*)calloc(cap, 1);
  g.mask = cap - 1;
  for (uint64_t s = 0; s <= m->mask; s++) {
    if (!m->used[s]) continue;
    uint64_t t = mix64((uint64_t)m->keys[s]) & g.mask;
    while (g.used[t]) t = (t + 1) & g.mask;
    g.used[t] = 1;
    g.keys[t] = m->keys[s];
    g.vals[t] = m->vals[s];
  }
  hmap_free(m);
  *m = g;
}
/* returns the slot for key; *found set when already present */
static uint64_t hmap_slot(hmap_t* m, int64_t key, int* found) {
  uint64_t s = mix64((uint64_t)key) & m->mask;
  while (m->used[s]) {
    if (m->keys[s] == key) { *found = 1; return s; }
    s = (s + 1) & m->mask;
  }
  *found = 0;
  return s;
}

/* ------------------------------------------------------------------------ */
/* GROUP BY — src/exec/groupby.rs:295-833, groupby_common.rs:204-420        */
/* ------------------------------------------------------------------------ */

#define NULL_GROUP_KEY INT64_MAX /* groupby.rs:144 */

/* groupby_common.rs:302,325 canonicalisation then bit pattern */
static uint64_t key_bits(const bo_col* c, int64_t i) {
  switch (c->dtype) {
    case BO_I32: case BO_DATE32: /* Date32 keys (rejected by the reference, groupby_common.rs:117)
                                    are packed like Int32 */
      return (uint64_t)(uint32_t)((const int32_t*)c->data)[i];
    case BO_I64: return (uint64_t)((const int64_t*)c->data)[i];
    case BO_DEC128: /* a Decimal128 key that fits i64 (checked by the caller) packs as its low word */
      return (uint64_t)((const int64_t*)c->data)[2 * i];
    case BO_F32: {
      float f = ((const float*)c->data)[i];
      uint32_t b;
      if (isnan(f)) b = 0x7FC00000u; else if (f == 0.0f) b = 0; else memcpy(&b, &f, 4);
      return b;
    }
    case BO_F64: {
      double d = ((const double*)c->data)[i];
      uint64_t b;
      if (isnan(d)) b = 0x7FF8000000000000ull; else if (d == 0.0) b = 0; else memcpy(&b, &d, 8);
      return b;
    }
  }
  return 0;
}

typedef struct {
  i128 isum;       /* integer / decimal SUM (running checked) */
  neum_t fsum;     /* float SUM / AVG numerator */
  int64_t cnt;     /* non-null values seen */
  int64_t imm;     /* integer MIN/MAX */
  i128 dmm;        /* decimal MIN/MAX */
  double fmm;      /* float MIN/MAX */
  welford_t w;
} gacc_t;

static int cmp_i64(const void* a, const void* b) {
  int64_t x = ((const int64_t*)a)[0], y = ((const int64_t*)b)[0];
  return x < y ? -1 : (x > y ? 1 : 0);
}

int32_t bo_groupby(const bo_col* keys, int32_t n_keys, const bo_col* inputs, const int32_t* agg_ops,
                   const int32_t* agg_cols, int32_t n_aggs, int64_t* n_groups, int64_t* out_keys,
                   uint8_t* out_key_null, int64_t** out_i64, double** out_f64, int64_t** out_hi,
                   uint8_t** out_null, int64_t out_cap) {
  if (n_keys < 1 || n_keys > 2) return BO_ERR_PLAN;
  int64_t n = keys[0].n;
  /* Decimal128 keys: one column, every valid value within i64 (the GPU
   * operator declines wider ones; the reference rejects Decimal keys,
   * groupby_common.rs:117) */
  if (keys[0].dtype == BO_DEC128) {
    if (n_keys != 1) return BO_ERR_PLAN;
    for (int64_t i = 0; i < n; i++)
      if (is_valid(&keys[0], i)) {
        const int64_t* w = (const int64_t*)keys[0].data + 2 * i;
        if (w[1] != (w[0] >> 63)) return BO_ERR_PLAN;
      }
  }
  if (n_keys == 2) {
    int w0 = (keys[0].dtype == BO_I32 || keys[0].dtype == BO_F32);
    int w1 = (keys[1].dtype == BO_I32 || keys[1].dtype == BO_F32);
    if (!w0 || !w1) return BO_ERR_PLAN; /* > 64 bits: wide path, not modelled here */
  }
  /* pack (groupby_common.rs:352): single column in the low bits; two columns
   * with the first in the high 32 bits */
  int64_t* packed = (int64_t*)malloc((size_t)(n > 0 ? n : 1) * 8);
  uint8_t* keep = (uint8_t*)malloc((size_t)(n > 0 ? n : 1));
  int any_null = 0;
  for (int64_t i = 0; i < n; i++) {
    int v = 1;
    for (int k = 0; k < n_keys; k++) v &= is_valid(&keys[k], i);
    keep[i] = (uint8_t)v;
    any_null |= !v;
    if (n_keys == 1) packed[i] = (int64_t)key_bits(&keys[0], i);
    else packed[i] = (int64_t)((key_bits(&keys[0], i) << 32) | key_bits(&keys[1], i));
  }
  /* NULL group (groupby.rs:528-606): single column only; a genuine key equal
   * to the sentinel routes to the valid-flag executor, which drops NULL rows */
  int synth_null = (n_keys == 1 && any_null);
  if (synth_null) {
    for (int64_t i = 0; i < n; i++)
      if (keep[i] && packed[i] == NULL_GROUP_KEY) { synth_null = 0; break; }
  }
  /* accumulators, keys and the map are sized by the groups seen so far
   * (doubling), so memory follows the group count, not the row count */
  const size_t na = (size_t)(n_aggs > 0 ? n_aggs : 1);
  int64_t gcap = 1024;
  hmap_t m;
  hmap_init(&m, gcap);
  gacc_t* acc = (gacc_t*)calloc((size_t)gcap * na, sizeof(gacc_t));
  int64_t* gkey = (int64_t*)malloc((size_t)gcap * 8);
  int64_t ng = 0;
  int32_t status = BO_OK;
  for (int64_t i = 0; i < n && status == BO_OK; i++) {
    int64_t key;
    if (keep[i]) key = packed[i];
    else if (synth_null) key = NULL_GROUP_KEY;
    else continue;
    int found;
    uint64_t s = hmap_slot(&m, key, &found);
    int64_t g;
    if (found) g = m.vals[s];
    else {
      if (ng == gcap) {
        acc = (gacc_t*)realloc(acc, (size_t)gcap * 2 * na * sizeof(gacc_t));
        memset(acc + (size_t)gcap * na, 0, (size_t)gcap * na * sizeof(gacc_t));
        gkey = (int64_t*)realloc(gkey, (size_t)gcap * 2 * 8);
        gcap *= 2;
      }
      if ((uint64_t)(ng + 1) * 2 > m.mask + 1) {
        hmap_grow(&m);
        s = hmap_slot(&m, key, &found);
      }
      m.used[s] = 1; m.keys[s] = key; m.vals[s] = ng; gkey[ng] = key; g = ng++;
    }
    for (int32_t a = 0; a < n_aggs; a++) {
      gacc_t* A = &acc[g * n_aggs + a];
      int32_t op = agg_ops[a];
      if (op == BO_COUNT_STAR) { A->cnt++; continue; }
      const bo_col* c = &inputs[agg_cols[a]];
      if (!is_valid(c, i)) continue;
      int first = (A->cnt == 0);
      A->cnt++;
      switch (op) {
        case BO_SUM:
          if (is_int_dtype(c->dtype)) {
            /* groupby.rs:1775 checked_group_sum: running i64 checked add */
            int64_t s64 = (int64_t)A->isum;
            if (__builtin_add_overflow(s64, get_i64(c, i), &s64)) status = BO_ERR_TYPE;
            A->isum = s64;
          } else if (c->dtype == BO_DEC128) {
            if (!checked_add_i128(A->isum, get_i128(c, i), &A->isum)) status = BO_ERR_TYPE;
          } else {
            neum_add(&A->fsum, get_f64(c, i));
          }
          break;
        case BO_AVG: neum_add(&A->fsum, get_f64(c, i)); break;
        case BO_MIN: case BO_MAX:
          if (is_float_dtype(c->dtype)) {
            /* float_atomics.rs:27-62 NaN-largest total order */
            double v = get_f64(c, i);
            int cmp = first ? 0 : total_cmp_f64(v, A->fmm);
            if (first || (op == BO_MIN && cmp < 0) || (op == BO_MAX && cmp > 0)) A->fmm = v;
          } else if (c->dtype == BO_DEC128) {
            i128 v = get_i128(c, i);
            if (first || (op == BO_MIN ? v < A->dmm : v > A->dmm)) A->dmm = v;
          } else {
            int64_t v = get_i64(c, i);
            if (first || (op == BO_MIN ? v < A->imm : v > A->imm)) A->imm = v;
          }
          break;
        case BO_COUNT: break;
        default: welford_push(&A->w, get_f64(c, i)); break;
      }
    }
  }
  if (status != BO_OK) goto done;
  if (ng > out_cap) {
    *n_groups = ng;
    status = BO_ERR_CAP;
    goto done;
  }
  /* sort groups by packed key (groupby.rs:800) */
  int64_t* order = (int64_t*)malloc((size_t)(ng > 0 ? ng : 1) * 16);
  for (int64_t g = 0; g < ng; g++) { order[2 * g] = gkey[g]; order[2 * g + 1] = g; }
  qsort(order, (size_t)ng, 16, cmp_i64);
  for (int64_t r = 0; r < ng; r++) {
    int64_t g = order[2 * r + 1];
    out_keys[r] = gkey[g];
    out_key_null[r] = (uint8_t)(synth_null && gkey[g] == NULL_GROUP_KEY);
    for (int32_t a = 0; a < n_aggs; a++) {
      gacc_t* A = &acc[g * n_aggs + a];
      int32_t op = agg_ops[a];
      const bo_col* c = op == BO_COUNT_STAR ? NULL : &inputs[agg_cols[a]];
      out_null[a][r] = 0;
      out_i64[a][r] = 0;
      out_f64[a][r] = 0;
      out_hi[a][r] = 0;
      if (op == BO_COUNT || op == BO_COUNT_STAR) { out_i64[a][r] = A->cnt; continue; }
      if (A->cnt == 0) { out_null[a][r] = 1; continue; }
      switch (op) {
        case BO_SUM:
          if (is_int_dtype(c->dtype) || c->dtype == BO_DEC128) {
            out_i64[a][r] = (int64_t)(uint64_t)A->isum;
            out_hi[a][r] = (int64_t)(A->isum >> 64);
          } else {
            out_f64[a][r] = neum_result(&A->fsum);
            if (c->dtype == BO_F32) out_f64[a][r] = (double)(float)out_f64[a][r];
          }
          break;
        case BO_AVG: out_f64[a][r] = neum_result(&A->fsum) / (double)A->cnt; break;
        case BO_MIN: case BO_MAX:
          if (is_float_dtype(c->dtype)) out_f64[a][r] = A->fmm;
          else if (c->dtype == BO_DEC128) { out_i64[a][r] = (int64_t)(uint64_t)A->dmm; out_hi[a][r] = (int64_t)(A->dmm >> 64); }
          else out_i64[a][r] = A->imm;
          break;
        default:
          if (!welford_finalize(&A->w, op, &out_f64[a][r])) out_null[a][r] = 1;
          break;
      }
    }
  }
  free(order);
  *n_groups = ng;
done:
  free(packed); free(keep); free(acc); free(gkey);
  hmap_free(&m);
  (void)I64_MAX_128; (void)I64_MIN_128;
  return status;
}

/* ------------------------------------------------------------------------ */
/* INNER hash join — src/exec/join.rs:169-270, 844-860                      */
/* ------------------------------------------------------------------------ */

int64_t bo_inner_join(const bo_col* left_key, const bo_col* right_key, uint32_t* out_left,
                      uint32_t* out_right, int64_t cap, int64_t* needed) {
  int64_t nl = left_key->n, nr = right_key->n;
  *needed = 0;
  if (nl == 0 || nr == 0) return 0;
  int build_is_left = nl <= nr; /* join.rs:192 */
  const bo_col* b = build_is_left ? left_key : right_key;
  const bo_col* p = build_is_left ? right_key : left_key;
  /* build: key -> head row; rows chained in build order (BuildSlot push) */
  hmap_t m;
  hmap_init(&m, b->n);
  int64_t* next = (int64_t*)malloc((size_t)b->n * 8);
  int64_t* tail = (int64_t*)malloc((size_t)b->n * 8);
  for (int64_t i = 0; i < b->n; i++) {
    next[i] = -1;
    if (!is_valid(b, i)) continue; /* NULL build keys skipped (join.rs:851) */
    int found;
    uint64_t s = hmap_slot(&m, get_i64(b, i), &found);
    if (!found) { m.used[s] = 1; m.keys[s] = get_i64(b, i); m.vals[s] = i; tail[i] = i; }
    else { int64_t h = m.vals[s]; next[tail[h]] = i; tail[h] = i; }
  }
  int64_t k = 0;
  for (int64_t j = 0; j < p->n; j++) {
    if (!is_valid(p, j)) continue;
    int found;
    uint64_t s = hmap_slot(&m, get_i64(p, j), &found);
    if (!found) continue;
    for (int64_t r = m.vals[s]; r >= 0; r = next[r]) {
      if (k < cap) {
        if (build_is_left) { out_left[k] = (uint32_t)r; out_right[k] = (uint32_t)j; }
        else { out_left[k] = (uint32_t)j; out_right[k] = (uint32_t)r; }
      }
      k++;
    }
  }
  free(next); free(tail);
  hmap_free(&m);
  *needed = k;
  return k <= cap ? k : -1;
}

/* one join key as an i64 that is equal exactly when the keys are (SQL
 * equality of extract_key, src/exec/join.rs:1045; floats by canonical bits,
 * gpu_join.rs:682-705; two 32-bit columns packed first-high, :788).
 * Returns 0 for a NULL key. */
static int join_key(const bo_col* cols, int32_t n_keys, int64_t i, int64_t* key) {
  for (int c = 0; c < n_keys; c++)
    if (!is_valid(&cols[c], i)) return 0;
  if (n_keys > 2 || (n_keys == 2 && (cols[0].dtype == BO_I64 || cols[1].dtype == BO_I64))) {
    /* the wrapping sum of the tuple: collision-prone on purpose ((1, 2) and
     * (2, 1) share it), so join_tuple_eq, which separates such tuples, is
     * exercised by every test that has multi-column keys */
    uint64_t h = 0;
    for (int c = 0; c < n_keys; c++) h += (uint64_t)get_i64(&cols[c], i);
    *key = (int64_t)h;
    return 1;
  }
  if (n_keys == 2) {
    uint32_t hi = (uint32_t)((const int32_t*)cols[0].data)[i];
    uint32_t lo = (uint32_t)((const int32_t*)cols[1].data)[i];
    *key = (int64_t)(((uint64_t)hi << 32) | lo);
    return 1;
  }
  const bo_col* c = &cols[0];
  switch (c->dtype) {
    case BO_BOOL: *key = ((const uint8_t*)c->data)[i] != 0; return 1;
    case BO_F32: {
      float f = ((const float*)c->data)[i];
      uint32_t b;
      if (isnan(f)) b = 0x7FC00000u;      /* f32::NAN.to_bits() */
      else if (f == 0.0f) b = 0;          /* -0.0 == +0.0 */
      else memcpy(&b, &f, 4);
      *key = (int64_t)(uint64_t)b;
      return 1;
    }
    case BO_F64: {
      double d = ((const double*)c->data)[i];
      uint64_t b;
      if (isnan(d)) b = 0x7FF8000000000000ull;
      else if (d == 0.0) b = 0;
      else memcpy(&b, &d, 8);
      *key = (int64_t)b;
      return 1;
    }
    default: *key = get_i64(c, i); return 1;
  }
}

/* exact SQL equality of two non-NULL key tuples (JoinKey::Many, join.rs:1062) */
static int join_tuple_eq(const bo_col* a, int64_t i, const bo_col* b, int64_t j, int32_t n_keys) {
  if (n_keys == 1) return 1; /* one column: the key itself is exact */
  for (int c = 0; c < n_keys; c++)
    if (get_i64(&a[c], i) != get_i64(&b[c], j)) return 0;
  return 1;
}

int64_t bo_join(const bo_col* left_keys, const bo_col* right_keys, int32_t n_keys, int32_t join_type,
                uint32_t* out_left, uint32_t* out_right, int64_t cap, int64_t* needed) {
  const uint32_t NUL = 0xFFFFFFFFu;
  int64_t nl = left_keys[0].n, nr = right_keys[0].n;
  *needed = 0;
  /* build side (join.rs:192 for INNER, :305 for OUTER) */
  int build_is_left = join_type == 0 ? nl <= nr : join_type == 2;
  int preserve_probe = join_type != 0, preserve_build = join_type == 3;
  if (join_type == 0 && (nl == 0 || nr == 0)) return 0;
  const bo_col* b = build_is_left ? left_keys : right_keys;
  const bo_col* p = build_is_left ? right_keys : left_keys;
  int64_t nb = b[0].n, np = p[0].n;
  hmap_t m;
  hmap_init(&m, nb > 0 ? nb : 1);
  int64_t* next = (int64_t*)malloc((size_t)(nb > 0 ? nb : 1) * 8);
  int64_t* tail = (int64_t*)malloc((size_t)(nb > 0 ? nb : 1) * 8);
  uint8_t* matched = (uint8_t*)calloc((size_t)(nb > 0 ? nb : 1), 1);
  for (int64_t i = 0; i < nb; i++) { /* build_hash_map (join.rs:844): NULL keys skipped */
    next[i] = -1;
    int64_t k;
    if (!join_key(b, n_keys, i, &k)) continue;
    int found;
    uint64_t s = hmap_slot(&m, k, &found);
    if (!found) { m.used[s] = 1; m.keys[s] = k; m.vals[s] = i; tail[i] = i; }
    else { int64_t h = m.vals[s]; next[tail[h]] = i; tail[h] = i; }
  }
  int64_t n = 0;
#define EMIT(BROW, PROW)                                                               \
  do {                                                                                 \
    if (n < cap) {                                                                     \
      if (build_is_left) { out_left[n] = (BROW); out_right[n] = (PROW); }              \
      else { out_left[n] = (PROW); out_right[n] = (BROW); }                            \
    }                                                                                  \
    n++;                                                                               \
  } while (0)
  for (int64_t j = 0; j < np; j++) { /* join.rs:396-437 */
    int64_t k, head = -1;
    if (join_key(p, n_keys, j, &k)) {
      int found;
      uint64_t s = hmap_slot(&m, k, &found);
      if (found) head = m.vals[s];
    }
    int hit = 0;
    for (int64_t r = head; r >= 0; r = next[r]) {
      if (!join_tuple_eq(b, r, p, j, n_keys)) continue;
      EMIT((uint32_t)r, (uint32_t)j);
      matched[r] = 1;
      hit = 1;
    }
    if (!hit && preserve_probe) EMIT(NUL, (uint32_t)j);
  }
  if (preserve_build) /* join.rs:440-449 */
    for (int64_t r = 0; r < nb; r++)
      if (!matched[r]) EMIT((uint32_t)r, NUL);
#undef EMIT
  free(next); free(tail); free(matched);
  hmap_free(&m);
  *needed = n;
  return n <= cap ? n : -1;
}

/* ------------------------------------------------------------------------ */
/* ORDER BY — src/exec/sort.rs:161-230, gpu_sort.rs:1070-1180               */
/* ------------------------------------------------------------------------ */

/* order-preserving u64 image of one key value (ascending) */
static uint64_t sort_image(const bo_col* c, int64_t i) {
  switch (c->dtype) {
    case BO_I32: case BO_DATE32:
      return (uint64_t)((uint32_t)((const int32_t*)c->data)[i] ^ 0x80000000u); /* gpu_sort.rs:1075 */
    case BO_I64: return (uint64_t)((const int64_t*)c->data)[i] ^ 0x8000000000000000ull; /* :1091 */
    case BO_F32: {
      float f = ((const float*)c->data)[i];
      uint32_t b;
      if (isnan(f)) b = 0x7FC00000u; else if (f == 0.0f) b = 0; else memcpy(&b, &f, 4);
      uint32_t mask = (uint32_t)((int32_t)b >> 31) | 0x80000000u; /* :1150 */
      return (uint64_t)(b ^ mask);
    }
    case BO_F64: {
      double d = ((const double*)c->data)[i];
      uint64_t b;
      if (isnan(d)) b = 0x7FF8000000000000ull; else if (d == 0.0) b = 0; else memcpy(&b, &d, 8);
      uint64_t mask = (uint64_t)((int64_t)b >> 63) | 0x8000000000000000ull; /* :1169 */
      return b ^ mask;
    }
  }
  return 0;
}

typedef struct { const bo_col* keys; int32_t n_keys; const int32_t* desc; const int32_t* nf; } sort_ctx_t;

static int sort_cmp_rows(const sort_ctx_t* S, uint32_t a, uint32_t b) {
  for (int32_t k = 0; k < S->n_keys; k++) {
    const bo_col* c = &S->keys[k];
    int va = is_valid(c, a), vb = is_valid(c, b);
    if (va != vb) {
      /* the NULL row goes first iff nulls_first */
      int a_first = (!va) ? S->nf[k] : !S->nf[k];
      return a_first ? -1 : 1;
    }
    if (!va) continue;
    int lt;
    if (c->dtype == BO_DEC128) {
      /* Decimal128 orders as its signed 128-bit value (arrow-ord lexsort,
         the comparator of src/exec/sort.rs:222) */
      const uint64_t* p = (const uint64_t*)c->data;
      i128 xa = (i128)(((unsigned __int128)p[2 * a + 1] << 64) | p[2 * a]);
      i128 xb = (i128)(((unsigned __int128)p[2 * b + 1] << 64) | p[2 * b]);
      if (xa == xb) continue;
      lt = xa < xb;
    } else {
      uint64_t ia = sort_image(c, a), ib = sort_image(c, b);
      if (ia == ib) continue;
      lt = ia < ib;
    }
    if (S->desc[k]) lt = !lt;
    return lt ? -1 : 1;
  }
  return 0;
}

static void merge_sort(const sort_ctx_t* S, uint32_t* a, uint32_t* tmp, int64_t n) {
  /* bottom-up stable merge sort */
  for (int64_t w = 1; w < n; w *= 2) {
    for (int64_t lo = 0; lo < n; lo += 2 * w) {
      int64_t mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
      int64_t i = lo, j = mid, k = lo;
      while (i < mid && j < hi) tmp[k++] = sort_cmp_rows(S, a[j], a[i]) < 0 ? a[j++] : a[i++];
      while (i < mid) tmp[k++] = a[i++];
      while (j < hi) tmp[k++] = a[j++];
    }
    memcpy(a, tmp, (size_t)n * 4);
  }
}

void bo_sort_indices(const bo_col* keys, int32_t n_keys, const int32_t* descending,
                     const int32_t* nulls_first, int64_t n_rows, uint32_t* out_idx) {
  sort_ctx_t S = {keys, n_keys, descending, nulls_first};
  for (int64_t i = 0; i < n_rows; i++) out_idx[i] = (uint32_t)i;
  if (n_rows < 2) return;
  uint32_t* tmp = (uint32_t*)malloc((size_t)n_rows * 4);
  merge_sort(&S, out_idx, tmp, n_rows);
  free(tmp);
}

/* ------------------------------------------------------------------------ */
/* Filter + INNER join + scalar aggregates (the reference composition)      */
/* ------------------------------------------------------------------------ */

int32_t bo_filter_join_agg(const int64_t* build_key, const int32_t* build_payload, int64_t n_build,
                           const int64_t* probe_key, const double* probe_amount, int64_t n_probe,
                           double amount_gt, int64_t* sum_payload, double* sum_amount_naive,
                           double* sum_amount_neumaier, int64_t* count) {
  /* the filtered probe side is normally the larger one; the build side is
   * the smaller (join.rs:192).  With unique build keys each probe row
   * matches at most once, so the join output order is probe order. */
  hmap_t m;
  hmap_init(&m, n_build);
  int dup = 0;
  for (int64_t i = 0; i < n_build; i++) {
    int found;
    uint64_t s = hmap_slot(&m, build_key[i], &found);
    if (found) { dup = 1; continue; }
    m.used[s] = 1; m.keys[s] = build_key[i]; m.vals[s] = i;
  }
  if (dup) { hmap_free(&m); return BO_ERR_PLAN; }
  int64_t sp = 0, c = 0;
  neum_t a = {0, 0, 0, 0};
  int32_t status = BO_OK;
  for (int64_t j = 0; j < n_probe; j++) {
    if (!(probe_amount[j] > amount_gt)) continue;
    int found;
    uint64_t s = hmap_slot(&m, probe_key[j], &found);
    if (!found) continue;
    if (__builtin_add_overflow(sp, (int64_t)build_payload[m.vals[s]], &sp)) status = BO_ERR_TYPE;
    neum_add(&a, probe_amount[j]);
    c++;
  }
  hmap_free(&m);
  *sum_payload = sp;
  *sum_amount_naive = a.naive;
  *sum_amount_neumaier = neum_result(&a);
  *count = c;
  return status;
}

/* ------------------------------------------------------------------------ */
/* Per-row expression programs — src/plan/physical_plan.rs:81, ptx_gen.rs    */
/* ------------------------------------------------------------------------ */

enum { X_LOAD = 1, X_CONST, X_CAST, X_BINARY, X_SELECT, X_NOT, X_IS_NULL, X_STORE, X_EXTRACT, X_DATE_TRUNC };
enum { DF_YEAR = 1, DF_MONTH, DF_DAY, DF_HOUR, DF_MINUTE, DF_SECOND };
enum { B_ADD = 1, B_SUB, B_MUL, B_DIV, B_MOD, B_EQ, B_NE, B_LT, B_LE, B_GT, B_GE, B_AND, B_OR, B_BAND, B_BOR };

typedef struct {
  int32_t dt;
  int64_t i;    /* Bool / Int32 / Int64 / Date32 / Timestamp ticks */
  double f;     /* Float64; Float32 values are held rounded to float */
  __int128 d;   /* Decimal128 */
  int32_t u;    /* Timestamp unit: digits of a second (0, 3, 6, 9) */
} xval;

/* floor_div and civil_from_days of src/jit/date_scalar.rs:29-60 (the
 * proleptic Gregorian calendar on days since 1970-01-01), and the inverse */
static int64_t x_floor_div(int64_t a, int64_t b) {
  int64_t q = a / b, r = a % b;
  return (r != 0 && ((r < 0) != (b < 0))) ? q - 1 : q;
}
static void x_civil(int64_t days, int64_t* y, int64_t* m, int64_t* d) {
  int64_t z = days + 719468, era = x_floor_div(z, 146097), doe = z - era * 146097;
  int64_t yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365;
  int64_t doy = doe - (365 * yoe + yoe / 4 - yoe / 100), mp = (5 * doy + 2) / 153;
  *d = doy - (153 * mp + 2) / 5 + 1;
  *m = mp < 10 ? mp + 3 : mp - 9;
  *y = yoe + era * 400 + (*m <= 2);
}
static int64_t x_days(int64_t y, int64_t m, int64_t d) {
  y -= m <= 2;
  int64_t era = x_floor_div(y, 400), yoe = y - era * 400;
  int64_t doy = (153 * (m > 2 ? m - 3 : m + 9) + 2) / 5 + d - 1;
  return era * 146097 + yoe * 365 + yoe / 4 - yoe / 100 + doy - 719468;
}
static int64_t x_pow10(int s) { int64_t r = 1; while (s-- > 0) r *= 10; return r; }
static __int128 x_pow10x(int s) { __int128 r = 1; while (s-- > 0) r *= 10; return r; }
static double x_pow10f(int s) { double r = 1.0; while (s-- > 0) r *= 10.0; return r; }

/* date_scalar.rs:459 EXTRACT and :539 DATE_TRUNC on days (ts == 0) or ticks */
static int64_t x_extract(int64_t x, int ts, int unit, int field) {
  int64_t days = x, sod = 0;
  if (ts) {
    int64_t tps = x_pow10(unit), tpd = 86400 * tps;
    days = x_floor_div(x, tpd);
    const int64_t r = x % tpd; /* the floor remainder, without forming days * tpd */
    sod = (r < 0 ? r + tpd : r) / tps;
  }
  if (field == DF_HOUR) return sod / 3600;
  if (field == DF_MINUTE) return (sod / 60) % 60;
  if (field == DF_SECOND) return sod % 60;
  int64_t y, m, d;
  x_civil(days, &y, &m, &d);
  return field == DF_YEAR ? y : field == DF_MONTH ? m : d;
}
static int64_t x_trunc(int64_t x, int ts, int unit, int field) {
  int64_t tps = ts ? x_pow10(unit) : 1, tpd = ts ? 86400 * tps : 1;
  if (field >= DF_HOUR) {
    int64_t q = field == DF_HOUR ? 3600 * tps : field == DF_MINUTE ? 60 * tps : tps;
    return (int64_t)((uint64_t)x_floor_div(x, q) * (uint64_t)q); /* mul.lo.s64: wraps */
  }
  int64_t days = x_floor_div(x, tpd), y, m, d;
  x_civil(days, &y, &m, &d);
  const int64_t start = field == DF_YEAR ? x_days(y, 1, 1) : field == DF_MONTH ? x_days(y, m, 1) : days;
  return (int64_t)((uint64_t)start * (uint64_t)tpd);
}

/* F64ToI128 / I128ToF64 of src/plan/physical_plan.rs:496-523: round half away
 * from zero as trunc(|v| + 0.5), NaN -> 0, saturation at 2^127; the i128 as
 * (double)hi * 2^64 + (double)lo */
static __int128 x_f2d(double v) {
  if (isnan(v)) return 0;
  double a = fabs(v);
  if (a >= ldexp(1.0, 127)) return v < 0 ? (__int128)((unsigned __int128)1 << 127) : (__int128)(((unsigned __int128)1 << 127) - 1);
  double m = trunc(a + 0.5); /* below 2^127, where adding 0.5 cannot carry past it */
  __int128 r = (__int128)m;  /* an integer below 2^127: exact */
  return v < 0 ? -r : r;
}
static double x_d2f(__int128 x) {
  volatile double hi = (double)(int64_t)(x >> 64) * 18446744073709551616.0;
  volatile double lo = (double)(uint64_t)x;
  return hi + lo;
}

#define XREGS 256

static int x_int(int dt) { return dt == BO_I32 || dt == BO_I64 || dt == BO_DATE32; }
static int x_float(int dt) { return dt == BO_F32 || dt == BO_F64; }

/* two's-complement wrap of an integer result to the width of dt */
static int64_t x_wrap(int dt, uint64_t v) {
  if (dt == BO_I64) return (int64_t)v;
  return (int64_t)(int32_t)(uint32_t)v;
}

/* cvt.rzi.s32 / .s64: truncate, saturate, NaN -> 0 */
static int64_t x_f2i(double x, int to) {
  double lo = to == BO_I32 ? -2147483648.0 : -9223372036854775808.0;
  double hi = to == BO_I32 ? 2147483647.0 : 9223372036854775807.0;
  if (isnan(x)) return 0;
  if (x >= hi) return to == BO_I32 ? 2147483647LL : INT64_MAX;
  if (x <= lo) return to == BO_I32 ? -2147483648LL : INT64_MIN;
  return (int64_t)x;
}

/* sf / st: the source / target scales when they are Decimal128 (physical_plan.rs:2046-2060) */
static int x_cast(xval* r, const xval* a, int to, int sf, int st) {
  int from = a->dt;
  xval v = {to, 0, 0.0, 0, 0};
  if (from == BO_DEC128 || to == BO_DEC128) {
    if (sf < 0 || sf > 38 || st < 0 || st > 38) return BO_ERR_TYPE;
    if (from == BO_DEC128 && to == BO_DEC128) {
      if (st >= sf) v.d = (__int128)((unsigned __int128)a->d * (unsigned __int128)x_pow10x(st - sf));
      else v.d = a->d / x_pow10x(sf - st); /* truncates toward zero */
    } else if (to == BO_DEC128 && (from == BO_I32 || from == BO_I64 || from == BO_DATE32)) {
      v.d = (__int128)((unsigned __int128)(__int128)a->i * (unsigned __int128)x_pow10x(st));
    } else if (from == BO_DEC128 && (to == BO_I32 || to == BO_I64)) {
      __int128 q = sf > 0 ? a->d / x_pow10x(sf) : a->d;
      v.i = x_wrap(to, (uint64_t)(unsigned __int128)q); /* the low 64 / 32 bits */
    } else if (to == BO_DEC128 && x_float(from)) {
      if (st > 22) return BO_ERR_TYPE;
      v.d = x_f2d(a->f * x_pow10f(st));
    } else if (from == BO_DEC128 && x_float(to)) {
      if (sf > 22) return BO_ERR_TYPE;
      v.f = x_d2f(a->d) / x_pow10f(sf);
      if (to == BO_F32) v.f = (double)(float)v.f;
    } else {
      return BO_ERR_TYPE;
    }
    *r = v;
    return BO_OK;
  }
  if (from == to) { *r = *a; return BO_OK; }
  if (from == BO_DATE32 || to == BO_DATE32 || from == BO_TIMESTAMP || to == BO_TIMESTAMP) return BO_ERR_TYPE;
  if (to == BO_BOOL) {
    if (x_float(from)) v.i = (a->f < 0.0 || a->f > 0.0); /* ordered != 0: NaN -> 0 */
    else v.i = a->i != 0;
  } else if (to == BO_I32 || to == BO_I64) {
    if (x_float(from)) v.i = x_f2i(a->f, to);
    else v.i = x_wrap(to, (uint64_t)a->i);
  } else if (to == BO_F32) {
    v.f = x_float(from) ? (double)(float)a->f : (double)(float)a->i;
  } else {
    v.f = x_float(from) ? a->f : (double)a->i;
  }
  *r = v;
  return BO_OK;
}

static int x_binary(xval* r, const xval* a, const xval* b, int op, int dt) {
  xval v = {dt, 0, 0.0, 0, 0};
  if (a->dt != dt || b->dt != dt) return BO_ERR_TYPE;
  if (dt == BO_TIMESTAMP && (!(op >= B_EQ && op <= B_GE) || a->u != b->u)) return BO_ERR_TYPE; /* compare only */
  if (op >= B_EQ && op <= B_GE) {
    int c;
    v.dt = BO_BOOL;
    if (x_float(dt)) {
      double x = a->f, y = b->f;
      switch (op) {
        case B_EQ: c = x == y; break;
        case B_NE: c = x < y || x > y; break; /* ordered: NaN -> false */
        case B_LT: c = x < y; break;
        case B_LE: c = x <= y; break;
        case B_GT: c = x > y; break;
        default: c = x >= y; break;
      }
    } else if (dt == BO_DEC128) {
      __int128 x = a->d, y = b->d;
      c = op == B_EQ ? x == y : op == B_NE ? x != y : op == B_LT ? x < y : op == B_LE ? x <= y : op == B_GT ? x > y : x >= y;
    } else {
      int64_t x = a->i, y = b->i;
      c = op == B_EQ ? x == y : op == B_NE ? x != y : op == B_LT ? x < y : op == B_LE ? x <= y : op == B_GT ? x > y : x >= y;
    }
    v.i = c;
    *r = v;
    return BO_OK;
  }
  switch (op) {
    case B_ADD: case B_SUB: case B_MUL:
      if (dt == BO_DATE32) {
        if (op != B_SUB) return BO_ERR_TYPE;
        v.dt = BO_I32;
        v.i = x_wrap(BO_I32, (uint64_t)a->i - (uint64_t)b->i);
      } else if (x_float(dt)) {
        double x = a->f, y = b->f, z = op == B_ADD ? x + y : op == B_SUB ? x - y : x * y;
        if (dt == BO_F32) { float xf = (float)x, yf = (float)y; z = op == B_ADD ? xf + yf : op == B_SUB ? xf - yf : xf * yf; z = (float)z; }
        v.f = z;
      } else if (dt == BO_DEC128) {
        unsigned __int128 x = (unsigned __int128)a->d, y = (unsigned __int128)b->d;
        v.d = (__int128)(op == B_ADD ? x + y : op == B_SUB ? x - y : x * y);
      } else if (x_int(dt)) {
        uint64_t x = (uint64_t)a->i, y = (uint64_t)b->i;
        v.i = x_wrap(dt, op == B_ADD ? x + y : op == B_SUB ? x - y : x * y);
      } else return BO_ERR_TYPE;
      break;
    case B_DIV: case B_MOD:
      if (x_float(dt) && op == B_DIV) {
        v.f = dt == BO_F32 ? (double)((float)a->f / (float)b->f) : a->f / b->f;
      } else if (dt == BO_I32 || dt == BO_I64) {
        int64_t x = a->i, y = b->i, mn = dt == BO_I32 ? -2147483648LL : INT64_MIN;
        if (y == 0) v.i = 0;                            /* x / 0 and x % 0 -> 0 */
        else if (x == mn && y == -1) v.i = op == B_DIV ? mn : 0; /* overflow corner */
        else v.i = op == B_DIV ? x / y : x % y;
      } else return BO_ERR_TYPE;
      break;
    case B_AND: case B_OR:
      if (dt != BO_BOOL) return BO_ERR_TYPE;
      v.i = op == B_AND ? (a->i & b->i) : (a->i | b->i);
      break;
    case B_BAND: case B_BOR:
      if (dt != BO_I32 && dt != BO_I64) return BO_ERR_TYPE;
      v.i = op == B_BAND ? (a->i & b->i) : (a->i | b->i);
      break;
    default: return BO_ERR_PLAN;
  }
  *r = v;
  return BO_OK;
}

static xval x_load(const bo_col* c, int64_t i) {
  xval v = {c->dtype, 0, 0.0, 0, 0};
  switch (c->dtype) {
    case BO_BOOL: v.i = ((const uint8_t*)c->data)[i] != 0; break;
    case BO_I32: case BO_DATE32: v.i = ((const int32_t*)c->data)[i]; break;
    case BO_I64: v.i = ((const int64_t*)c->data)[i]; break;
    case BO_TIMESTAMP: v.i = ((const int64_t*)c->data)[i]; v.u = c->scale; break;
    case BO_F32: v.f = ((const float*)c->data)[i]; break;
    case BO_F64: v.f = ((const double*)c->data)[i]; break;
    case BO_DEC128: {
      const uint64_t* w = (const uint64_t*)c->data;
      v.d = (__int128)(((unsigned __int128)w[2 * i + 1] << 64) | w[2 * i]);
      break;
    }
  }
  return v;
}

static void x_store(bo_col* c, int64_t i, const xval* v) {
  switch (c->dtype) {
    case BO_BOOL: ((uint8_t*)c->data)[i] = (uint8_t)v->i; break;
    case BO_I32: case BO_DATE32: ((int32_t*)c->data)[i] = (int32_t)v->i; break;
    case BO_I64: case BO_TIMESTAMP: ((int64_t*)c->data)[i] = v->i; break;
    case BO_F32: ((float*)c->data)[i] = (float)v->f; break;
    case BO_F64: ((double*)c->data)[i] = v->f; break;
    case BO_DEC128: {
      uint64_t* w = (uint64_t*)c->data;
      w[2 * i] = (uint64_t)(unsigned __int128)v->d;
      w[2 * i + 1] = (uint64_t)((unsigned __int128)v->d >> 64);
      break;
    }
  }
}

int32_t bo_project(const bo_col* inputs, int32_t n_inputs, const bo_expr_op* prog, int32_t n_ops, bo_col* outputs,
                   int32_t n_outputs, int64_t n_rows) {
  xval* r = (xval*)calloc(XREGS, sizeof(xval));
  uint32_t* deps = (uint32_t*)calloc(XREGS, sizeof(uint32_t));
  int32_t st = BO_OK;
  for (int64_t i = 0; i < n_rows && st == BO_OK; i++) {
    for (int k = 0; k < n_ops && st == BO_OK; k++) {
      const bo_expr_op* o = &prog[k];
      if (o->kind != X_STORE && (o->dst < 0 || o->dst >= XREGS)) { st = BO_ERR_PLAN; break; }
      xval* d = &r[o->kind == X_STORE ? 0 : o->dst];
      switch (o->kind) {
        case X_LOAD:
          if (o->col < 0 || o->col >= n_inputs || inputs[o->col].dtype != o->dtype) { st = BO_ERR_TYPE; break; }
          *d = x_load(&inputs[o->col], i);
          deps[o->dst] = 1u << o->col;
          break;
        case X_CONST: {
          xval v = {o->dtype, 0, 0.0, 0, 0};
          if (o->dtype == BO_F32) v.f = (float)o->lit_f64;
          else if (o->dtype == BO_F64) v.f = o->lit_f64;
          else if (o->dtype == BO_DEC128) v.d = (__int128)(((unsigned __int128)(uint64_t)o->lit_hi << 64) | (uint64_t)o->lit_i64);
          else if (o->dtype == BO_BOOL) v.i = o->lit_i64 != 0;
          else if (o->dtype == BO_TIMESTAMP) { v.i = o->lit_i64; v.u = o->op; }
          else v.i = x_wrap(o->dtype == BO_I64 ? BO_I64 : BO_I32, (uint64_t)o->lit_i64);
          *d = v;
          deps[o->dst] = 0;
          break;
        }
        case X_CAST: {
          const int dec_src = r[o->a].dt == BO_DEC128, dec_dst = o->dtype == BO_DEC128;
          if ((dec_src || dec_dst) && (r[o->a].dt == BO_BOOL || r[o->a].dt == BO_TIMESTAMP || o->dtype == BO_BOOL ||
                                       o->dtype == BO_DATE32 || o->dtype == BO_TIMESTAMP)) { st = BO_ERR_TYPE; break; }
          const int u = r[o->a].u;
          st = x_cast(d, &r[o->a], o->dtype, dec_src ? o->c : 0, dec_dst ? o->op : 0);
          d->u = u;
          deps[o->dst] = deps[o->a];
          break;
        }
        case X_EXTRACT:
        case X_DATE_TRUNC: {
          const xval* a = &r[o->a];
          if ((a->dt != BO_DATE32 && a->dt != BO_TIMESTAMP) || o->op < DF_YEAR || o->op > DF_SECOND ||
              (a->dt == BO_DATE32 && o->op >= DF_HOUR)) { st = BO_ERR_TYPE; break; }
          const int ts = a->dt == BO_TIMESTAMP;
          xval v = {BO_I64, 0, 0.0, 0, 0};
          if (o->kind == X_EXTRACT) {
            v.i = x_extract(a->i, ts, a->u, o->op);
          } else {
            v.dt = a->dt;
            v.u = a->u;
            v.i = x_trunc(a->i, ts, a->u, o->op);
            if (!ts) v.i = x_wrap(BO_I32, (uint64_t)v.i);
          }
          *d = v;
          deps[o->dst] = deps[o->a];
          break;
        }
        case X_BINARY: st = x_binary(d, &r[o->a], &r[o->b], o->op, o->dtype); deps[o->dst] = deps[o->a] | deps[o->b]; break;
        case X_SELECT:
          if (r[o->a].dt != BO_BOOL || r[o->b].dt != r[o->c].dt ||
              (r[o->b].dt == BO_TIMESTAMP && r[o->b].u != r[o->c].u)) { st = BO_ERR_TYPE; break; }
          *d = r[o->a].i ? r[o->b] : r[o->c];
          deps[o->dst] = deps[o->a] | deps[o->b] | deps[o->c];
          break;
        case X_NOT:
          if (r[o->a].dt != BO_BOOL) { st = BO_ERR_TYPE; break; }
          d->dt = BO_BOOL;
          d->i = !r[o->a].i;
          deps[o->dst] = deps[o->a];
          break;
        case X_IS_NULL: {
          int isnull = !is_valid(&inputs[o->col], i);
          d->dt = BO_BOOL;
          d->i = o->op ? isnull : !isnull;
          deps[o->dst] = 0;
          break;
        }
        case X_STORE: {
          bo_col* c = &outputs[o->col];
          if (c->dtype != r[o->a].dt || (c->dtype == BO_TIMESTAMP && c->scale != r[o->a].u)) { st = BO_ERR_TYPE; break; }
          x_store(c, i, &r[o->a]);
          if (c->validity) {
            int ok = 1;
            for (int q = 0; q < n_inputs; q++)
              if ((deps[o->a] >> q) & 1u) ok = ok && is_valid(&inputs[q], i);
            uint8_t* vb = (uint8_t*)c->validity;
            if (ok) vb[i >> 3] |= (uint8_t)(1u << (i & 7));
            else vb[i >> 3] &= (uint8_t)~(1u << (i & 7));
          }
          break;
        }
        default: st = BO_ERR_PLAN;
      }
    }
  }
  (void)n_outputs;
  free(r);
  free(deps);
  return st;
}
