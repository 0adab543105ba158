/*
 * datagen.c — seeded synthetic tables for the operator workloads.
 *
 * Every generator draws from one Xorshift64Star stream (the fixture PRNG of
 * tests/common/mod.rs:46-90: seed + 0x9E3779B97F4A7C15, shifts 12/25/27,
 * multiplier 0x2545F4914F6CDD1D, 53-bit unit float) and consumes a fixed
 * number of draws per row, so row r of a table depends only on (seed, r).
 * That lets the table be filled by many host threads: the xorshift state
 * update is linear over GF(2), so the state after k steps is M^k applied to
 * the seed state, and each thread jumps to its first row with a 64x64 bit
 * matrix power instead of stepping through the rows before it.
 *
 * Host-only code (gcc -O3 -fopenmp); it allocates nothing on the device.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define XS_MUL 0x2545F4914F6CDD1Dull

static uint64_t xs_seed(uint64_t seed) {
  uint64_t s = seed + 0x9E3779B97F4A7C15ull;
  return s ? s : 0xDEADBEEFCAFEBABEull;
}
static inline uint64_t xs_step(uint64_t x) {
  x ^= x >> 12;
  x ^= x << 25;
  x ^= x >> 27;
  return x;
}
static inline uint64_t xs_next(uint64_t* s) {
  *s = xs_step(*s);
  return *s * XS_MUL;
}
static inline double xs_unit(uint64_t* s) { return (double)(xs_next(s) >> 11) * (1.0 / 9007199254740992.0); }

/* ---- GF(2) jump-ahead ---- */
typedef struct { uint64_t col[64]; } gf2m; /* col[i] = image of basis bit i */

static uint64_t gf2_apply(const gf2m* a, uint64_t v) {
  uint64_t r = 0;
  while (v) {
    int i = __builtin_ctzll(v);
    r ^= a->col[i];
    v &= v - 1;
  }
  return r;
}
static void gf2_mul(gf2m* out, const gf2m* a, const gf2m* b) { /* out = a o b */
  gf2m t;
  for (int i = 0; i < 64; i++) t.col[i] = gf2_apply(a, b->col[i]);
  *out = t;
}
static void gf2_pow_step(gf2m* out, uint64_t k) { /* out = T^k */
  gf2m base, acc;
  for (int i = 0; i < 64; i++) {
    base.col[i] = xs_step(1ull << i);
    acc.col[i] = 1ull << i;
  }
  while (k) {
    if (k & 1) gf2_mul(&acc, &base, &acc);
    gf2_mul(&base, &base, &base);
    k >>= 1;
  }
  *out = acc;
}
/* state after `k` draws of the stream seeded with `seed` */
uint64_t bolt_gen_state_at(uint64_t seed, uint64_t k) {
  gf2m j;
  gf2_pow_step(&j, k);
  return gf2_apply(&j, xs_seed(seed));
}

/* Runs body(row, state*) for rows [0,n) with `draws` draws per row, the
 * stream starting `offset` draws after the seed. */
#define PARALLEL_ROWS(seed, offset, n, draws, ...)                                \
  do {                                                                            \
    int nt_ = 1;                                                                  \
    _Pragma("omp parallel") {                                                     \
      _Pragma("omp single") nt_ = omp_get_num_threads_safe();                     \
    }                                                                             \
    int64_t chunk_ = ((n) + nt_ - 1) / nt_;                                       \
    _Pragma("omp parallel for schedule(static, 1)")                               \
    for (int t_ = 0; t_ < nt_; t_++) {                                            \
      int64_t lo_ = (int64_t)t_ * chunk_;                                         \
      int64_t hi_ = lo_ + chunk_ < (n) ? lo_ + chunk_ : (n);                      \
      if (lo_ >= hi_) continue;                                                   \
      uint64_t st_ = bolt_gen_state_at((seed), (uint64_t)(offset) + (uint64_t)lo_ * (draws)); \
      for (int64_t r = lo_; r < hi_; r++) {                                       \
        uint64_t* S = &st_;                                                       \
        __VA_ARGS__                                                               \
      }                                                                           \
    }                                                                             \
  } while (0)

static int omp_get_num_threads_safe(void) {
#ifdef _OPENMP
  return omp_get_num_threads();
#else
  return 1;
#endif
}

/* Fill `out` with draws [offset, offset+n) of the stream (parallel). */
void bolt_gen_draws(uint64_t seed, uint64_t offset, int64_t n, uint64_t* out) {
  PARALLEL_ROWS(seed, offset, n, 1, { out[r] = xs_next(S); });
}

/* splitmix64 finalizer: a bijection on 64-bit words, used to turn dense
 * ranks into scattered unique int64 keys. */
uint64_t bolt_scramble64(uint64_t x) {
  x ^= x >> 30; x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27; x *= 0x94d049bb133111ebull;
  x ^= x >> 31;
  return x;
}
/* 32-bit bijection (lowbias32-style xorshift-multiply) for int32 keys. */
uint32_t bolt_scramble32(uint32_t x) {
  x ^= x >> 16; x *= 0x7feb352dU;
  x ^= x >> 15; x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

/* ---- plain GROUP BY table: k int32 = draw % card, v int64 = draw % vmod ---- */
void bolt_gen_groupby_small(uint64_t seed, int64_t n, uint64_t card, uint64_t vmod, int32_t* k,
                            int64_t* v) {
  PARALLEL_ROWS(seed, 0, n, 2, {
    k[r] = (int32_t)(xs_next(S) % card);
    v[r] = (int64_t)(xs_next(S) % vmod);
  });
}

/* ---- join build side: key = scramble(r) (unique), payload = draw % pmod ---- */
void bolt_gen_join_build(uint64_t seed, int64_t n_build, uint64_t pmod, int64_t* key, int32_t* payload) {
  PARALLEL_ROWS(seed, 0, n_build, 1, {
    key[r] = (int64_t)bolt_scramble64((uint64_t)r);
    payload[r] = (int32_t)(xs_next(S) % pmod);
  });
}

/* ---- join probe side (stream continues after the build's n_build draws):
 *      draw 1: hit iff top bit set; draw 2: index; hit -> scramble(idx % n_build),
 *      miss -> scramble(n_build + idx % n_build) (never a build key, the
 *      scramble being a bijection); draw 3: amount = unit * amount_scale ---- */
void bolt_gen_join_probe(uint64_t seed, int64_t n_build, int64_t n_probe, double amount_scale,
                         int64_t* fk, double* amount) {
  PARALLEL_ROWS(seed, (uint64_t)n_build, n_probe, 3, {
    uint64_t coin = xs_next(S);
    uint64_t idx = xs_next(S) % (uint64_t)n_build;
    fk[r] = (int64_t)bolt_scramble64((coin >> 63) ? idx : (uint64_t)n_build + idx);
    amount[r] = xs_unit(S) * amount_scale;
  });
}

/* ---- sort input: key = raw draw over the full int64 range, rowid = r ---- */
void bolt_gen_sort_pairs(uint64_t seed, int64_t n, int64_t* key, uint32_t* rowid) {
  PARALLEL_ROWS(seed, 0, n, 1, {
    key[r] = (int64_t)xs_next(S);
    rowid[r] = (uint32_t)r;
  });
}

/* ---- Zipf(s) over ranks 1..n_ranks by inversion with a guide table
 *      (Chen & Asau indexed search): one uniform per sample. ---- */
typedef struct {
  double* cdf;     /* cdf[k] = P(rank <= k+1) */
  uint32_t* guide; /* guide[j] = first k with cdf[k] > j / m */
  int64_t n, m;
} zipf_t;

static void zipf_init(zipf_t* z, int64_t n_ranks, double s) {
  z->n = n_ranks;
  z->m = n_ranks;
  z->cdf = (double*)malloc((size_t)n_ranks * 8);
  z->guide = (uint32_t*)malloc((size_t)z->m * 4);
  double acc = 0;
  for (int64_t k = 0; k < n_ranks; k++) {
    acc += pow((double)(k + 1), -s);
    z->cdf[k] = acc;
  }
  for (int64_t k = 0; k < n_ranks; k++) z->cdf[k] /= acc;
  z->cdf[n_ranks - 1] = 1.0;
  int64_t k = 0;
  for (int64_t j = 0; j < z->m; j++) {
    double t = (double)j / (double)z->m;
    while (k < n_ranks - 1 && z->cdf[k] <= t) k++;
    z->guide[j] = (uint32_t)k;
  }
}
static inline int64_t zipf_rank(const zipf_t* z, double u) { /* 0-based rank */
  int64_t j = (int64_t)(u * (double)z->m);
  if (j >= z->m) j = z->m - 1;
  int64_t k = z->guide[j];
  while (k < z->n - 1 && z->cdf[k] <= u) k++;
  return k;
}
static void zipf_free(zipf_t* z) { free(z->cdf); free(z->guide); }

/* ---- high-cardinality GROUP BY input: key = scramble(rank) with rank ~
 *      Zipf(s) over n_ranks; value ~ N(0,1) by Box-Muller (cos branch) from
 *      two further uniforms.  Three draws per row. ---- */
void bolt_gen_zipf_groupby(uint64_t seed, int64_t n, int64_t n_ranks, double s, int64_t* key, double* val) {
  zipf_t z;
  zipf_init(&z, n_ranks, s);
  PARALLEL_ROWS(seed, 0, n, 3, {
    int64_t rk = zipf_rank(&z, xs_unit(S));
    double u1 = xs_unit(S), u2 = xs_unit(S);
    key[r] = (int64_t)bolt_scramble64((uint64_t)rk);
    val[r] = sqrt(-2.0 * log(1.0 - u1)) * cos(6.283185307179586 * u2);
  });
  zipf_free(&z);
}

/* ---- batched-query input: int32 key = scramble32(Zipf rank over card),
 *      v_i64 = draw % 2001 - 1000, v_f32 = unit*2-1.  Three draws per row. ---- */
void bolt_gen_zipf_small(uint64_t seed, int64_t n, int64_t card, double s, int32_t* key, int64_t* v_i64,
                         float* v_f32) {
  zipf_t z;
  zipf_init(&z, card, s);
  PARALLEL_ROWS(seed, 0, n, 3, {
    int64_t rk = zipf_rank(&z, xs_unit(S));
    key[r] = (int32_t)bolt_scramble32((uint32_t)rk);
    v_i64[r] = (int64_t)(xs_next(S) % 2001) - 1000;
    v_f32[r] = (float)(xs_unit(S) * 2.0 - 1.0);
  });
  zipf_free(&z);
}

/* ---- served-form star schema fact table: three int32 foreign keys uniform
 *      over the dimension sizes, price Decimal128(12,2) = cents uniform over
 *      [100, 1000000) as (lo, hi) int64 pairs, qty float64 uniform [1, 50),
 *      Date32 uniform over 2557 days (7 years) from day `date0`.  Six draws
 *      per row. ---- */
void bolt_gen_star_fact(uint64_t seed, int64_t n, int64_t n1, int64_t n2, int64_t n3, int32_t date0, int32_t* fk1,
                        int32_t* fk2, int32_t* fk3, int64_t* price, double* qty, int32_t* date) {
  PARALLEL_ROWS(seed, 0, n, 6, {
    fk1[r] = (int32_t)(xs_next(S) % (uint64_t)n1);
    fk2[r] = (int32_t)(xs_next(S) % (uint64_t)n2);
    fk3[r] = (int32_t)(xs_next(S) % (uint64_t)n3);
    price[2 * r] = (int64_t)(100 + xs_next(S) % (1000000 - 100));
    price[2 * r + 1] = 0;
    qty[r] = 1.0 + xs_unit(S) * 49.0;
    date[r] = date0 + (int32_t)(xs_next(S) % 2557);
  });
}
