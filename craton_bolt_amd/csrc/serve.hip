// serve.hip — served form: request loop, HBM-budget admission and
// continuous batching over a device-resident star schema.
//
// No reference counterpart (the reference's Flight server runs one query at
// a time behind Mutex<Engine>, src/flight/mod.rs:86,129-168).
//
// One server thread owns a stream.  The fact table is scanned in chunks by a
// shared circular scan: every step one launch processes the current chunk
// for every active query (each fact row is loaded once and then filtered,
// joined against the dimension hash tables and aggregated into each query's
// own HBM hash table with the exact atomics of groupby_dev.h).  Requests are
// admitted FIFO at chunk boundaries while a slot is free and their HBM need
// fits the budget next to the queries already running; a query starts at the
// current chunk, wraps around, and completes after one full lap, when its
// groups are extracted, ordered (ORDER BY aggregate DESC, group key ASC, via
// bolt_sort_limit: the first LIMIT rows of the radix-sort order) and those rows
// are copied out.  Its slot and memory are then free for the next request.
#include <algorithm>
#include <cmath>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/bolt_serve.h"
#include "groupby_dev.h"
#include "join_dev.h"

namespace bolt {

constexpr int kServeMaxAttrs = 8;
constexpr int kServeBlock = 256;

struct SDim {
  const uint4* tbl;
  uint64_t mask;
  const void* attr[kServeMaxAttrs];
  int32_t attr_dt[kServeMaxAttrs];
};
struct SDims {
  SDim d[BOLT_SERVE_MAX_DIMS];
};
struct SJoin {
  const void* fk;
  int32_t fk_dt;
  int32_t dim;
};
struct SGroupKey {
  const void* attr;
  int32_t dt;
  int32_t join;
};
// Up to kPreCols fact columns that the active queries' predicates read are
// loaded once per row for all of them (the union, rebuilt when the active
// set changes); a query's term then reads registers, so rows that fail every
// predicate cost a few ALU operations per query and no dependent loads.
constexpr int kPreCols = 4;
struct SPre {
  const void* d[kPreCols];
  const uint8_t* v[kPreCols];
  int32_t dt[kPreCols];
  int32_t n;
};
struct SQ {
  DevFilter f;
  int8_t term_slot[BOLT_MAX_PRED_TERMS];  // preloaded column of each term, -1 = read it directly
  // The conjunction as inclusive intervals over the order-preserving 128-bit
  // image of each preloaded column (pre_image): row passes iff, for every
  // slot c in cmask, lo[c] <= image_c <= hi[c], and every slot in vmask is
  // non-NULL.  generic = 1 when a term is not such an interval (NE, IS NULL,
  // a column outside the preloaded set): the terms are then evaluated one by
  // one.
  uint64_t lo_h[kPreCols], lo_l[kPreCols], hi_h[kPreCols], hi_l[kPreCols];
  uint32_t cmask, vmask;
  int32_t generic;
  GPlan P;
  SJoin j[BOLT_SERVE_MAX_DIMS];
  SGroupKey g[2];
  int32_t n_joins, n_group;
  GTable t;
  uint64_t stride;  // words per accumulator field (cap + 2)
};

__device__ __forceinline__ uint64_t attr_bits(const void* p, int dt, uint32_t row) {
  return dt == BOLT_I64 ? static_cast<const uint64_t*>(p)[row] : (uint64_t) static_cast<const uint32_t*>(p)[row];
}
__device__ __forceinline__ int64_t fk_value(const void* p, int dt, int64_t row) {
  return dt == BOLT_I64 ? static_cast<const int64_t*>(p)[row] : (int64_t) static_cast<const int32_t*>(p)[row];
}


// Step, part 1: every active query's predicate over one chunk.  Each wave
// owns rows [r0 + w*per_w, r0 + (w+1)*per_w) of the chunk (w its global
// wave index) and appends the rows that pass query q, in order, to its own
// segment (q, w) of the row lists.  The running length of segment (q, w) is
// lane q of one register (nq <= 64), so the append is a ballot, a popcount
// and a lane read: no LDS, no barrier, no global atomics.  Each lane takes
// kFilterRows rows per iteration, loading the predicate columns of all
// queries once per row, so a query's terms are fetched once per
// 64 * kFilterRows rows.  Part 2 then runs the joins and aggregation over
// those dense lists with every lane busy, instead of the few lanes per wave
// (a few percent of rows) that pass a selective predicate.
constexpr int kServeGrid = 2048;
constexpr int kServeWaves = kServeBlock / 64;
constexpr int64_t kServeAllWaves = (int64_t)kServeGrid * kServeWaves;
constexpr int kFilterRows = 4;

// rows per wave segment for a chunk of `chunk_rows`
inline int64_t serve_per_wave(int64_t chunk_rows) {
  const int64_t step = 64 * kFilterRows;
  return ((chunk_rows + kServeAllWaves - 1) / kServeAllWaves + step - 1) / step * step;
}

// Order-preserving 128-bit image (h, l) of a column value, under which every
// comparison of the column with a literal of its type is an interval:
// integers and Bool by sign flip; floats by the IEEE total-order map with
// -0.0 taken as +0.0 (NaN maps to the top, above the image of +inf, so no
// interval with a finite or infinite bound holds it); Decimal128 as the
// sign-flipped high word, then the low word.
__host__ __device__ __forceinline__ void value_image(int dt, int64_t vi, int64_t vhi, double vf, uint64_t* h,
                                                     uint64_t* l) {
  if (dt == BOLT_F32 || dt == BOLT_F64) {
    *h = 0;
    *l = f64_order_key(vf == 0.0 ? 0.0 : vf);
  } else if (dt == BOLT_DEC128) {
    *h = (uint64_t)vhi ^ 0x8000000000000000ull;
    *l = (uint64_t)vi;
  } else {
    *h = 0;
    *l = (uint64_t)vi ^ 0x8000000000000000ull;
  }
}

__device__ __forceinline__ void pre_image(const Raw& r, int dt, uint64_t* h, uint64_t* l) {
  int64_t vi, vhi;
  double vf;
  raw_value(r, dt, vi, vhi, vf);
  value_image(dt, vi, vhi, vf, h, l);
}

// Images of preloaded slot C for the kFilterRows rows rc[] into h[] / l[]
// (one uniform dtype switch per slot and iteration; branch-free per row).
// A macro rather than a function so that the arrays stay in registers.
#define BOLT_PRELOAD_IMAGES(C, h, l)                                                                         \
  do {                                                                                                       \
    const void* d_ = pre.d[C];                                                                               \
    switch (pre.dt[C]) {                                                                                     \
      case BOLT_I32: case BOLT_DATE32:                                                                       \
        _Pragma("unroll") for (int r = 0; r < kFilterRows; r++) {                                            \
          const int32_t v_ = (int32_t)__builtin_nontemporal_load(static_cast<const uint32_t*>(d_) + rc[r]); \
          h[r] = 0;                                                                                          \
          l[r] = (uint64_t)(int64_t)v_ ^ 0x8000000000000000ull;                                              \
        }                                                                                                    \
        break;                                                                                               \
      case BOLT_I64:                                                                                         \
        _Pragma("unroll") for (int r = 0; r < kFilterRows; r++) {                                            \
          h[r] = 0;                                                                                          \
          l[r] = __builtin_nontemporal_load(static_cast<const uint64_t*>(d_) + rc[r]) ^ 0x8000000000000000ull; \
        }                                                                                                    \
        break;                                                                                               \
      case BOLT_F32:                                                                                         \
        _Pragma("unroll") for (int r = 0; r < kFilterRows; r++) {                                            \
          const double v_ =                                                                                  \
              (double)__uint_as_float(__builtin_nontemporal_load(static_cast<const uint32_t*>(d_) + rc[r])); \
          h[r] = 0;                                                                                          \
          l[r] = f64_order_key(v_ == 0.0 ? 0.0 : v_);                                                        \
        }                                                                                                    \
        break;                                                                                               \
      case BOLT_F64:                                                                                         \
        _Pragma("unroll") for (int r = 0; r < kFilterRows; r++) {                                            \
          const double v_ = as_f64(__builtin_nontemporal_load(static_cast<const uint64_t*>(d_) + rc[r]));    \
          h[r] = 0;                                                                                          \
          l[r] = f64_order_key(v_ == 0.0 ? 0.0 : v_);                                                        \
        }                                                                                                    \
        break;                                                                                               \
      case BOLT_DEC128:                                                                                      \
        _Pragma("unroll") for (int r = 0; r < kFilterRows; r++) {                                            \
          const uint64_t* w_ = static_cast<const uint64_t*>(d_) + 2 * rc[r];                                 \
          l[r] = __builtin_nontemporal_load(w_);                                                             \
          h[r] = __builtin_nontemporal_load(w_ + 1) ^ 0x8000000000000000ull;                                 \
        }                                                                                                    \
        break;                                                                                               \
      default:                                                                                               \
        _Pragma("unroll") for (int r = 0; r < kFilterRows; r++) {                                            \
          h[r] = 0;                                                                                          \
          l[r] = (uint64_t)static_cast<const uint8_t*>(d_)[rc[r]] ^ 0x8000000000000000ull;                   \
        }                                                                                                    \
        break;                                                                                               \
    }                                                                                                        \
    if (pre.v[C]) {                                                                                          \
      _Pragma("unroll") for (int r = 0; r < kFilterRows; r++) if (!bit_valid(pre.v[C], rc[r])) vm[r] &= ~(1u << C); \
    }                                                                                                        \
  } while (0)

__device__ __forceinline__ bool in_interval(uint64_t h, uint64_t l, uint64_t lh, uint64_t ll, uint64_t hh,
                                            uint64_t hl) {
  const bool ge = h > lh || (h == lh && l >= ll);
  const bool le = h < hh || (h == hh && l <= hl);
  return ge && le;
}

__global__ __launch_bounds__(kServeBlock) void serve_filter_kernel(const SQ* __restrict__ qs, int nq, SPre pre,
                                                                   int64_t r0, int64_t r1, int64_t per_w,
                                                                   uint32_t* __restrict__ lists,
                                                                   uint32_t* __restrict__ counts) {
  const int lane = threadIdx.x & 63;
  const int64_t gw = (int64_t)blockIdx.x * kServeWaves + (threadIdx.x >> 6);
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const int64_t w0 = r0 + gw * per_w;
  const int64_t w1 = min(r1, w0 + per_w);
  uint32_t run = 0;  // lane q: rows appended to segment (q, gw)
  for (int64_t base = w0; base < w1; base += 64 * kFilterRows) {  // wave-uniform trip count
    // each preloaded value's image, computed once per row for every query;
    // the dtype switch is taken once per column and iteration, not per row
    uint64_t h0[kFilterRows], l0[kFilterRows], h1[kFilterRows], l1[kFilterRows];
    uint64_t h2[kFilterRows], l2[kFilterRows], h3[kFilterRows], l3[kFilterRows];
    uint32_t vm[kFilterRows];  // bit c: preloaded slot c is non-NULL
    int64_t rc[kFilterRows];
    bool in[kFilterRows];
#pragma unroll
    for (int r = 0; r < kFilterRows; r++) {
      const int64_t row = base + r * 64 + lane;
      in[r] = row < w1;
      rc[r] = in[r] ? row : w1 - 1;  // clamped: every load is in range
      h0[r] = l0[r] = h1[r] = l1[r] = h2[r] = l2[r] = h3[r] = l3[r] = 0;
      vm[r] = 0xFu;
    }
    if (pre.n > 0) BOLT_PRELOAD_IMAGES(0, h0, l0);
    if (pre.n > 1) BOLT_PRELOAD_IMAGES(1, h1, l1);
    if (pre.n > 2) BOLT_PRELOAD_IMAGES(2, h2, l2);
    if (pre.n > 3) BOLT_PRELOAD_IMAGES(3, h3, l3);
    for (int qi = 0; qi < nq; qi++) {
      const SQ& Q = qs[qi];
      uint32_t* seg = lists + ((uint64_t)qi * kServeAllWaves + gw) * (uint64_t)per_w;
      const uint32_t cm = Q.cmask, vq = Q.vmask;
      const bool generic = Q.generic != 0;
#pragma unroll
      for (int r = 0; r < kFilterRows; r++) {
        bool pass = in[r];
        if (generic) {
          for (int t = 0; t < Q.f.n; t++) pass = pass && eval_term(Q.f.t[t], rc[r]);
        } else {
          pass = pass && (vm[r] & vq) == vq;
          if (cm & 1u) pass = pass && in_interval(h0[r], l0[r], Q.lo_h[0], Q.lo_l[0], Q.hi_h[0], Q.hi_l[0]);
          if (cm & 2u) pass = pass && in_interval(h1[r], l1[r], Q.lo_h[1], Q.lo_l[1], Q.hi_h[1], Q.hi_l[1]);
          if (cm & 4u) pass = pass && in_interval(h2[r], l2[r], Q.lo_h[2], Q.lo_l[2], Q.hi_h[2], Q.hi_l[2]);
          if (cm & 8u) pass = pass && in_interval(h3[r], l3[r], Q.lo_h[3], Q.lo_l[3], Q.hi_h[3], Q.hi_l[3]);
        }
        const uint64_t m = __ballot(pass);
        if (m) {
          const uint32_t off = (uint32_t)__builtin_amdgcn_readlane((int)run, qi);
          if (pass) seg[off + (uint32_t)__popcll(m & lt)] = (uint32_t)(rc[r] - r0);
          if (lane == qi) run += (uint32_t)__popcll(m);
        }
      }
    }
  }
  if (lane < nq) counts[(uint64_t)lane * kServeAllWaves + gw] = run;
}

// Step, part 2: block b runs the joins and aggregation for the rows in the
// segments (q, w) of its own waves w = 4b .. 4b + 3, taken as one index
// space so that every lane has a row.
__global__ __launch_bounds__(kServeBlock) void serve_probe_kernel(const SQ* __restrict__ qs, int nq, SDims dims,
                                                                  int64_t r0, int64_t per_w,
                                                                  const uint32_t* __restrict__ lists,
                                                                  const uint32_t* __restrict__ counts) {
  for (int qi = 0; qi < nq; qi++) {
    const SQ& Q = qs[qi];
    const uint64_t seg0 = (uint64_t)qi * kServeAllWaves + (uint64_t)blockIdx.x * kServeWaves;
    uint32_t pfx[kServeWaves + 1];
    pfx[0] = 0;
#pragma unroll
    for (int w = 0; w < kServeWaves; w++) pfx[w + 1] = pfx[w] + counts[seg0 + w];
    const uint32_t cnt = pfx[kServeWaves];
    for (uint32_t i = threadIdx.x; i < cnt; i += kServeBlock) {
      int w = 0;
#pragma unroll
      for (int k = 1; k < kServeWaves; k++) w += i >= pfx[k] ? 1 : 0;
      const uint32_t off = i - (w == 0 ? 0u : w == 1 ? pfx[1] : w == 2 ? pfx[2] : pfx[3]);
      const int64_t row = r0 + (int64_t)lists[(seg0 + w) * (uint64_t)per_w + off];
      uint32_t d0 = 0, d1 = 0, d2 = 0;
      bool ok = true;
      for (int j = 0; j < Q.n_joins && ok; j++) {
        const SDim& D = dims.d[Q.j[j].dim];
        const int64_t k = fk_value(Q.j[j].fk, Q.j[j].fk_dt, row);
        bool hit = false;
        uint32_t v = 0;
        probe_chain(D.tbl, D.mask, k, true, [&](uint32_t wv) {
          hit = true;
          v = wv;
        });
        ok = hit;
        if (j == 0) d0 = v;
        else if (j == 1) d1 = v;
        else d2 = v;
      }
      if (!ok) continue;
      auto drow = [&](int j) { return j == 0 ? d0 : (j == 1 ? d1 : d2); };
      uint64_t key = attr_bits(Q.g[0].attr, Q.g[0].dt, drow(Q.g[0].join));
      if (Q.n_group == 2)
        key = (key << 32) | (attr_bits(Q.g[1].attr, Q.g[1].dt, drow(Q.g[1].join)) & 0xFFFFFFFFull);
      const int64_t sl = gslot(Q.t, 0, key);
      if (sl >= 0) update_slot(Q.P, Q.t.words, Q.stride, (uint32_t)sl, row);
    }
  }
}

__global__ void serve_init_kernel(GTable t, GPlan P, uint64_t stride) {
  const size_t total = t.cap + 2;
  const size_t st = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += st) {
    if (i < t.cap) t.keys[i] = kEmpty;
    for (int w = 0; w < P.words; w++) t.words[(size_t)w * stride + i] = word_identity(P, w);
  }
  if (blockIdx.x == 0 && threadIdx.x < 2) t.used_special[threadIdx.x] = 0;
  if (blockIdx.x == 0 && threadIdx.x < 4) t.err[threadIdx.x] = 0;
}

struct SOut {
  OutCol key[2];
  OutCol agg[BOLT_SERVE_MAX_AGGS];
};

// Occupied slots -> dense output columns (hash order).
__global__ __launch_bounds__(256) void serve_extract_kernel(GTable t, GPlan P, uint64_t stride, int n_group, SOut o,
                                                            unsigned long long* __restrict__ counter) {
  const uint64_t total = t.cap + 2;
  const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t rounds = (total + step - 1) / step;  // uniform trip count (wave_append is wave-wide)
  for (uint64_t r = 0; r < rounds; r++) {
    const uint64_t s = r * step + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool occ = false;
    uint64_t packed = 0;
    if (s < t.cap) {
      packed = t.keys[s];
      occ = packed != kEmpty;
    } else if (s < total) {
      occ = t.used_special[s - t.cap] != 0;
      packed = kEmpty;
    }
    const unsigned long long pos = wave_append(occ, counter);
    if (!occ) continue;
    if (n_group == 1) {
      write_key(o.key[0], pos, packed);
    } else {
      write_key(o.key[0], pos, packed >> 32);
      write_key(o.key[1], pos, packed & 0xFFFFFFFFull);
    }
    for (int a = 0; a < P.n; a++) finalize_to(t, P.a[a], o.agg[a], stride, s, pos);
  }
}

// Host: a query's terms on preloaded columns as intervals over value_image
// (a conjunction of comparisons on one column is one interval).
static void compile_intervals(SQ& d) {
  d.cmask = d.vmask = 0;
  d.generic = 0;
  for (int c = 0; c < kPreCols; c++) {
    d.lo_h[c] = d.lo_l[c] = 0;
    d.hi_h[c] = d.hi_l[c] = ~0ull;
  }
  auto less = [](uint64_t ah, uint64_t al, uint64_t bh, uint64_t bl) { return ah < bh || (ah == bh && al < bl); };
  for (int t = 0; t < d.f.n; t++) {
    const DevTerm& tm = d.f.t[t];
    const int sl = d.term_slot[t];
    if (sl < 0 || tm.op == BOLT_CMP_IS_NULL || tm.op == BOLT_CMP_NE) {
      d.generic = 1;
      continue;
    }
    d.vmask |= 1u << sl;
    if (tm.op == BOLT_CMP_IS_NOT_NULL) continue;
    const bool fl = tm.dtype == BOLT_F32 || tm.dtype == BOLT_F64;
    if (fl && std::isnan(tm.lit_f64)) {  // every comparison with NaN is false
      d.generic = 1;
      continue;
    }
    if (!(d.cmask & (1u << sl)) && fl) {  // a constrained float column never holds NaN
      uint64_t h, l;
      value_image(tm.dtype, 0, 0, HUGE_VAL, &h, &l);
      d.hi_h[sl] = h;
      d.hi_l[sl] = l;
    }
    d.cmask |= 1u << sl;
    uint64_t h, l;
    value_image(tm.dtype, tm.lit_i64, tm.lit_hi, tm.lit_f64, &h, &l);
    uint64_t lh = h, ll = l, hh = h, hl = l;  // the term's own interval
    bool lower = true, upper = true, empty = false;
    switch (tm.op) {
      case BOLT_CMP_GE: upper = false; break;
      case BOLT_CMP_LE: lower = false; break;
      case BOLT_CMP_GT:
        upper = false;
        if (h == ~0ull && l == ~0ull) empty = true;
        else { ll = l + 1; lh = h + (ll == 0 ? 1 : 0); }
        break;
      case BOLT_CMP_LT:
        lower = false;
        if (h == 0 && l == 0) empty = true;
        else { hl = l - 1; hh = h - (l == 0 ? 1 : 0); }
        break;
      default: break;  // EQ: both bounds
    }
    if (empty) {  // lo > hi: nothing passes
      d.lo_h[sl] = d.lo_l[sl] = ~0ull;
      d.hi_h[sl] = d.hi_l[sl] = 0;
      continue;
    }
    if (lower && less(d.lo_h[sl], d.lo_l[sl], lh, ll)) { d.lo_h[sl] = lh; d.lo_l[sl] = ll; }
    if (upper && less(hh, hl, d.hi_h[sl], d.hi_l[sl])) { d.hi_h[sl] = hh; d.hi_l[sl] = hl; }
  }
}

__global__ void serve_gather_kernel(const uint8_t* __restrict__ src, int width, const uint32_t* __restrict__ perm,
                                    int64_t n, uint8_t* __restrict__ dst) {
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += st) {
    const int64_t p = perm[i];
    for (int b = 0; b < width; b += 4)
      *reinterpret_cast<uint32_t*>(dst + i * width + b) = *reinterpret_cast<const uint32_t*>(src + p * width + b);
  }
}

using SClock = std::chrono::steady_clock;

struct SReq {
  int64_t ticket;
  bolt_serve_query_t q;
  double submit_ms;
  int attempt;
  int64_t need;
  uint64_t cap;
  uint64_t min_cap;  // on a re-run: 4x the table that filled up
};

struct SActive {
  SReq r;
  double admit_ms = 0;
  int64_t chunks_done = 0;
  GPlan P;
  int32_t out_dt[kMaxAggs];
  DevBuf<unsigned long long> keys, words;
  DevBuf<uint32_t> misc;  // used_special[2], err[4]
  SQ desc;
};

struct SDone {
  bool ready = false;
  bolt_serve_result_t res{};
  std::vector<bolt_scalar_t> cells;
  std::string err;
};

}  // namespace bolt

using namespace bolt;

struct bolt_server {
  hipStream_t s = nullptr;
  int dev = 0;
  std::vector<bolt_column_t> fact;
  int64_t n_rows = 0;
  std::vector<bolt_join_table_t*> dim_tables;
  std::vector<std::vector<bolt_column_t>> dim_attrs;
  SDims sdims;
  bolt_serve_config_t cfg{};
  int64_t n_chunks = 0;
  int64_t cur_chunk = 0;
  SClock::time_point t0;
  // request state
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<SReq> queue;
  std::map<int64_t, SDone> done;
  int64_t next_ticket = 1;
  bool stop = false;
  // server-thread state
  std::list<std::unique_ptr<SActive>> active;
  int64_t bytes_in_use = 0;
  SQ* h_desc = nullptr;
  DevBuf<SQ> d_desc;
  DevBuf<uint32_t> d_lists;   // [slots][grid][per] chunk-local rows passing each query's predicate
  DevBuf<uint32_t> d_counts;  // [slots][grid]
  SPre pre{};
  bool desc_dirty = true;
  std::thread th;
  // finisher: extraction / ORDER BY / copy-out of queries that completed
  // their lap run here, on their own stream, while the scan goes on
  hipStream_t fs = nullptr;
  std::thread fin_th;
  std::deque<std::unique_ptr<SActive>> fin_q;  // under mu
  int fin_pending = 0;                          // under mu
  bool fin_stop = false;                        // under mu
  std::condition_variable cv_fin;
  std::atomic<int64_t> step_us{0}, finish_us{0};
  // stats
  std::atomic<int64_t> steps{0}, peak_bytes{0};
  int64_t completed = 0;  // under mu
  std::atomic<int32_t> peak_active{0};

  double now_ms() const { return std::chrono::duration<double, std::milli>(SClock::now() - t0).count(); }

  ~bolt_server() {
    if (h_desc) (void)hipHostFree(h_desc);
  }

  // Validates the query against the schema; fills cap / need.  Runs on the
  // submitting thread (no device work).
  int plan_request(SReq& r) {
    const bolt_serve_query_t& q = r.q;
    if (q.n_terms < 0 || q.n_terms > BOLT_SERVE_MAX_TERMS) return set_error(BOLT_ERR_PLAN, "serve: bad n_terms");
    if (q.n_joins < 1 || q.n_joins > BOLT_SERVE_MAX_DIMS) return set_error(BOLT_ERR_PLAN, "serve: 1..3 joins required");
    for (int j = 0; j < q.n_joins; j++) {
      if (q.join_fk[j] < 0 || q.join_fk[j] >= (int)fact.size())
        return set_error(BOLT_ERR_PLAN, "serve: join %d: bad fact column", j);
      const int dt = fact[q.join_fk[j]].dtype;
      if (dt != BOLT_I32 && dt != BOLT_I64) return set_error(BOLT_ERR_TYPE, "serve: join key must be Int32 / Int64");
      if (q.join_dim[j] < 0 || q.join_dim[j] >= (int)dim_tables.size())
        return set_error(BOLT_ERR_PLAN, "serve: join %d: bad dimension", j);
      if (fact[q.join_fk[j]].validity) return set_error(BOLT_ERR_UNSUPPORTED, "serve: nullable foreign keys");
    }
    if (q.n_group < 1 || q.n_group > 2) return set_error(BOLT_ERR_PLAN, "serve: 1 or 2 group attributes");
    for (int g = 0; g < q.n_group; g++) {
      if (q.group_dim[g] < 0 || q.group_dim[g] >= q.n_joins) return set_error(BOLT_ERR_PLAN, "serve: bad group join");
      const auto& attrs = dim_attrs[q.join_dim[q.group_dim[g]]];
      if (q.group_attr[g] < 0 || q.group_attr[g] >= (int)attrs.size())
        return set_error(BOLT_ERR_PLAN, "serve: bad group attribute");
      const int dt = attrs[q.group_attr[g]].dtype;
      if (q.n_group == 2 && dtype_size(dt) != 4) return set_error(BOLT_ERR_UNSUPPORTED, "serve: two-column keys must be 32-bit");
    }
    if (q.n_aggs < 1 || q.n_aggs > BOLT_SERVE_MAX_AGGS) return set_error(BOLT_ERR_PLAN, "serve: 1..4 aggregates");
    if (q.order_agg < 0 || q.order_agg >= q.n_aggs) return set_error(BOLT_ERR_PLAN, "serve: bad ORDER BY aggregate");
    if (q.limit < 1 || q.limit > BOLT_SERVE_MAX_LIMIT) return set_error(BOLT_ERR_PLAN, "serve: LIMIT 1..1024");
    GPlan P;
    int32_t odt[kMaxAggs];
    BOLT_TRY(make_gplan(q.aggs, q.n_aggs, fact.data(), (int32_t)fact.size(), n_rows, &P, odt));
    for (int a = 0; a < q.n_aggs; a++)
      if (P.a[a].kind == K_VAR) return set_error(BOLT_ERR_UNSUPPORTED, "serve: VAR / STDDEV aggregates");
    const int od = odt[q.order_agg];
    if (od == BOLT_BOOL) return set_error(BOLT_ERR_UNSUPPORTED, "serve: ORDER BY a Bool aggregate");
    bolt_filter_t flt{q.terms, q.n_terms};
    DevFilter f;
    BOLT_TRY(make_dev_filter(&flt, fact.data(), (int32_t)fact.size(), n_rows, &f));
    const double est = q.est_groups > 0 ? (double)q.est_groups : 1024.0;
    r.cap = std::max(next_pow2((uint64_t)std::max(1024.0, est * 2.5)), r.min_cap);
    const int64_t table = (int64_t)(r.cap + 2) * 8 * (1 + P.words) + 8 * (int64_t)r.cap;
    const int64_t scratch = (int64_t)(r.cap + 2) * (16 * (q.n_group + q.n_aggs) + 40);
    r.need = table + scratch;
    return BOLT_OK;
  }

  int admit(SActive& A) {
    const bolt_serve_query_t& q = A.r.q;
    BOLT_TRY(make_gplan(q.aggs, q.n_aggs, fact.data(), (int32_t)fact.size(), n_rows, &A.P, A.out_dt));
    bolt_filter_t flt{q.terms, q.n_terms};
    SQ& d = A.desc;
    memset(&d, 0, sizeof d);
    BOLT_TRY(make_dev_filter(&flt, fact.data(), (int32_t)fact.size(), n_rows, &d.f));
    d.P = A.P;
    d.n_joins = q.n_joins;
    for (int j = 0; j < q.n_joins; j++) {
      d.j[j].fk = fact[q.join_fk[j]].data;
      d.j[j].fk_dt = fact[q.join_fk[j]].dtype;
      d.j[j].dim = q.join_dim[j];
    }
    d.n_group = q.n_group;
    for (int g = 0; g < q.n_group; g++) {
      const bolt_column_t& c = dim_attrs[q.join_dim[q.group_dim[g]]][q.group_attr[g]];
      d.g[g].attr = c.data;
      d.g[g].dt = c.dtype;
      d.g[g].join = q.group_dim[g];
    }
    const uint64_t cap = A.r.cap;
    BOLT_TRY(A.keys.alloc(cap, s));
    BOLT_TRY(A.words.alloc((cap + 2) * (size_t)(A.P.words > 0 ? A.P.words : 1), s));
    BOLT_TRY(A.misc.alloc(8, s));
    d.t.keys = A.keys.ptr;
    d.t.words = A.words.ptr;
    d.t.used_special = A.misc.ptr;
    d.t.err = A.misc.ptr + 2;
    d.t.cap = cap;
    d.stride = cap + 2;
    serve_init_kernel<<<grid_for((int64_t)cap + 2, 256, 4, 2048), 256, 0, s>>>(d.t, A.P, d.stride);
    BOLT_HIP(hipGetLastError());
    return BOLT_OK;
  }

  // Extract, order, limit and copy out the result of a query that has seen
  // every chunk.  Returns 1 when the query must be re-run with larger tables.
  int finish(SActive& A, SDone& out, hipStream_t s) {
    const bolt_serve_query_t& q = A.r.q;
    uint32_t err[4];
    BOLT_HIP(hipMemcpyAsync(err, A.misc.ptr + 2, 16, hipMemcpyDeviceToHost, s));
    BOLT_HIP(hipStreamSynchronize(s));
    if (err[0]) return 1;
    const uint64_t total = A.r.cap + 2;
    std::vector<DevBuf<uint8_t>> cols(q.n_group + q.n_aggs);
    std::vector<DevBuf<uint32_t>> vcols(q.n_group + q.n_aggs);
    SOut o;
    memset(&o, 0, sizeof o);
    std::vector<int32_t> dts(q.n_group + q.n_aggs);
    for (int c = 0; c < q.n_group + q.n_aggs; c++) {
      int dt;
      if (c < q.n_group) dt = dim_attrs[q.join_dim[q.group_dim[c]]][q.group_attr[c]].dtype;
      else dt = A.out_dt[c - q.n_group];
      dts[c] = dt;
      BOLT_TRY(cols[c].alloc(total * dtype_size(dt), s));
      OutCol oc{cols[c].ptr, nullptr, dt, 0};
      if (c >= q.n_group) {
        const GAgg& g = A.P.a[c - q.n_group];
        if (g.nw >= 0 && g.kind != K_COUNT) {
          BOLT_TRY(vcols[c].alloc(total / 32 + 2, s));
          BOLT_TRY(fill_u32(vcols[c].ptr, 0, total / 32 + 2, s));
          oc.valid = vcols[c].ptr;
        }
        o.agg[c - q.n_group] = oc;
      } else {
        o.key[c] = oc;
      }
    }
    DevBuf<unsigned long long> counter;
    BOLT_TRY(counter.alloc(1, s));
    BOLT_TRY(fill_u32(counter.ptr, 0, 2, s));
    serve_extract_kernel<<<grid_for((int64_t)total, 256, 4, 2048), 256, 0, s>>>(A.desc.t, A.P, A.desc.stride,
                                                                                q.n_group, o, counter.ptr);
    BOLT_HIP(hipGetLastError());
    unsigned long long ng = 0;
    BOLT_HIP(hipMemcpyAsync(&ng, counter.ptr, 8, hipMemcpyDeviceToHost, s));
    BOLT_HIP(hipMemcpyAsync(err, A.misc.ptr + 2, 16, hipMemcpyDeviceToHost, s));
    BOLT_HIP(hipStreamSynchronize(s));
    if (err[2] & 1u) return set_error(BOLT_ERR_TYPE, "SUM(integer) overflow: accumulator exceeds i64 range");
    if (err[2] & 2u) return set_error(BOLT_ERR_TYPE, "SUM(Decimal128) precision overflow: accumulator exceeds i128 range");
    const int64_t n_groups = (int64_t)ng;
    const int64_t L = std::min<int64_t>(q.limit, n_groups);
    const int nc = q.n_group + q.n_aggs;
    out.cells.assign((size_t)L * nc, bolt_scalar_t{});
    out.res.n_groups = n_groups;
    out.res.n_rows = (int32_t)L;
    out.res.n_cols = nc;
    if (L > 0) {
      // ORDER BY aggs[order_agg] DESC NULLS LAST, then the group key ascending
      std::vector<bolt_column_t> sk;
      std::vector<int32_t> desc, nf;
      const int oc = q.n_group + q.order_agg;
      const uint8_t* ov = reinterpret_cast<const uint8_t*>(vcols[oc].ptr);
      sk.push_back(bolt_column_t{cols[oc].ptr, ov, n_groups, dts[oc], 0});
      desc.push_back(1);
      nf.push_back(0);
      for (int g = 0; g < q.n_group; g++) {
        sk.push_back(bolt_column_t{cols[g].ptr, nullptr, n_groups, dts[g], 0});
        desc.push_back(0);
        nf.push_back(0);
      }
      // the first L rows of the order only (a radix select above 65536 groups)
      DevBuf<uint32_t> perm;
      int64_t got = 0;
      BOLT_TRY(perm.alloc(L, s));
      BOLT_TRY(bolt_sort_limit((bolt_stream_t)s, sk.data(), (int32_t)sk.size(), desc.data(), nf.data(), 0, L,
                               perm.ptr, &got));
      if (got != L) return set_error(BOLT_ERR_OTHER, "serve: ORDER BY + LIMIT returned %lld of %lld rows",
                                     (long long)got, (long long)L);
      // gather the first L rows of every column, then one copy each
      size_t row_bytes = 0;
      for (int c = 0; c < nc; c++) row_bytes += dtype_size(dts[c]);
      DevBuf<uint8_t> small;
      BOLT_TRY(small.alloc((size_t)L * row_bytes + 64, s));
      std::vector<size_t> offs(nc);
      size_t off = 0;
      for (int c = 0; c < nc; c++) {
        offs[c] = off;
        serve_gather_kernel<<<grid_for(L, 256, 1), 256, 0, s>>>(cols[c].ptr, dtype_size(dts[c]), perm.ptr, L,
                                                                small.ptr + off);
        off += (size_t)L * dtype_size(dts[c]);
      }
      BOLT_HIP(hipGetLastError());
      std::vector<uint8_t> h(off);
      std::vector<uint32_t> hperm(L);
      BOLT_HIP(hipMemcpyAsync(h.data(), small.ptr, off, hipMemcpyDeviceToHost, s));
      BOLT_HIP(hipMemcpyAsync(hperm.data(), perm.ptr, L * 4, hipMemcpyDeviceToHost, s));
      std::vector<std::vector<uint32_t>> hv(nc);
      for (int c = 0; c < nc; c++)
        if (vcols[c].ptr) {
          hv[c].resize(total / 32 + 2);
          BOLT_HIP(hipMemcpyAsync(hv[c].data(), vcols[c].ptr, hv[c].size() * 4, hipMemcpyDeviceToHost, s));
        }
      BOLT_HIP(hipStreamSynchronize(s));
      for (int64_t i = 0; i < L; i++)
        for (int c = 0; c < nc; c++) {
          bolt_scalar_t& v = out.cells[(size_t)i * nc + c];
          const uint8_t* p = h.data() + offs[c] + (size_t)i * dtype_size(dts[c]);
          v.dtype = dts[c];
          const uint32_t pr = hperm[i];
          v.is_null = (!hv[c].empty() && !((hv[c][pr >> 5] >> (pr & 31)) & 1u)) ? 1 : 0;
          switch (dts[c]) {
            case BOLT_I32: case BOLT_DATE32: v.i64 = *reinterpret_cast<const int32_t*>(p); break;
            case BOLT_I64: v.i64 = *reinterpret_cast<const int64_t*>(p); break;
            case BOLT_F32: v.f64 = *reinterpret_cast<const float*>(p); break;
            case BOLT_F64: v.f64 = *reinterpret_cast<const double*>(p); break;
            case BOLT_DEC128:
              v.i64 = *reinterpret_cast<const int64_t*>(p);
              v.hi = *reinterpret_cast<const int64_t*>(p + 8);
              break;
          }
        }
    }
    out.res.status = BOLT_OK;
    return BOLT_OK;
  }

  void complete(SActive& A, int st, SDone&& d) {
    std::lock_guard<std::mutex> lk(mu);
    SDone& slot = done[A.r.ticket];
    slot = std::move(d);
    slot.ready = true;
    slot.res.submit_ms = A.r.submit_ms;
    slot.res.admit_ms = A.admit_ms;
    slot.res.done_ms = now_ms();
    slot.res.hbm_bytes = A.r.need;
    if (st != BOLT_OK) {
      slot.res.status = st;
      slot.err = bolt_last_error();
      slot.res.n_rows = 0;
      slot.cells.clear();
    }
    slot.res.cells = slot.cells.empty() ? nullptr : slot.cells.data();
    slot.res.error = slot.err.c_str();
    completed++;
    cv_done.notify_all();
  }

  void fail_request(const SReq& r, int st, const std::string& msg) {
    std::lock_guard<std::mutex> lk(mu);
    SDone& slot = done[r.ticket];
    slot.ready = true;
    slot.err = msg;
    slot.res.status = st;
    slot.res.submit_ms = r.submit_ms;
    slot.res.admit_ms = slot.res.done_ms = now_ms();
    slot.res.error = slot.err.c_str();
    cv_done.notify_all();
  }

  void loop() {
    (void)hipSetDevice(dev);
    for (;;) {
      std::vector<SReq> admitted;
      {
        std::unique_lock<std::mutex> lk(mu);
        if (active.empty() && queue.empty()) {
          if (stop && fin_pending == 0) return;
          cv_work.wait(lk, [&] { return !queue.empty() || (stop && fin_pending == 0); });
          if (queue.empty()) continue;
        }
        // FIFO admission at this chunk boundary: a slot and the HBM budget
        int64_t bytes = bytes_in_use;
        size_t n_act = active.size();
        while (!queue.empty() && (int)(n_act + admitted.size()) < cfg.slots) {
          const SReq& r = queue.front();
          // queries still being finished hold their memory too (bytes_in_use
          // counts them); submit already declined any query over the budget
          if (bytes + r.need > cfg.hbm_budget) break;
          bytes += r.need;
          admitted.push_back(r);
          queue.pop_front();
        }
      }
      for (SReq& r : admitted) {
        auto A = std::make_unique<SActive>();
        A->r = r;
        A->admit_ms = now_ms();
        clear_error();
        int st = admit(*A);
        if (st != BOLT_OK) {
          SDone d;
          complete(*A, st, std::move(d));
          continue;
        }
        {
          std::lock_guard<std::mutex> lk(mu);
          bytes_in_use += r.need;
        }
        active.push_back(std::move(A));
        desc_dirty = true;
      }
      if ((int32_t)active.size() > peak_active.load()) peak_active = (int32_t)active.size();
      if (bytes_in_use > peak_bytes.load()) peak_bytes = bytes_in_use;
      if (active.empty()) continue;
      if (n_chunks > 0) {
        if (desc_dirty) {
          // union of the predicate columns of the active queries
          memset(&pre, 0, sizeof pre);
          int i = 0;
          for (auto& a : active) {
            SQ& d = a->desc;
            for (int t = 0; t < BOLT_MAX_PRED_TERMS; t++) d.term_slot[t] = -1;
            for (int t = 0; t < d.f.n; t++) {
              const DevTerm& tm = d.f.t[t];
              int slot = -1;
              for (int c = 0; c < pre.n; c++)
                if (pre.d[c] == tm.data) slot = c;
              if (slot < 0 && pre.n < kPreCols) {
                slot = pre.n++;
                pre.d[slot] = tm.data;
                pre.v[slot] = tm.validity;
                pre.dt[slot] = tm.dtype;
              }
              d.term_slot[t] = (int8_t)slot;
            }
            compile_intervals(d);
            h_desc[i++] = d;
          }
          (void)hipMemcpyAsync(d_desc.ptr, h_desc, sizeof(SQ) * active.size(), hipMemcpyHostToDevice, s);
          desc_dirty = false;
        }
        const int64_t r0 = cur_chunk * cfg.chunk_rows;
        const int64_t r1 = std::min(n_rows, r0 + cfg.chunk_rows);
        const auto t_step = SClock::now();
        const int64_t per_w = serve_per_wave(cfg.chunk_rows);
        serve_filter_kernel<<<kServeGrid, kServeBlock, 0, s>>>(d_desc.ptr, (int)active.size(), pre, r0, r1, per_w,
                                                               d_lists.ptr, d_counts.ptr);
        serve_probe_kernel<<<kServeGrid, kServeBlock, 0, s>>>(d_desc.ptr, (int)active.size(), sdims, r0, per_w,
                                                              d_lists.ptr, d_counts.ptr);
        (void)hipStreamSynchronize(s);
        step_us += std::chrono::duration_cast<std::chrono::microseconds>(SClock::now() - t_step).count();
        cur_chunk = (cur_chunk + 1) % n_chunks;
        steps++;
      }
      for (auto it = active.begin(); it != active.end();) {
        SActive& A = **it;
        A.chunks_done++;
        if (A.chunks_done < n_chunks) {
          ++it;
          continue;
        }
        {
          std::lock_guard<std::mutex> lk(mu);
          fin_q.push_back(std::move(*it));
          fin_pending++;
        }
        cv_fin.notify_one();
        it = active.erase(it);
        desc_dirty = true;
      }
    }
  }

  void finisher() {
    (void)hipSetDevice(dev);
    for (;;) {
      std::unique_ptr<SActive> Ap;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_fin.wait(lk, [&] { return !fin_q.empty() || fin_stop; });
        if (fin_q.empty()) return;
        Ap = std::move(fin_q.front());
        fin_q.pop_front();
      }
      SActive& A = *Ap;
      const auto t0f = SClock::now();
      clear_error();
      SDone d;
      int st = finish(A, d, fs);
      bool retry = false;
      SReq r = A.r;
      const int64_t admitted = A.r.need;  // what admission added to bytes_in_use
      if (st == 1 && A.r.attempt < 3) {
        // a group table filled up: run the query again with 4x the table,
        // unless the larger table no longer fits the whole budget (then it
        // could never be admitted and would block the queue head)
        r.attempt++;
        r.min_cap = A.r.cap * 4;
        const int pst = plan_request(r);
        if (pst != BOLT_OK) st = pst;
        else if (r.need > cfg.hbm_budget)
          st = set_error(BOLT_DECLINE, "serve: the re-planned group table (%lld bytes) exceeds the HBM budget",
                         (long long)r.need);
        else retry = true;
      } else if (st == 1) {
        st = set_error(BOLT_ERR_MEMORY, "serve: group table overflow (est_groups far too low)");
      }
      if (!retry) complete(A, st, std::move(d));
      Ap.reset();  // frees the query's tables (stream-ordered on the scan stream)
      finish_us += std::chrono::duration_cast<std::chrono::microseconds>(SClock::now() - t0f).count();
      {
        std::lock_guard<std::mutex> lk(mu);
        bytes_in_use -= admitted;  // the retry is charged again when it is re-admitted
        if (retry) queue.push_front(r);
        fin_pending--;
      }
      cv_work.notify_all();
    }
  }
};

extern "C" int bolt_server_create(bolt_stream_t stream, const bolt_column_t* fact, int32_t n_fact,
                                  const bolt_serve_dim_t* dims, int32_t n_dims, const bolt_serve_config_t* cfg,
                                  bolt_server_t** out) {
  BOLT_CHECK_DEVICE();
  clear_error();
  (void)stream;
  if (!out) return set_error(BOLT_ERR_PLAN, "bolt_server_create: out is NULL");
  *out = nullptr;
  if (n_fact < 1 || !fact) return set_error(BOLT_ERR_PLAN, "bolt_server_create: no fact columns");
  if (n_dims < 1 || n_dims > BOLT_SERVE_MAX_DIMS || !dims)
    return set_error(BOLT_ERR_PLAN, "bolt_server_create: 1..3 dimensions required");
  auto srv = std::make_unique<bolt_server>();
  srv->t0 = SClock::now();
  BOLT_HIP(hipGetDevice(&srv->dev));
  BOLT_HIP(hipStreamCreateWithFlags(&srv->s, hipStreamNonBlocking));
  srv->fact.assign(fact, fact + n_fact);
  srv->n_rows = fact[0].length;
  for (int c = 0; c < n_fact; c++)
    if (fact[c].length != srv->n_rows) return set_error(BOLT_ERR_PLAN, "bolt_server_create: fact columns differ in length");
  srv->cfg.slots = cfg && cfg->slots > 0 ? cfg->slots : 8;
  srv->cfg.hbm_budget = cfg && cfg->hbm_budget > 0 ? cfg->hbm_budget : (int64_t)4 << 30;
  srv->cfg.chunk_rows = cfg && cfg->chunk_rows > 0 ? cfg->chunk_rows : (int64_t)8 << 20;
  if (srv->cfg.slots > BOLT_SERVE_MAX_SLOTS)
    return set_error(BOLT_ERR_PLAN, "bolt_server_create: at most %d slots", BOLT_SERVE_MAX_SLOTS);
  // a chunk never exceeds the table (its row lists are allocated per slot) nor 2^31 rows
  srv->cfg.chunk_rows = std::min<int64_t>(srv->cfg.chunk_rows, std::max<int64_t>(srv->n_rows, 1));
  srv->cfg.chunk_rows = std::min<int64_t>(srv->cfg.chunk_rows, (int64_t)1 << 31);
  srv->n_chunks = (srv->n_rows + srv->cfg.chunk_rows - 1) / srv->cfg.chunk_rows;
  memset(&srv->sdims, 0, sizeof srv->sdims);
  auto cleanup = [&]() {
    for (auto* t : srv->dim_tables) bolt_join_table_release(t);
    srv->dim_tables.clear();
    srv->d_desc.reset();
    srv->d_lists.reset();
    srv->d_counts.reset();
    (void)hipStreamSynchronize(srv->s);
    cache_forget_stream(srv->s);
    (void)hipStreamDestroy(srv->s);
  };
  for (int d = 0; d < n_dims; d++) {
    if (dims[d].n_attrs < 1 || dims[d].n_attrs > kServeMaxAttrs || !dims[d].attrs) {
      cleanup();
      return set_error(BOLT_ERR_PLAN, "dimension %d: 1..8 attribute columns required", d);
    }
    if (dims[d].key.validity) {
      cleanup();
      return set_error(BOLT_ERR_UNSUPPORTED, "dimension %d: nullable keys", d);
    }
    bolt_join_table_t* t = nullptr;
    int st = bolt_join_build((bolt_stream_t)srv->s, &dims[d].key, nullptr, &t);
    if (st != BOLT_OK) {
      cleanup();
      return st;
    }
    srv->dim_tables.push_back(t);
    int32_t uniq = 0;
    st = bolt_join_table_unique(t, &uniq);
    if (st != BOLT_OK || !uniq) {
      cleanup();
      return st != BOLT_OK ? st : set_error(BOLT_ERR_UNSUPPORTED, "dimension %d: keys are not unique", d);
    }
    srv->dim_attrs.emplace_back(dims[d].attrs, dims[d].attrs + dims[d].n_attrs);
    SDim& sd = srv->sdims.d[d];
    sd.tbl = t->tbl;
    sd.mask = t->mask;
    for (int a = 0; a < dims[d].n_attrs; a++) {
      const bolt_column_t& c = dims[d].attrs[a];
      if (c.length != dims[d].key.length || c.validity ||
          !(c.dtype == BOLT_I32 || c.dtype == BOLT_I64 || c.dtype == BOLT_DATE32)) {
        cleanup();
        return set_error(BOLT_ERR_UNSUPPORTED, "dimension %d attribute %d: non-null Int32 / Int64 / Date32 required", d, a);
      }
      sd.attr[a] = c.data;
      sd.attr_dt[a] = c.dtype;
    }
  }
  int st = srv->d_desc.alloc(srv->cfg.slots, srv->s);
  {
    const int64_t per_w = serve_per_wave(srv->cfg.chunk_rows);
    if (st == BOLT_OK) st = srv->d_lists.alloc((size_t)srv->cfg.slots * kServeAllWaves * (size_t)per_w, srv->s);
    if (st == BOLT_OK) st = srv->d_counts.alloc((size_t)srv->cfg.slots * kServeAllWaves, srv->s);
  }
  if (st == BOLT_OK && hipHostMalloc(reinterpret_cast<void**>(&srv->h_desc), sizeof(SQ) * srv->cfg.slots,
                                     hipHostMallocDefault) != hipSuccess)
    st = set_error(BOLT_ERR_MEMORY, "bolt_server_create: pinned descriptor buffer");
  if (st == BOLT_OK && hipStreamSynchronize(srv->s) != hipSuccess)
    st = set_error(BOLT_ERR_DEVICE, "bolt_server_create: dimension build failed");
  if (st != BOLT_OK) {
    cleanup();
    return st;
  }
  if (hipStreamCreateWithFlags(&srv->fs, hipStreamNonBlocking) != hipSuccess) {
    cleanup();
    return set_error(BOLT_ERR_DEVICE, "bolt_server_create: finisher stream");
  }
  bolt_server* p = srv.release();
  p->th = std::thread([p] { p->loop(); });
  p->fin_th = std::thread([p] { p->finisher(); });
  *out = p;
  return BOLT_OK;
}

extern "C" int bolt_server_submit(bolt_server_t* srv, const bolt_serve_query_t* q, int64_t* ticket) {
  clear_error();
  if (!srv || !q || !ticket) return set_error(BOLT_ERR_PLAN, "bolt_server_submit: NULL argument");
  SReq r;
  r.q = *q;
  r.attempt = 0;
  r.min_cap = 0;
  r.submit_ms = srv->now_ms();
  int st = srv->plan_request(r);
  std::lock_guard<std::mutex> lk(srv->mu);
  r.ticket = srv->next_ticket++;
  *ticket = r.ticket;
  if (st != BOLT_OK) {
    SDone& d = srv->done[r.ticket];
    d.ready = true;
    d.err = bolt_last_error();
    d.res.status = st;
    d.res.submit_ms = d.res.admit_ms = d.res.done_ms = r.submit_ms;
    d.res.error = d.err.c_str();
    return BOLT_OK;
  }
  if (r.need > srv->cfg.hbm_budget) {
    SDone& d = srv->done[r.ticket];
    d.ready = true;
    d.err = "serve: the query's hash tables exceed the whole HBM budget";
    d.res.status = BOLT_DECLINE;
    d.res.submit_ms = d.res.admit_ms = d.res.done_ms = r.submit_ms;
    d.res.error = d.err.c_str();
    return BOLT_OK;
  }
  srv->queue.push_back(r);
  srv->cv_work.notify_one();
  return BOLT_OK;
}

extern "C" int bolt_server_wait(bolt_server_t* srv, int64_t ticket, double timeout_ms, bolt_serve_result_t* res) {
  if (!srv || !res) return set_error(BOLT_ERR_PLAN, "bolt_server_wait: NULL argument");
  std::unique_lock<std::mutex> lk(srv->mu);
  auto ready = [&] {
    auto it = srv->done.find(ticket);
    return it != srv->done.end() && it->second.ready;
  };
  if (timeout_ms < 0) srv->cv_done.wait(lk, ready);
  else if (!srv->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready))
    return BOLT_HIP_ERROR_NOT_READY;
  SDone& d = srv->done[ticket];
  *res = d.res;
  res->cells = d.cells.empty() ? nullptr : d.cells.data();
  res->error = d.err.c_str();
  return BOLT_OK;
}

extern "C" void bolt_server_result_release(bolt_server_t* srv, int64_t ticket) {
  if (!srv) return;
  std::lock_guard<std::mutex> lk(srv->mu);
  srv->done.erase(ticket);
}

extern "C" double bolt_server_now_ms(const bolt_server_t* srv) { return srv ? srv->now_ms() : 0.0; }

extern "C" int bolt_server_stats(const bolt_server_t* srv, int64_t* steps, int64_t* completed, int32_t* peak_active,
                                 int64_t* peak_bytes, double* step_ms, double* finish_ms) {
  if (!srv) return set_error(BOLT_ERR_PLAN, "bolt_server_stats: server is NULL");
  bolt_server* s = const_cast<bolt_server*>(srv);
  std::lock_guard<std::mutex> lk(s->mu);
  if (steps) *steps = s->steps.load();
  if (completed) *completed = s->completed;
  if (peak_active) *peak_active = s->peak_active.load();
  if (peak_bytes) *peak_bytes = s->peak_bytes.load();
  if (step_ms) *step_ms = s->step_us.load() / 1000.0;
  if (finish_ms) *finish_ms = s->finish_us.load() / 1000.0;
  return BOLT_OK;
}

extern "C" void bolt_server_destroy(bolt_server_t* srv) {
  if (!srv) return;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stop = true;
  }
  srv->cv_work.notify_all();
  if (srv->th.joinable()) srv->th.join();  // returns once the queue, the scan and the finisher are drained
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->fin_stop = true;
  }
  srv->cv_fin.notify_all();
  if (srv->fin_th.joinable()) srv->fin_th.join();
  srv->active.clear();
  srv->d_desc.reset();
  srv->d_lists.reset();
  srv->d_counts.reset();
  for (auto* t : srv->dim_tables) bolt_join_table_release(t);
  (void)hipStreamSynchronize(srv->s);
  cache_forget_stream(srv->s);
  (void)hipStreamDestroy(srv->s);
  if (srv->fs) {
    cache_forget_stream(srv->fs);
    (void)hipStreamDestroy(srv->fs);
  }
  delete srv;
}
