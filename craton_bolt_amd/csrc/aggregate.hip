// aggregate.hip — scalar aggregates (no GROUP BY) and the host helpers shared
// by every operator that evaluates predicates or aggregates.
//
// Replaces execute_aggregate (src/exec/aggregate.rs:154) / reduce_gpu_vec
// (:1596) / the bolt_reduce kernels (src/jit/agg_kernels.rs:242,818) and
// agg_with_pre::try_execute_resident (src/exec/agg_with_pre.rs:157).  The
// reference launches one kernel per aggregate and finishes the per-block
// partials on the host; here one pass evaluates the filter and up to four
// aggregates per launch, and a second one-block kernel folds the partials in
// a fixed order on the device.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <vector>

#include "jit.h"
#include "reduce_common.h"

namespace bolt {

// ---- host helpers -----------------------------------------------------------

int make_dev_filter(const bolt_filter_t* f, const bolt_column_t* cols, int32_t n_cols, int64_t n_rows,
                    DevFilter* out) {
  memset(out, 0, sizeof *out);
  if (f == nullptr || f->n_terms == 0) return BOLT_OK;
  if (f->n_terms < 0 || f->n_terms > BOLT_MAX_PRED_TERMS)
    return set_error(BOLT_ERR_PLAN, "filter: %d terms (max %d)", f->n_terms, BOLT_MAX_PRED_TERMS);
  if (f->terms == nullptr) return set_error(BOLT_ERR_PLAN, "filter: terms is NULL");
  for (int i = 0; i < f->n_terms; i++) {
    const bolt_pred_term_t& t = f->terms[i];
    if (t.col < 0 || t.col >= n_cols)
      return set_error(BOLT_ERR_PLAN, "filter term %d: column %d out of range (%d columns)", i, t.col, n_cols);
    const bolt_column_t& c = cols[t.col];
    if (c.length != n_rows)
      return set_error(BOLT_ERR_PLAN, "filter term %d: column length %lld != %lld rows", i,
                       (long long)c.length, (long long)n_rows);
    if (t.op < BOLT_CMP_LT || t.op > BOLT_CMP_IS_NOT_NULL)
      return set_error(BOLT_ERR_PLAN, "filter term %d: unknown comparison %d", i, t.op);
    if (dtype_size(c.dtype) == 0) return set_error(BOLT_ERR_TYPE, "filter term %d: unknown dtype %d", i, c.dtype);
    if (c.dtype == BOLT_TIMESTAMP)
      return set_error(BOLT_ERR_TYPE, "filter term %d: a Timestamp is compared inside an expression program", i);
    DevTerm& d = out->t[i];
    d.data = c.data;
    d.validity = c.validity;
    d.dtype = c.dtype;
    d.op = t.op;
    d.lit_i64 = t.lit_i64;
    d.lit_f64 = t.lit_f64;
    d.lit_hi = t.lit_hi;
  }
  out->n = f->n_terms;
  return BOLT_OK;
}

int make_agg_list(const bolt_agg_t* aggs, int32_t n_aggs, const bolt_column_t* cols, int32_t n_cols,
                  int32_t carry_dtype, AggList* out) {
  memset(out, 0, sizeof *out);
  if (n_aggs < 0 || n_aggs > kMaxAggs)
    return set_error(BOLT_ERR_PLAN, "%d aggregates (max %d)", n_aggs, kMaxAggs);
  for (int i = 0; i < n_aggs; i++) {
    int op = aggs[i].op;
    if (op < BOLT_AGG_SUM || op > BOLT_AGG_STDDEV_SAMP)
      return set_error(BOLT_ERR_PLAN, "aggregate %d: unknown op %d", i, op);
    AggDesc& d = out->a[i];
    d.op = op;
    if (op == BOLT_AGG_COUNT_STAR) {
      d.dtype = BOLT_I64;
      continue;
    }
    int dt;
    if (aggs[i].col == BOLT_JOIN_CARRY) {
      if (carry_dtype == 0) return set_error(BOLT_ERR_PLAN, "aggregate %d reads the join carry, but the table has none", i);
      dt = carry_dtype;
      d.src = 1;
    } else {
      if (aggs[i].col < 0 || aggs[i].col >= n_cols)
        return set_error(BOLT_ERR_PLAN, "aggregate %d: input column %d out of range (%d columns)", i, aggs[i].col, n_cols);
      const bolt_column_t& c = cols[aggs[i].col];
      dt = c.dtype;
      d.data = c.data;
      d.validity = c.validity;
    }
    d.dtype = dt;
    if (op == BOLT_AGG_COUNT) continue;
    if (dt == BOLT_BOOL) return set_error(BOLT_ERR_TYPE, "aggregate %d: Bool inputs are not supported", i);
    if (dt == BOLT_TIMESTAMP)
      return set_error(BOLT_ERR_TYPE, "aggregate %d: a Timestamp enters through an expression program (EXTRACT, "
                       "DATE_TRUNC, comparisons)", i);
    if (op == BOLT_AGG_SUM && dt == BOLT_DATE32)
      // src/plan/logical_plan.rs:2135-2146
      return set_error(BOLT_ERR_TYPE, "SUM over a temporal type is undefined in SQL (Date32)");
    if ((op == BOLT_AGG_AVG || op >= BOLT_AGG_VAR_POP) && dt == BOLT_DATE32)
      return set_error(BOLT_ERR_TYPE, "AVG/VAR/STDDEV over Date32 is not supported");
    if (dt == BOLT_DEC128 && op != BOLT_AGG_SUM)
      return set_error(BOLT_ERR_UNSUPPORTED, "aggregate %d: only SUM is supported over Decimal128", i);
  }
  out->n = n_aggs;
  return BOLT_OK;
}

int acc_finalize(const Acc& a, int op, int dt, bolt_scalar_t* r) {
  memset(r, 0, sizeof *r);
  const uint64_t n = acc_count(a);
  const bool ovf = (a.n & kOvfBit) != 0;
  switch (op) {
    case BOLT_AGG_COUNT:
    case BOLT_AGG_COUNT_STAR:
      r->dtype = BOLT_I64;
      r->i64 = (int64_t)n;
      return BOLT_OK;
    case BOLT_AGG_SUM:
      r->dtype = dt == BOLT_I32 ? BOLT_I64 : dt;
      if (n == 0) { r->is_null = 1; return BOLT_OK; }
      if (dt == BOLT_DEC128) {
        if (ovf) return set_error(BOLT_ERR_TYPE, "SUM(Decimal128) precision overflow: accumulator exceeds i128 range");
        r->i64 = (int64_t)a.a;
        r->hi = (int64_t)a.b;
      } else if (dtype_is_int(dt)) {
        // the exact 128-bit total must fit i64 (src/exec/aggregate.rs:2120)
        if (ovf || (int64_t)a.b != ((int64_t)a.a >> 63))
          return set_error(BOLT_ERR_TYPE, "SUM(integer) overflow: accumulator exceeds i64 range");
        r->i64 = (int64_t)a.a;
      } else {
        // Neumaier result; non-finite terms fall back to the plain fold, which
        // is the running leading part (src/exec/aggregate.rs:2022)
        double s = as_f64(a.a), c = as_f64(a.b);
        double v = (std::isfinite(s) && std::isfinite(c)) ? s + c : s;
        r->f64 = dt == BOLT_F32 ? (double)(float)v : v;
      }
      return BOLT_OK;
    case BOLT_AGG_AVG: {
      r->dtype = BOLT_F64;
      if (n == 0) { r->is_null = 1; return BOLT_OK; }
      double s = as_f64(a.a), c = as_f64(a.b);
      r->f64 = ((std::isfinite(s) && std::isfinite(c)) ? s + c : s) / (double)n;
      return BOLT_OK;
    }
    case BOLT_AGG_MIN:
    case BOLT_AGG_MAX:
      r->dtype = dt;
      if (n == 0) { r->is_null = 1; return BOLT_OK; }
      if (dtype_is_float(dt)) r->f64 = f64_from_order_key(a.a);
      else r->i64 = (int64_t)(a.a ^ 0x8000000000000000ull);
      return BOLT_OK;
    default: {
      // Welford finalize (src/exec/welford.rs:104-134)
      r->dtype = BOLT_F64;
      bool samp = (op == BOLT_AGG_VAR_SAMP || op == BOLT_AGG_STDDEV_SAMP);
      if (n == 0 || (samp && n <= 1)) { r->is_null = 1; return BOLT_OK; }
      double var = as_f64(a.b) / (double)(samp ? n - 1 : n);
      r->f64 = (op == BOLT_AGG_STDDEV_POP || op == BOLT_AGG_STDDEV_SAMP) ? std::sqrt(var) : var;
      return BOLT_OK;
    }
  }
}

int launch_finish(int na, const Acc* parts, int n_parts, int stride, int first, const AggList& aggs, Acc* out,
                  hipStream_t s) {
  switch (na) {
    case 1: finish_partials_kernel<1><<<1, kReduceBlock, 0, s>>>(parts, n_parts, stride, first, aggs, out); break;
    case 2: finish_partials_kernel<2><<<1, kReduceBlock, 0, s>>>(parts, n_parts, stride, first, aggs, out); break;
    case 3: finish_partials_kernel<3><<<1, kReduceBlock, 0, s>>>(parts, n_parts, stride, first, aggs, out); break;
    case 4: finish_partials_kernel<4><<<1, kReduceBlock, 0, s>>>(parts, n_parts, stride, first, aggs, out); break;
    default: return set_error(BOLT_ERR_OTHER, "launch_finish: bad chunk size %d", na);
  }
  BOLT_HIP(hipGetLastError());
  return BOLT_OK;
}

// ---- scalar aggregate kernel -------------------------------------------------

// Grid-stride pass, four rows per thread per iteration with every load
// issued before any use (indices are clamped to the last row so the loads
// need no branch); each thread folds its rows into NA register accumulators
// and each block writes one partial.  The row-to-thread mapping is fixed by
// the grid, so float results are reproducible.
constexpr int kAggUnroll = 4;

template <int NT, int NA>
__global__ __launch_bounds__(kReduceBlock) void scalar_agg_kernel(DevFilter f, AggList aggs, int64_t n,
                                                                 Acc* __restrict__ parts) {
  Acc acc[NA];
#pragma unroll
  for (int a = 0; a < NA; a++) acc_init(acc[a], aggs.a[a].op);
  const int64_t step = (int64_t)gridDim.x * kReduceBlock * kAggUnroll;
  for (int64_t base = (int64_t)blockIdx.x * kReduceBlock * kAggUnroll + threadIdx.x; base < n; base += step) {
    bool keep[kAggUnroll];
    Raw v[kAggUnroll][NA];
#pragma unroll
    for (int u = 0; u < kAggUnroll; u++) {
      int64_t i = base + (int64_t)u * kReduceBlock;
      int64_t ic = i < n ? i : n - 1;
      keep[u] = i < n && eval_filter_nt<NT>(f, ic);
#pragma unroll
      for (int a = 0; a < NA; a++)
        if (aggs.a[a].op != BOLT_AGG_COUNT_STAR) v[u][a] = load_raw(aggs.a[a].data, aggs.a[a].dtype, ic);
    }
#pragma unroll
    for (int u = 0; u < kAggUnroll; u++) {
      if (!keep[u]) continue;
      int64_t i = base + (int64_t)u * kReduceBlock;
#pragma unroll
      for (int a = 0; a < NA; a++) {
        const AggDesc& d = aggs.a[a];
        if (d.op == BOLT_AGG_COUNT_STAR) { acc[a].n++; continue; }
        if (!bit_valid(d.validity, i)) continue;
        int64_t vi, vhi;
        double vf;
        raw_value(v[u][a], d.dtype, vi, vhi, vf);
        acc_update(acc[a], d.op, d.dtype, vi, vhi, vf);
      }
    }
  }
  block_reduce_store<NA>(acc, aggs, parts + (int64_t)blockIdx.x * NA);
}

template <int NT>
static int launch_scalar_nt(int na, const DevFilter& f, const AggList& chunk, int64_t n, unsigned grid, Acc* parts,
                            hipStream_t s) {
  switch (na) {
    case 1: scalar_agg_kernel<NT, 1><<<grid, kReduceBlock, 0, s>>>(f, chunk, n, parts); break;
    case 2: scalar_agg_kernel<NT, 2><<<grid, kReduceBlock, 0, s>>>(f, chunk, n, parts); break;
    case 3: scalar_agg_kernel<NT, 3><<<grid, kReduceBlock, 0, s>>>(f, chunk, n, parts); break;
    case 4: scalar_agg_kernel<NT, 4><<<grid, kReduceBlock, 0, s>>>(f, chunk, n, parts); break;
    default: return set_error(BOLT_ERR_OTHER, "scalar aggregate: bad chunk size %d", na);
  }
  BOLT_HIP(hipGetLastError());
  return BOLT_OK;
}

static int launch_scalar_chunk(int na, const DevFilter& f, const AggList& chunk, int64_t n, unsigned grid,
                               Acc* parts, hipStream_t s) {
  switch (f.n) {
    case 0: return launch_scalar_nt<0>(na, f, chunk, n, grid, parts, s);
    case 1: return launch_scalar_nt<1>(na, f, chunk, n, grid, parts, s);
    case 2: return launch_scalar_nt<2>(na, f, chunk, n, grid, parts, s);
    default: return launch_scalar_nt<-1>(na, f, chunk, n, grid, parts, s);
  }
}

// ---- row-order overflow of SUM(Int64) ----------------------------------------
//
// The reference folds an integer SUM with checked_add in row order and fails
// as soon as a running total leaves i64 (src/exec/aggregate.rs:2120-2140), so
// [i64::MAX, 1, -1] is an error although its total fits.  The fold's state
// over a run of rows is (sum, highest running total, lowest running total),
// and runs combine in order:
//   (s1, hi1, lo1) + (s2, hi2, lo2) = (s1 + s2, max(hi1, s1 + hi2), min(lo1, s1 + lo2)).
// Each workgroup folds one contiguous range: 256 rows per step, combined in
// lane order by a shuffle tree, then across the four waves in order; the host
// folds the workgroups' triples in order.  Sums stay below 2^31 * 2^63, so
// 128-bit arithmetic is exact, and a run with no row has hi = -kPfxNone,
// lo = +kPfxNone, far outside any real total.
struct Pfx {
  __int128 s, hi, lo;
};
constexpr __int128 kPfxNone = (__int128)1 << 120;

__host__ __device__ __forceinline__ Pfx pfx_cat(const Pfx& a, const Pfx& b) {
  Pfx r;
  r.s = a.s + b.s;
  const __int128 h = a.s + b.hi, l = a.s + b.lo;
  r.hi = a.hi > h ? a.hi : h;
  r.lo = a.lo < l ? a.lo : l;
  return r;
}

__device__ __forceinline__ __int128 shfl_down_i128(__int128 v, int off) {
  const uint64_t lo = (uint64_t)v, hi = (uint64_t)(v >> 64);
  const uint64_t l2 = __shfl_down(lo, off), h2 = __shfl_down(hi, off);
  return (__int128)(((unsigned __int128)h2 << 64) | l2);
}

__global__ __launch_bounds__(kReduceBlock) void sum_prefix_kernel(DevFilter f, const int64_t* __restrict__ x,
                                                                 const uint8_t* __restrict__ valid, int64_t n,
                                                                 int64_t per, Pfx* __restrict__ out) {
  __shared__ Pfx sh[kWaves];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  Pfx run{0, -kPfxNone, kPfxNone};
  const int64_t r0 = (int64_t)blockIdx.x * per;
  const int64_t r1 = r0 + per < n ? r0 + per : n;
  for (int64_t base = r0; base < r1; base += kReduceBlock) {
    const int64_t i = base + threadIdx.x;
    Pfx m{0, -kPfxNone, kPfxNone};
    if (i < r1 && bit_valid(valid, i) && eval_filter_nt<-1>(f, i)) {
      const __int128 v = x[i];
      m = Pfx{v, v, v};
    }
    // in-order tree over the wave: lane L ends holding lanes [L, L + 2 * off)
    for (int off = 1; off < 64; off <<= 1) {
      Pfx o;
      o.s = shfl_down_i128(m.s, off);
      o.hi = shfl_down_i128(m.hi, off);
      o.lo = shfl_down_i128(m.lo, off);
      if ((lane & (2 * off - 1)) == 0 && lane + off < 64) m = pfx_cat(m, o);
    }
    if (lane == 0) sh[wave] = m;
    __syncthreads();
    if (threadIdx.x == 0)
      for (int w = 0; w < kWaves; w++) run = pfx_cat(run, sh[w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = run;
}

// BOLT_OK when no running total of the filtered, non-NULL rows of `col`
// (Int64), in row order, leaves i64; the reference's overflow error otherwise.
constexpr unsigned kPfxMaxGrid = 1024;

// `scratch` (room for kPfxMaxGrid triples) or NULL to allocate.
static int check_sum_prefix(hipStream_t s, const bolt_column_t& col, const DevFilter& f, Pfx* scratch) {
  const int64_t n = col.length;
  const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(kPfxMaxGrid, (n + 65535) / 65536));
  const int64_t per = (n + grid - 1) / grid;
  DevBuf<Pfx> parts;
  Pfx* pp = scratch;
  if (!pp) {
    BOLT_TRY(parts.alloc(grid, s));
    pp = parts.ptr;
  }
  sum_prefix_kernel<<<grid, kReduceBlock, 0, s>>>(f, static_cast<const int64_t*>(col.data), col.validity, n, per, pp);
  BOLT_HIP(hipGetLastError());
  Pfx h[kPfxMaxGrid];
  BOLT_HIP(hipMemcpyAsync(h, pp, sizeof(Pfx) * grid, hipMemcpyDeviceToHost, s));
  BOLT_HIP(hipStreamSynchronize(s));
  Pfx run{0, -kPfxNone, kPfxNone};
  for (unsigned b = 0; b < grid; b++) run = pfx_cat(run, h[b]);
  if (run.hi > (__int128)INT64_MAX || run.lo < (__int128)INT64_MIN)
    return set_error(BOLT_ERR_TYPE, "SUM(integer) overflow: accumulator exceeds i64 range");
  return BOLT_OK;
}

// Device buffers of one scalar-aggregate call: per-workgroup partials, the
// finished accumulators and the row-order prefix triples.
struct AggScratch {
  Acc* parts;
  Acc* outs;
  Pfx* pfx;
};

// bolt_aggregate, with its buffers allocated per call (scratch == NULL; the
// generated kernel when hipRTC is available) or taken from `scratch` (the
// precompiled kernels only, which allocate nothing).
static int aggregate_impl(hipStream_t s, const bolt_column_t* inputs, int32_t n_inputs, const bolt_agg_t* aggs,
                          int32_t n_aggs, const bolt_filter_t* filter, bolt_scalar_t* results,
                          const AggScratch* scratch);

}  // namespace bolt

using namespace bolt;

extern "C" int bolt_aggregate(bolt_stream_t stream, const bolt_column_t* inputs, int32_t n_inputs,
                              const bolt_agg_t* aggs, int32_t n_aggs, const bolt_filter_t* filter,
                              bolt_scalar_t* results) {
  BOLT_CHECK_DEVICE();
  clear_error();
  return aggregate_impl((hipStream_t)stream, inputs, n_inputs, aggs, n_aggs, filter, results, nullptr);
}

// ---- allocation-free variant (bolt_aggregate_noalloc) ------------------------
//
// For callers that cannot allocate device memory (HIP graph capture, serving
// loops, harnesses that forbid hipMalloc): the partials, the finished
// accumulators and the prefix triples live in one file-scope device array,
// `scratch`, of kScratchBytes.  Each device holds its own copy at its own
// address, looked up once per device with hipGetSymbolAddress; host code uses
// only that address.  Nothing in it needs to start at zero (every kernel
// writes what it later reads), and calls that overlap on one device share it
// and therefore race.
constexpr size_t kScratchBytes = (size_t)1 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// The current device's copy of `scratch` (looked up on the device's first call).
int scratch_for_current_device(void** out) {
  int dev = 0;
  BOLT_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxScratchDevices)
    return set_error(BOLT_ERR_DEVICE, "bolt_aggregate_noalloc: device %d beyond %d", dev, kMaxScratchDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      return set_error(BOLT_ERR_DEVICE, "bolt_aggregate_noalloc: address of the scratch array: %s",
                       hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return BOLT_OK;
}
}  // namespace

extern "C" int bolt_aggregate_noalloc(bolt_stream_t stream, const bolt_column_t* inputs, int32_t n_inputs,
                                      const bolt_agg_t* aggs, int32_t n_aggs, const bolt_filter_t* filter,
                                      bolt_scalar_t* results) {
  BOLT_CHECK_DEVICE();
  clear_error();
  if (n_aggs < 0 || n_aggs > kMaxAggs)
    return set_error(BOLT_ERR_PLAN, "%d aggregates (max %d)", n_aggs, kMaxAggs);
  const int64_t n = n_inputs >= 1 && inputs ? inputs[0].length : 0;
  // room for the most the call can use: kMaxAggs accumulators (the caller's
  // and the hidden MIN / MAX of SUM(Int64)) per workgroup, in chunks of 4
  const unsigned grid = grid_for(n, kReduceBlock, kAggUnroll * 4, 256 * 8);
  const size_t parts_b = align256((size_t)grid * kMaxAggs * sizeof(Acc));
  const size_t outs_b = align256((size_t)kMaxAggs * sizeof(Acc));
  const size_t pfx_b = (size_t)kPfxMaxGrid * sizeof(Pfx);
  if (parts_b + outs_b + pfx_b > kScratchBytes)
    return set_error(BOLT_ERR_MEMORY, "bolt_aggregate_noalloc: needs %zu bytes of scratch, over its maximum of %zu",
                     parts_b + outs_b + pfx_b, kScratchBytes);
  void* base = nullptr;
  BOLT_TRY(scratch_for_current_device(&base));
  uint8_t* b = static_cast<uint8_t*>(base);
  AggScratch sc{reinterpret_cast<Acc*>(b), reinterpret_cast<Acc*>(b + parts_b),
                reinterpret_cast<Pfx*>(b + parts_b + outs_b)};
  return aggregate_impl((hipStream_t)stream, inputs, n_inputs, aggs, n_aggs, filter, results, &sc);
}

static int bolt::aggregate_impl(hipStream_t s, const bolt_column_t* inputs, int32_t n_inputs, const bolt_agg_t* aggs,
                                int32_t n_aggs, const bolt_filter_t* filter, bolt_scalar_t* results,
                                const AggScratch* scratch) {
  if (n_inputs < 1 || inputs == nullptr)
    return set_error(BOLT_ERR_PLAN, "bolt_aggregate: at least one input column is required");
  int64_t n = inputs[0].length;
  for (int i = 1; i < n_inputs; i++)
    if (inputs[i].length != n)
      return set_error(BOLT_ERR_PLAN, "bolt_aggregate: column %d has %lld rows, column 0 has %lld", i,
                       (long long)inputs[i].length, (long long)n);
  DevFilter f;
  BOLT_TRY(make_dev_filter(filter, inputs, n_inputs, n, &f));
  if (n_aggs == 0) {
    AggList none;
    return make_agg_list(aggs, n_aggs, inputs, n_inputs, 0, &none);
  }
  // Each SUM over an Int64 column also gets a hidden MIN and MAX of the
  // column: when count * max(|MIN|, |MAX|) fits i64, no running total can
  // leave i64, and the row-order check (check_sum_prefix) is skipped.
  // (fixed-size arrays: the allocation-free variant runs this too)
  const int32_t n_user = n_aggs;
  if (n_user < 0 || n_user > kMaxAggs || !aggs) return set_error(BOLT_ERR_PLAN, "%d aggregates (max %d)", n_user, kMaxAggs);
  bolt_agg_t ext[kMaxAggs];
  int n_ext = n_user;
  for (int i = 0; i < n_user; i++) ext[i] = aggs[i];
  int guard_sum[kMaxAggs], guard_mm[kMaxAggs];  // SUM index, index of its hidden MIN (-1: none)
  int n_guards = 0;
  for (int i = 0; i < n_user; i++)
    if (aggs[i].op == BOLT_AGG_SUM && aggs[i].col >= 0 && aggs[i].col < n_inputs &&
        inputs[aggs[i].col].dtype == BOLT_I64 && n >= 2) {
      guard_sum[n_guards] = i;
      guard_mm[n_guards] = -1;
      if (n_ext + 2 <= kMaxAggs) {
        guard_mm[n_guards] = n_ext;
        ext[n_ext++] = bolt_agg_t{BOLT_AGG_MIN, aggs[i].col};
        ext[n_ext++] = bolt_agg_t{BOLT_AGG_MAX, aggs[i].col};
      }
      n_guards++;
    }
  aggs = ext;
  n_aggs = n_ext;
  AggList all;
  BOLT_TRY(make_agg_list(aggs, n_aggs, inputs, n_inputs, 0, &all));

  // at most 1024 workgroups: a streaming read of 2 GiB measured 6.8-6.9 TB/s
  // at 1024 x 256 threads with nontemporal loads, 5.4-6.1 TB/s at 4096-16384
  unsigned grid = stream_grid(n, kReduceBlock, kAggUnroll * 4, 1024);
  DevBuf<Acc> parts, outs;
  Acc* outs_p = scratch ? scratch->outs : nullptr;
  if (!outs_p) {
    BOLT_TRY(outs.alloc((size_t)n_aggs, s));
    outs_p = outs.ptr;
  }
  bool done = false;
  if (!scratch && jit_enabled() && n > 0) {
    JitSpec sp;
    JitLits lits;
    JitCols cols;
    hipFunction_t fn;
    if (jit_spec_from(f, all, false, 0, false, 0, &sp, &lits, &cols) == BOLT_OK &&
        jit_get_keyed(jit_row_agg_key(sp), [&] { return jit_row_agg_source(sp); }, &fn) == BOLT_OK) {
      BOLT_TRY(parts.alloc((size_t)grid * n_aggs, s));
      prof_begin(s);
      BOLT_TRY(jit_launch_row_agg(fn, grid, s, nullptr, 0, nullptr, n, nullptr, nullptr, lits, cols, parts.ptr, nullptr, 0));
      prof_end(s);
      for (int first = 0; first < n_aggs; first += 4) {
        int na = n_aggs - first < 4 ? n_aggs - first : 4;
        BOLT_TRY(launch_finish(na, parts.ptr, (int)grid, n_aggs, first, agg_chunk(all, first, na), outs_p + first, s));
      }
      done = true;
    }
    clear_error();
  }
  if (!done) {
    int n_chunks = (n_aggs + 3) / 4;
    Acc* parts_p = scratch ? scratch->parts : nullptr;
    if (!parts_p) {
      BOLT_TRY(parts.alloc((size_t)grid * 4 * n_chunks, s));
      parts_p = parts.ptr;
    }
    for (int c = 0; c < n_chunks; c++) {
      int first = c * 4, na = n_aggs - first < 4 ? n_aggs - first : 4;
      AggList chunk = agg_chunk(all, first, na);
      Acc* p = parts_p + (size_t)c * grid * 4;
      BOLT_TRY(launch_scalar_chunk(na, f, chunk, n, grid, p, s));
      BOLT_TRY(launch_finish(na, p, (int)grid, na, 0, chunk, outs_p + first, s));
    }
  }
  Acc host[kMaxAggs];
  BOLT_HIP(hipMemcpyAsync(host, outs_p, sizeof(Acc) * n_aggs, hipMemcpyDeviceToHost, s));
  BOLT_HIP(hipStreamSynchronize(s));
  bolt_scalar_t res[kMaxAggs];
  for (int i = 0; i < n_aggs; i++) BOLT_TRY(acc_finalize(host[i], all.a[i].op, all.a[i].dtype, &res[i]));
  for (int gi = 0; gi < n_guards; gi++) {
    const int i = guard_sum[gi], mm = guard_mm[gi];
    if (res[i].is_null || acc_count(host[i]) < 2) continue;
    if (mm >= 0) {
      const __int128 lo = res[mm].i64, hi = res[mm + 1].i64;
      const unsigned __int128 a = (unsigned __int128)(lo < 0 ? -lo : lo);
      const unsigned __int128 b = (unsigned __int128)(hi < 0 ? -hi : hi);
      if ((a > b ? a : b) * (unsigned __int128)acc_count(host[i]) <= (unsigned __int128)INT64_MAX) continue;
    }
    BOLT_TRY(check_sum_prefix(s, inputs[aggs[i].col], f, scratch ? scratch->pfx : nullptr));
  }
  for (int i = 0; i < n_user; i++) results[i] = res[i];
  return BOLT_OK;
}
