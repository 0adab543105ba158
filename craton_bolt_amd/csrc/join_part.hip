// join_part.hip — radix-partitioned form of the fused filter + hash-join
// probe + aggregate, used by bolt_join_probe_aggregate when the table is far
// larger than the L2 caches.
//
// The reference probes its table at random (src/jit/hash_join_kernel.rs:711,
// src/exec/gpu_join.rs:1583).  On MI355X a random probe into a table that
// misses the 4 MB per-XCD L2 costs one 128-byte line per 16-byte slot read,
// and the chip sustains about 58 G such requests/s, so a 2 GiB table caps the
// probe far below the HBM roofline.  Here the probe side is partitioned by the
// table region its keys hash to, and each region is probed by workgroups of
// one XCD while it sits in that XCD's L2:
//
//   1. jp_hist_kernel     per (workgroup chunk, partition) row counts (keys only)
//   2. jp_offsets_kernel  per-partition prefix over chunks, partition totals
//   3. jp_base_kernel     partition bases (one workgroup)
//   4. jp_scatter_kernel  filter, then key + the probe columns the aggregates
//                         read, staged by partition in LDS and written as runs
//   5. jp_probe_kernel    persistent: workgroup b serves XCD b % 8 (workgroups
//                         are dispatched to the XCDs round-robin), whose work
//                         items (a few chunk runs of one partition each,
//                         partitions in order) it takes from a per-XCD queue,
//                         so at any time an XCD's workgroups probe one or two
//                         partitions: a table region of a few MB that stays in
//                         its L2
//
// A partition is the slot range [p << shift, (p + 1) << shift) of the table,
// whose home slots are fmix64(key) & mask, so the table is the one the
// region build writes and the random probe reads; a chain that runs past its
// partition's end simply reads the next slots.  Counts from step 1 include
// rows the filter drops: each (chunk, partition) run has room for all of them
// and step 4 records how many it wrote.
#include <vector>

#include "jit.h"
#include "join_dev.h"
#include "reduce_common.h"
// after the headers above, which bring in the HIP runtime and device_acc.h
#include "jp_tile.h"

namespace bolt {

constexpr int kJpGrid = 256;       // chunks: one per histogram / scatter workgroup
constexpr int kJpProbeBlock = 256;
constexpr int kJpProbeUnroll = 4;  // rows in flight per lane in the probe
constexpr int kJpProbeRuns = 4;    // chunk runs per probe work item (kJpGrid is a multiple)

// Probe columns carried with the key: sources and their partitioned copies.
struct JpPay {
  const void* src[kJpMaxPay];
  int32_t dtype[kJpMaxPay];
  JpDst dst;
  int32_t n;
};

template <typename KeyT>
__device__ __forceinline__ int64_t jp_key(const KeyT* k, int64_t i) {
  return (int64_t)__builtin_nontemporal_load(k + i);
}

// counts[g][p] = rows of chunk g with a valid key whose home slot lies in partition p
template <typename KeyT>
__global__ __launch_bounds__(kJpBlock) void jp_hist_kernel(const KeyT* __restrict__ keys,
                                                           const uint8_t* __restrict__ kvalid, int64_t n, int64_t per,
                                                           uint64_t mask, int shift, int P,
                                                           uint32_t* __restrict__ counts) {
  __shared__ uint32_t sc[kJpMaxParts];
  for (int p = threadIdx.x; p < P; p += kJpBlock) sc[p] = 0;
  __syncthreads();
  const int64_t lo = (int64_t)blockIdx.x * per;
  const int64_t hi = lo + per < n ? lo + per : n;
  for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += 4 * kJpBlock) {
    int64_t k[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t i = i0 + u * kJpBlock;
      k[u] = i < hi ? jp_key(keys, i) : 0;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int64_t i = i0 + u * kJpBlock;
      if (i < hi && bit_valid(kvalid, i)) atomicAdd(&sc[(fmix64((uint64_t)k[u]) & mask) >> shift], 1u);
    }
  }
  __syncthreads();
  for (int p = threadIdx.x; p < P; p += kJpBlock) counts[(int64_t)blockIdx.x * P + p] = sc[p];
}

// counts[g][p] becomes the rows of partition p in chunks before g; totals[p] its rows
__global__ __launch_bounds__(256) void jp_offsets_kernel(uint32_t* __restrict__ counts, int G, int P,
                                                         uint32_t* __restrict__ totals) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= P) return;
  uint32_t run = 0;
  for (int g0 = 0; g0 < G; g0 += 16) {
    uint32_t c[16];
#pragma unroll
    for (int j = 0; j < 16; j++) c[j] = g0 + j < G ? counts[(int64_t)(g0 + j) * P + p] : 0u;
#pragma unroll
    for (int j = 0; j < 16; j++) {
      if (g0 + j < G) counts[(int64_t)(g0 + j) * P + p] = run;
      run += c[j];
    }
  }
  totals[p] = run;
}

// base[p] = rows of partitions before p (exclusive scan of P <= 1024 totals)
__global__ __launch_bounds__(kJpMaxParts) void jp_base_kernel(const uint32_t* __restrict__ totals, int P,
                                                              uint32_t* __restrict__ base) {
  __shared__ uint32_t s[kJpMaxParts];
  const int t = threadIdx.x;
  const uint32_t v = t < P ? totals[t] : 0u;
  s[t] = v;
  __syncthreads();
  for (int d = 1; d < kJpMaxParts; d <<= 1) {
    const uint32_t a = t >= d ? s[t - d] : 0u;
    __syncthreads();
    s[t] += a;
    __syncthreads();
  }
  if (t < P) base[t] = s[t] - v;
}

// Chunk g, in tiles of kJpBlock * ITEMS rows: the rows that pass the filter
// (and have a valid key) are ranked within their partition with LDS atomics,
// staged in LDS in partition order, and written to the partitioned key and
// payload arrays at each partition's running cursor, so a tile's rows of one
// partition leave as one contiguous run.  actual[g][p] = rows written.
template <typename KeyT, int NT, int NPAY, int ITEMS>
__global__ __launch_bounds__(kJpBlock) void jp_scatter_kernel(const KeyT* __restrict__ keys,
                                                              const uint8_t* __restrict__ kvalid, int64_t n,
                                                              int64_t per, DevFilter f, JpPay pay, uint64_t mask,
                                                              int shift, int P, const uint32_t* __restrict__ counts,
                                                              const uint32_t* __restrict__ base,
                                                              int64_t* __restrict__ okeys,
                                                              uint32_t* __restrict__ actual) {
  constexpr int T = kJpBlock * ITEMS;
  __shared__ JpTileLds<T, NPAY> L;
  const int g = blockIdx.x;
  jp_tile_begin(L, P, g, counts, base);
  const int64_t lo = (int64_t)g * per;
  const int64_t hi = lo + per < n ? lo + per : n;
  for (int64_t t0 = lo; t0 < hi; t0 += T) {
    i64_t k[ITEMS];
    u64_t pv[ITEMS][NPAY > 0 ? NPAY : 1];
    int part[ITEMS];
#pragma unroll
    for (int u = 0; u < ITEMS; u++) {
      const int64_t i = t0 + u * kJpBlock + threadIdx.x;
      k[u] = jp_key(keys, i < hi ? i : hi - 1);
    }
    // the width test is uniform, so each column's ITEMS loads issue together
#pragma unroll
    for (int c = 0; c < NPAY; c++) {
      if (pay.dst.w[c] == 8) {
#pragma unroll
        for (int u = 0; u < ITEMS; u++) {
          const int64_t i = t0 + u * kJpBlock + threadIdx.x;
          pv[u][c] = __builtin_nontemporal_load(static_cast<const uint64_t*>(pay.src[c]) + (i < hi ? i : hi - 1));
        }
      } else {
#pragma unroll
        for (int u = 0; u < ITEMS; u++) {
          const int64_t i = t0 + u * kJpBlock + threadIdx.x;
          pv[u][c] = __builtin_nontemporal_load(static_cast<const uint32_t*>(pay.src[c]) + (i < hi ? i : hi - 1));
        }
      }
    }
#pragma unroll
    for (int u = 0; u < ITEMS; u++) {
      const int64_t i = t0 + u * kJpBlock + threadIdx.x;
      const bool keep = i < hi && bit_valid(kvalid, i) && eval_filter_nt<NT>(f, i);
      part[u] = keep ? (int)((fmix64((uint64_t)k[u]) & mask) >> shift) : -1;
    }
    jp_tile_stage<T, NPAY, ITEMS>(L, P, k, part, pv);
    jp_tile_flush<T, NPAY>(L, P, reinterpret_cast<i64_t*>(okeys), pay.dst);
  }
  jp_tile_end<T, NPAY>(L, P, g, counts, base, actual);
}

// Probe.  Workgroup b serves XCD x = b % 8 (workgroups are dispatched to the
// XCDs round-robin).  The work of XCD x is its partitions x, x + 8, ... in
// order, each cut into items of CG consecutive chunk runs, which the XCD's
// workgroups take from a per-XCD queue, so at any time they cover one or two
// partitions, whose table region stays in the XCD's L2.  An item's CG run
// lengths are read by every lane (broadcast loads) and its rows are flattened
// across the runs, so a lane's four rows in flight may come from different
// runs.  Aggregates read the carried build column from the table (src 1) or
// the partitioned probe columns (AggDesc data points at them).
template <int NA, int CG>
__global__ __launch_bounds__(kJpProbeBlock) void jp_probe_kernel(
    const uint4* __restrict__ tbl, uint64_t mask, const uint32_t* __restrict__ nonunique,
    const int64_t* __restrict__ pkeys, int P, int G, const uint32_t* __restrict__ counts,
    const uint32_t* __restrict__ base, const uint32_t* __restrict__ actual, uint32_t* __restrict__ queue,
    AggList aggs, int carry_dtype, Acc* __restrict__ parts) {
  Acc acc[NA];
#pragma unroll
  for (int a = 0; a < NA; a++) acc_init(acc[a], aggs.a[a].op);
  __shared__ uint32_t sh_item;
  const bool first_only = (*nonunique == 0u);
  const int x = blockIdx.x & 7;
  const uint32_t per_part = (uint32_t)(G / CG);
  const uint32_t items = (uint32_t)(P / 8) * per_part;
  for (;;) {
    if (threadIdx.x == 0) sh_item = atomicAdd(&queue[x], 1u);
    __syncthreads();
    const uint32_t item = sh_item;
    __syncthreads();
    if (item >= items) break;
    const int p = x + 8 * (int)(item / per_part);
    const int g0 = (int)(item % per_part) * CG;
    int64_t st[CG];
    uint32_t pre[CG + 1];
    pre[0] = 0;
#pragma unroll
    for (int c = 0; c < CG; c++) {
      const int64_t cell = (int64_t)(g0 + c) * P + p;
      st[c] = (int64_t)base[p] + counts[cell];
      pre[c + 1] = pre[c] + actual[cell];
    }
    const uint32_t len = pre[CG];
    for (uint32_t r0 = 0; r0 < len; r0 += kJpProbeBlock * kJpProbeUnroll) {
      int64_t k[kJpProbeUnroll], idx[kJpProbeUnroll];
      uint64_t s[kJpProbeUnroll];
      uint4 e[kJpProbeUnroll];
#pragma unroll
      for (int u = 0; u < kJpProbeUnroll; u++) {
        uint32_t v = r0 + u * kJpProbeBlock + threadIdx.x;
        v = v < len ? v : len - 1;
        int64_t i = st[0] + v;
#pragma unroll
        for (int c = 1; c < CG; c++)
          if (v >= pre[c]) i = st[c] + (v - pre[c]);
        idx[u] = i;
        k[u] = __builtin_nontemporal_load(pkeys + i);
        s[u] = fmix64((uint64_t)k[u]) & mask;
      }
#pragma unroll
      for (int u = 0; u < kJpProbeUnroll; u++) e[u] = tbl[s[u]];
#pragma unroll
      for (int u = 0; u < kJpProbeUnroll; u++) {
        if (r0 + u * kJpProbeBlock + threadIdx.x >= len) continue;
        for (uint64_t step = 0; e[u].w != 0u && step <= mask; step++) {
          if (slot_key(e[u]) == k[u]) {
#pragma unroll
            for (int a = 0; a < NA; a++) {
              const AggDesc& d = aggs.a[a];
              if (d.op == BOLT_AGG_COUNT_STAR) {
                acc[a].n++;
                continue;
              }
              int64_t vi, vhi = 0;
              double vf;
              // a probe column's value is read only for a matching row: the
              // runs are in probe order, so matching lanes read neighbours
              if (d.src == 1) carry_value(e[u].z, carry_dtype, vi, vf);
              else raw_value(load_raw(d.data, d.dtype, idx[u]), d.dtype, vi, vhi, vf);
              acc_update(acc[a], d.op, d.dtype, vi, vhi, vf);
            }
            if (first_only) break;
          }
          s[u] = (s[u] + 1) & mask;
          e[u] = tbl[s[u]];
        }
      }
    }
  }
  block_reduce_store<NA>(acc, aggs, parts + (int64_t)blockIdx.x * NA);
}

// Whether the partitioned form runs for this table and probe side.  It is
// opt-in (BOLT_JOIN_PARTITIONED=1, where its limits allow): measured on the
// 16M x 256M fused probe its generated kernels take 6.8 ms (hist 0.35,
// scatter 2.72, probe 3.70) against 5.3 ms for the random probe, which stays
// the default.  Both probes are bound by the L2 requests one CU's L1 keeps in
// flight; the region-local one only shortens their latency.
bool join_partitioned_wanted(const bolt_join_table_t* t, int64_t n_probe, const AggList& aggs) {
  const char* e = getenv("BOLT_JOIN_PARTITIONED");
  const bool forced_on = e && e[0] == '1';
  if (!forced_on) return false;
  if (t->bloom || n_probe <= 0 || n_probe >= ((int64_t)1 << 32)) return false;
  int npay = 0;
  const void* seen[kMaxAggs];
  for (int a = 0; a < aggs.n; a++) {
    const AggDesc& d = aggs.a[a];
    if (d.src != 0 || d.op == BOLT_AGG_COUNT_STAR) continue;
    if (d.validity || !(dtype_size(d.dtype) == 4 || dtype_size(d.dtype) == 8)) return false;
    bool dup = false;
    for (int j = 0; j < npay; j++) dup |= seen[j] == d.data;
    if (!dup) {
      if (npay == kJpMaxPay) return false;
      seen[npay++] = d.data;
    }
  }
  return true;
}

template <typename KeyT, int NT, int NPAY>
static void launch_scatter(const bolt_column_t& key, int64_t per, const DevFilter& f, const JpPay& pay,
                           uint64_t mask, int shift, int P, const uint32_t* counts, const uint32_t* base,
                           int64_t* okeys, uint32_t* actual, hipStream_t s) {
  constexpr int ITEMS = NPAY <= 1 ? 8 : 4;
  jp_scatter_kernel<KeyT, NT, NPAY, ITEMS><<<kJpGrid, kJpBlock, 0, s>>>(
      static_cast<const KeyT*>(key.data), key.validity, key.length, per, f, pay, mask, shift, P, counts, base, okeys,
      actual);
}

template <typename KeyT, int NT>
static void launch_scatter_np(int npay, const bolt_column_t& key, int64_t per, const DevFilter& f, const JpPay& pay,
                              uint64_t mask, int shift, int P, const uint32_t* counts, const uint32_t* base,
                              int64_t* okeys, uint32_t* actual, hipStream_t s) {
  switch (npay) {
    case 0: launch_scatter<KeyT, NT, 0>(key, per, f, pay, mask, shift, P, counts, base, okeys, actual, s); break;
    case 1: launch_scatter<KeyT, NT, 1>(key, per, f, pay, mask, shift, P, counts, base, okeys, actual, s); break;
    default: launch_scatter<KeyT, NT, 2>(key, per, f, pay, mask, shift, P, counts, base, okeys, actual, s); break;
  }
}

template <typename KeyT>
static void launch_scatter_all(int nt, int npay, const bolt_column_t& key, int64_t per, const DevFilter& f,
                               const JpPay& pay, uint64_t mask, int shift, int P, const uint32_t* counts,
                               const uint32_t* base, int64_t* okeys, uint32_t* actual, hipStream_t s) {
  switch (nt) {
    case 0: launch_scatter_np<KeyT, 0>(npay, key, per, f, pay, mask, shift, P, counts, base, okeys, actual, s); break;
    case 1: launch_scatter_np<KeyT, 1>(npay, key, per, f, pay, mask, shift, P, counts, base, okeys, actual, s); break;
    case 2: launch_scatter_np<KeyT, 2>(npay, key, per, f, pay, mask, shift, P, counts, base, okeys, actual, s); break;
    default: launch_scatter_np<KeyT, -1>(npay, key, per, f, pay, mask, shift, P, counts, base, okeys, actual, s); break;
  }
}

// Shape of the generated scatter: filter columns and carried columns share one
// column set, so a column that is both is loaded once.
static int jp_scatter_spec(const DevFilter& f, const JpPay& pay, const bolt_column_t& key, JpJitSpec* sp,
                           JitLits* lits, JitCols* cols) {
  *sp = JpJitSpec();
  memset(lits, 0, sizeof *lits);
  memset(cols, 0, sizeof *cols);
  sp->items = pay.n <= 1 ? 8 : 4;
  if (const char* e = getenv("BOLT_JOIN_SCATTER_ITEMS")) {  // rows per lane per tile: 2, 4 or 8 (LDS permitting)
    const int v = atoi(e);
    if ((v == 2 || v == 4 || v == 8) && (v < 8 || pay.n <= 1)) sp->items = v;
  }
  sp->key_dtype = key.dtype == BOLT_I64 ? BOLT_I64 : BOLT_I32;
  sp->key_valid = key.validity != nullptr;
  for (int t = 0; t < f.n; t++) {
    const int c = jit_row_agg_col(f.t[t].data, f.t[t].validity, f.t[t].dtype, &sp->cols, cols);
    if (c < 0) return set_error(BOLT_ERR_UNSUPPORTED, "jit: too many columns");
    sp->terms.push_back({c, f.t[t].op});
    lits->i[t] = f.t[t].lit_i64;
    lits->f[t] = f.t[t].lit_f64;
    lits->hi[t] = f.t[t].lit_hi;
  }
  for (int j = 0; j < pay.n; j++) {
    const int c = jit_row_agg_col(pay.src[j], nullptr, pay.dtype[j], &sp->cols, cols);
    if (c < 0) return set_error(BOLT_ERR_UNSUPPORTED, "jit: too many columns");
    sp->pay.push_back(c);
  }
  return BOLT_OK;
}

static std::string jp_probe_jit_key(const JitSpec& sp, int runs) {
  std::string k = "jpp:" + std::to_string(runs) + ':' + std::to_string(sp.unroll) + ':' +
                  std::to_string(sp.probe_slots) + ':' + std::to_string(sp.carry_dtype) + '|';
  for (const JitSpec::Col& c : sp.cols) k += std::to_string(c.dtype) + (c.has_valid ? "v," : "n,");
  k += '|';
  for (const JitSpec::Agg& g : sp.aggs)
    k += std::to_string(g.op) + '.' + std::to_string(g.src) + '.' + std::to_string(g.col) + '.' +
         std::to_string(g.dtype) + ',';
  return k;
}

// The partitioned probe + aggregate: `all` as make_agg_list built it over the
// probe columns; writes the n_aggs finished accumulators to `outs` (device).
int join_probe_aggregate_partitioned(hipStream_t s, const bolt_join_table_t* t, const bolt_column_t& key,
                                     const DevFilter& f, const AggList& all, Acc* outs) {
  const int n_aggs = all.n;
  const int64_t n = key.length;
  int L = 0;
  while ((1ull << L) < t->mask + 1) L++;
  // partitions of about 2^19 slots (8 MB), 8 to 1024 of them: at 2 MB (2^17
  // slots, half an XCD's L2) the scatter's runs are too short (1024 regions
  // for the 16M x 256M join: 4.1 ms scatter), at 8 MB they are four times
  // longer (3.36 ms) while the probe, whose regions then sit in the MALL
  // rather than the L2, is as fast (4.58 against 4.48 ms)
  int pbits = L - 19;
  if (const char* e = getenv("BOLT_JOIN_PART_BITS")) pbits = atoi(e);
  pbits = pbits < 3 ? 3 : (pbits > 10 ? 10 : pbits);
  if (pbits > L) pbits = L;
  const int P = 1 << pbits;
  const int shift = L - pbits;
  // probe columns read by the aggregates, carried with the key
  JpPay pay;
  memset(&pay, 0, sizeof pay);
  AggList pa = all;  // the aggregates over the partitioned copies
  std::vector<DevBuf<uint64_t>> pbufs(kJpMaxPay);
  for (int a = 0; a < n_aggs; a++) {
    AggDesc& d = pa.a[a];
    if (d.src != 0 || d.op == BOLT_AGG_COUNT_STAR) continue;
    int c = 0;
    while (c < pay.n && pay.src[c] != d.data) c++;
    if (c == pay.n) {
      if (pay.n == kJpMaxPay) return set_error(BOLT_ERR_OTHER, "partitioned join: too many probe columns");
      pay.src[c] = d.data;
      pay.dtype[c] = d.dtype;
      pay.dst.w[c] = dtype_size(d.dtype);
      BOLT_TRY(pbufs[c].alloc((size_t)((n * pay.dst.w[c] + 7) / 8), s));
      pay.dst.d[c] = pbufs[c].ptr;
      pay.n++;
    }
    d.data = pay.dst.d[c];
  }
  const int64_t per = (n + kJpGrid - 1) / kJpGrid;
  DevBuf<uint32_t> counts, actual, totals, base, queue;
  DevBuf<int64_t> okeys;
  DevBuf<Acc> parts;
  BOLT_TRY(counts.alloc((size_t)kJpGrid * P, s));
  BOLT_TRY(actual.alloc((size_t)kJpGrid * P, s));
  BOLT_TRY(totals.alloc((size_t)P, s));
  BOLT_TRY(base.alloc((size_t)P, s));
  BOLT_TRY(queue.alloc(8, s));
  BOLT_TRY(okeys.alloc((size_t)n, s));
  BOLT_TRY(fill_u32(queue.ptr, 0, 8, s));
  int dev = 0, cus = 256;
  BOLT_HIP(hipGetDevice(&dev));
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  // four 256-thread workgroups per CU, a multiple of 8 so every XCD has its share
  unsigned pgrid = (unsigned)((cus * 4 + 7) / 8 * 8);
  if (const char* e = getenv("BOLT_JOIN_PART_PROBE_GRID")) pgrid = (unsigned)((atoi(e) + 7) / 8 * 8);
  if (pgrid < 8) pgrid = 8;
  BOLT_TRY(parts.alloc((size_t)pgrid * n_aggs, s));
  const int nt = f.n;
  // generated scatter and probe (hipRTC), else the precompiled kernels
  hipFunction_t jscatter = nullptr, jprobe = nullptr;
  JpJitSpec ssp;
  JitLits slits;
  JitCols scols;
  JitSpec psp;
  JitLits plits;
  JitCols pcols;
  if (jit_enabled()) {
    DevFilter none;
    memset(&none, 0, sizeof none);
    if (jp_scatter_spec(f, pay, key, &ssp, &slits, &scols) == BOLT_OK &&
        jit_spec_from(none, pa, true, BOLT_I64, false, t->carry_dtype, &psp, &plits, &pcols) == BOLT_OK &&
        jit_get_keyed(jp_scatter_jit_key(ssp), [&] { return jit_jp_scatter_source(ssp); }, &jscatter) == BOLT_OK &&
        jit_get_keyed(jp_probe_jit_key(psp, kJpProbeRuns), [&] { return jit_jp_probe_source(psp, kJpProbeRuns); },
                      &jprobe) == BOLT_OK) {
    } else {
      jscatter = jprobe = nullptr;
    }
    clear_error();
  }
  prof_begin(s);
  if (key.dtype == BOLT_I64) {
    jp_hist_kernel<int64_t><<<kJpGrid, kJpBlock, 0, s>>>(static_cast<const int64_t*>(key.data), key.validity, n,
                                                         per, t->mask, shift, P, counts.ptr);
  } else {
    jp_hist_kernel<int32_t><<<kJpGrid, kJpBlock, 0, s>>>(static_cast<const int32_t*>(key.data), key.validity, n,
                                                         per, t->mask, shift, P, counts.ptr);
  }
  jp_offsets_kernel<<<(P + 255) / 256, 256, 0, s>>>(counts.ptr, kJpGrid, P, totals.ptr);
  jp_base_kernel<<<1, kJpMaxParts, 0, s>>>(totals.ptr, P, base.ptr);
  if (jscatter) {
    const void* kp = key.data;
    const uint8_t* kv = key.validity;
    long long nn = n, pp = per;
    unsigned long long m = t->mask;
    int sh = shift, PP = P;
    JpDst dst = pay.dst;
    void* args[] = {&kp, &kv, &nn, &pp, &slits, &scols, &m, &sh, &PP, &counts.ptr, &base.ptr, &okeys.ptr, &dst,
                    &actual.ptr};
    BOLT_HIP(hipModuleLaunchKernel(jscatter, kJpGrid, 1, 1, kJpBlock, 1, 1, 0, s, args, nullptr));
  } else if (key.dtype == BOLT_I64) {
    launch_scatter_all<int64_t>(nt, pay.n, key, per, f, pay, t->mask, shift, P, counts.ptr, base.ptr, okeys.ptr,
                                actual.ptr, s);
  } else {
    launch_scatter_all<int32_t>(nt, pay.n, key, per, f, pay, t->mask, shift, P, counts.ptr, base.ptr, okeys.ptr,
                                actual.ptr, s);
  }
  BOLT_HIP(hipGetLastError());
  if (jprobe) {
    const uint4* tb = t->tbl;
    unsigned long long m = t->mask;
    const uint32_t* fl = t->flags;
    int PP = P, GG = kJpGrid;
    void* args[] = {&tb, &m, &fl, &okeys.ptr, &PP, &GG, &counts.ptr, &base.ptr, &actual.ptr, &queue.ptr, &pcols,
                    &parts.ptr};
    BOLT_HIP(hipModuleLaunchKernel(jprobe, pgrid, 1, 1, kJpProbeBlock, 1, 1, 0, s, args, nullptr));
    for (int first = 0; first < n_aggs; first += 4) {
      const int na = n_aggs - first < 4 ? n_aggs - first : 4;
      BOLT_TRY(launch_finish(na, parts.ptr, (int)pgrid, n_aggs, first, agg_chunk(all, first, na), outs + first, s));
    }
    prof_end(s);
    return BOLT_OK;
  }
  for (int first = 0; first < n_aggs; first += 4) {
    const int na = n_aggs - first < 4 ? n_aggs - first : 4;
    AggList chunk = agg_chunk(pa, first, na);
    if (first > 0) BOLT_TRY(fill_u32(queue.ptr, 0, 8, s));
#define JP_PROBE(NA)                                                                                                 \
  jp_probe_kernel<NA, kJpProbeRuns><<<pgrid, kJpProbeBlock, 0, s>>>(t->tbl, t->mask, t->flags, okeys.ptr, P, kJpGrid, \
                                                                    counts.ptr, base.ptr, actual.ptr, queue.ptr,      \
                                                                    chunk, t->carry_dtype,                            \
                                                                    parts.ptr + (size_t)first * pgrid)
    switch (na) {
      case 1: JP_PROBE(1); break;
      case 2: JP_PROBE(2); break;
      case 3: JP_PROBE(3); break;
      default: JP_PROBE(4); break;
    }
#undef JP_PROBE
    BOLT_HIP(hipGetLastError());
    BOLT_TRY(launch_finish(na, parts.ptr + (size_t)first * pgrid, (int)pgrid, na, 0, agg_chunk(all, first, na),
                           outs + first, s));
  }
  prof_end(s);
  return BOLT_OK;
}

}  // namespace bolt
