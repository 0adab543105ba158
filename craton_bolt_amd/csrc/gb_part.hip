// gb_part.hip — the radix-partitioned HBM tier of hash GROUP BY.
//
// Replaces the reference's Tier-2 partitioned path (execute_tier2_sum,
// src/exec/groupby_tier2_orchestrator.rs:214: partition kernels
// src/jit/partition_kernel.rs:239,374, offsets src/exec/partition_offsets.rs:364,
// scatter src/jit/scatter_kernel.rs:99, per-partition LDS reduce
// src/jit/partition_reduce_kernel.rs:139-241), which is capped at
// 4096 x 1024 groups and one aggregate family per launch.  Here it takes
// what the LDS tier could not finish — its partial records and its spilled
// rows — so every aggregate of the query goes through one set of passes and
// there is no group cap (a partition whose groups overflow its LDS table
// sends the rest to the global table directly):
//
//   prefix    of the workgroups' spill-segment counts (one workgroup); the
//             spilled rows are then read in place in their segments;
//   histogram every item (partial record or spilled row) gets a partition
//             id from its key hash; per-workgroup counts in LDS, totals by
//             one atomic per (workgroup, partition);
//   scan      partition starts (one workgroup);
//   scatter   each workgroup reserves its share of every partition with one
//             atomic each, then writes its items as merge records
//             (key, tag, accumulator words; a spilled row becomes the words
//             of a one-row group) ranked by LDS atomics;
//   reduce    one workgroup per partition folds its records into an LDS
//             table and writes each finished group into the global table:
//             the slot is claimed with one CAS, and since no other
//             workgroup holds the key, its words are plain stores.
//
// The global table that results is the one the non-partitioned path builds,
// so extraction, the VAR / STDDEV second pass and ordering are shared.
#include <algorithm>

#include "groupby_dev.h"

namespace bolt {

constexpr int kPartBlock = 1024;
constexpr int kPartMaxBits = 12;  // up to 4096 partitions (48 KiB of LDS for the scatter)

__device__ __forceinline__ uint32_t part_of(uint64_t key, int pbits) {
  return pbits == 0 ? 0u : (uint32_t)(mix64(key) >> (64 - pbits));
}

// Raw image of a 4- or 8-byte column value: the value `pre` loaded earlier
// when the aggregate reads the column (pdata, pdtype) it was loaded from.
__device__ __forceinline__ uint64_t raw_or_pre(const GAgg& g, int64_t row, const void* pdata, int pdtype,
                                               uint64_t pre) {
  if (g.data == pdata && g.dtype == pdtype) return pre;
  return g.dtype == BOLT_I64 || g.dtype == BOLT_F64 ? static_cast<const uint64_t*>(g.data)[row]
                                                    : (uint64_t) static_cast<const uint32_t*>(g.data)[row];
}

// Accumulator words of a group holding the single row `row` (the update of
// update_slot applied to the identity).  The value of column (pdata, pdtype)
// at `row` may come preloaded in `pre` (pdata = nullptr: none).
__device__ __forceinline__ void row_words(const GPlan& P, int64_t row, unsigned long long* dst, size_t stride,
                                          const void* pdata = nullptr, int pdtype = 0, uint64_t pre = 0) {
  for (int a = 0; a < P.n; a++) {
    const GAgg& g = P.a[a];
    unsigned long long* w = dst + (size_t)g.w * stride;
    if (g.op == BOLT_AGG_COUNT_STAR) { w[0] = 1; continue; }
    const bool valid = bit_valid(g.validity, row);
    if (g.kind == K_COUNT) { w[0] = valid ? 1 : 0; continue; }
    if (g.nw >= 0) dst[(size_t)g.nw * stride] = valid ? 1 : 0;
    switch (g.kind) {
      case K_ISUM: {
        int64_t v = 0;
        if (valid) {
          const uint64_t r = raw_or_pre(g, row, pdata, pdtype, pre);
          v = g.dtype == BOLT_I64 ? (int64_t)r : (int64_t)(int32_t)(uint32_t)r;
        }
        w[0] = (uint64_t)v;
        w[stride] = (uint64_t)(v >> 63);
        break;
      }
      case K_DSUM: {
        uint64_t lo = 0, hi = 0;
        if (valid) {
          const uint64_t* p = static_cast<const uint64_t*>(g.data) + 2 * row;
          lo = p[0];
          hi = p[1];
        }
        w[0] = lo;
        w[stride] = hi;
        w[2 * stride] = (uint64_t)((int64_t)hi >> 63);
        break;
      }
      case K_FSUM:
      case K_AVG:
      case K_VAR: {
        double v = 0.0;
        if (valid) {
          const uint64_t r = raw_or_pre(g, row, pdata, pdtype, pre);
          v = g.dtype == BOLT_F64 ? as_f64(r)
            : g.dtype == BOLT_F32 ? (double)__uint_as_float((uint32_t)r)
            : g.dtype == BOLT_I64 ? (double)(int64_t)r
                                  : (double)(int32_t)(uint32_t)r;
        }
        w[0] = as_u64(v);
        break;
      }
      case K_MINMAX: {
        uint64_t k = g.op == BOLT_AGG_MIN ? ~0ull : 0ull;
        if (valid) {
          const uint64_t r = raw_or_pre(g, row, pdata, pdtype, pre);
          switch (g.dtype) {
            case BOLT_F64: k = f64_order_key(as_f64(r)); break;
            case BOLT_F32: k = f64_order_key((double)__uint_as_float((uint32_t)r)); break;
            case BOLT_I64: k = r ^ 0x8000000000000000ull; break;
            default: k = (uint64_t)(int64_t)(int32_t)(uint32_t)r ^ 0x8000000000000000ull; break;
          }
        }
        w[0] = k;
        break;
      }
    }
  }
}

// The items: partial records [0, n_parts), then the spilled rows (PartSrc).
// Segment of spilled row j: the last b with spre[b] <= j (so empty segments
// are passed over), j < n_spill.
__device__ __forceinline__ int spill_seg_find(const PartSrc& src, int64_t j) {
  int lo = 0, hi = src.nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((int64_t)src.spre[mid] <= j) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// A thread's walk over the spill segments, for spilled rows visited in
// non-decreasing order: the current segment's range [lo, hi) of j and its
// base (position = base + j) stay in registers, so a row costs a compare
// unless it crosses into a later segment.
struct SpillWalk {
  int b = -1;
  int64_t lo = 0, hi = 0, base = 0;
};

__device__ __forceinline__ void spill_walk_set(const PartSrc& src, int b, SpillWalk* w) {
  w->b = b;
  w->lo = (int64_t)src.spre[b];
  w->hi = b + 1 < src.nseg ? (int64_t)src.spre[b + 1] : INT64_MAX;
  w->base = (int64_t)b * src.per - w->lo;
}

// Position of spilled row j in the spill segments.
__device__ __forceinline__ int64_t spill_at(const PartSrc& src, int64_t j, SpillWalk* w) {
  if (w->b < 0 || j < w->lo) spill_walk_set(src, spill_seg_find(src, j), w);
  while (j >= w->hi) spill_walk_set(src, w->b + 1, w);
  return w->base + j;
}

// Key and tag of item i; *w is the calling thread's spill-segment walk.
__device__ __forceinline__ void item_key(const PartSrc& src, const KeySpec& ks, int64_t i, uint64_t* key,
                                         uint8_t* tag, SpillWalk* w) {
  if (i < src.n_parts) {
    *key = src.pkeys[i];
    *tag = src.ptag[i];
    return;
  }
  const int64_t at = spill_at(src, i - src.n_parts, w);
  uint64_t k = 0;
  int kr = 1;
  if (src.keys) k = src.keys[at];
  else kr = row_key(ks, (int64_t)src.rows[at], &k);
  // spilled rows carry regular keys; the NULL group is tagged for safety
  *tag = kr == 2 ? 1 : (k == kEmpty ? 2 : 0);
  *key = k;
}

// spre[b] = rows in the spill segments before b (spre[nseg] = all of them);
// one workgroup.
__global__ __launch_bounds__(kPartBlock) void part_spill_prefix_kernel(const uint32_t* __restrict__ seg, int nseg,
                                                                       uint64_t* __restrict__ spre) {
  __shared__ uint64_t wsum[kPartBlock / 64];
  __shared__ uint64_t carry;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nseg; base += kPartBlock) {
    const int b = base + threadIdx.x;
    const uint64_t v = b < nseg ? seg[b] : 0;
    uint64_t inc = v;
    for (int d = 1; d < 64; d <<= 1) {
      const uint64_t u = __shfl_up(inc, d);
      if (lane >= d) inc += u;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint64_t before = carry;
    for (int w = 0; w < wave; w++) before += wsum[w];
    if (b < nseg) spre[b] = before + inc - v;
    __syncthreads();
    if (threadIdx.x == kPartBlock - 1) carry = before + inc;
    __syncthreads();
  }
  if (threadIdx.x == 0) spre[nseg] = carry;
}

__global__ __launch_bounds__(kPartBlock) void part_hist_kernel(PartSrc src, KeySpec ks, int pbits, int64_t chunk,
                                                               uint16_t* __restrict__ pid,
                                                               uint32_t* __restrict__ hist,
                                                               uint32_t* __restrict__ total) {
  __shared__ uint32_t lh[1 << kPartMaxBits];
  const int np = 1 << pbits;
  for (int p = threadIdx.x; p < np; p += kPartBlock) lh[p] = 0;
  __syncthreads();
  const int64_t i0 = (int64_t)blockIdx.x * chunk;
  const int64_t i1 = std::min<int64_t>(i0 + chunk, src.n_items);
  SpillWalk sw;
  if (src.keys) {
    // spilled keys captured: four items per lane, their loads issued together
    // (a partial record reads spilled key 0 and a spilled row partial record
    // 0, both in their buffers, and the selects keep what applies)
    for (int64_t ib = i0 + threadIdx.x; ib < i1; ib += 4 * kPartBlock) {
      uint64_t kk[4];
      uint8_t tg[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        int64_t i = ib + (int64_t)u * kPartBlock;
        i = i < i1 ? i : i1 - 1;
        const bool sp = i >= src.n_parts;
        const int64_t at = sp ? spill_at(src, i - src.n_parts, &sw) : 0;
        const int64_t pi = sp ? 0 : i;
        const uint64_t ks_ = src.keys[at], kp = src.pkeys[pi];
        const uint8_t tp = src.ptag[pi];
        kk[u] = sp ? ks_ : kp;
        tg[u] = sp ? (uint8_t)(ks_ == kEmpty ? 2 : 0) : tp;
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int64_t i = ib + (int64_t)u * kPartBlock;
        if (i >= i1) continue;
        if (tg[u]) {  // the NULL group / empty-marker key: merged straight into the global table by the scatter
          pid[i] = 0xFFFFu;
          continue;
        }
        const uint32_t p = part_of(kk[u], pbits);
        pid[i] = (uint16_t)p;
        atomicAdd(&lh[p], 1u);
      }
    }
  } else
  for (int64_t i = i0 + threadIdx.x; i < i1; i += kPartBlock) {
    uint64_t key;
    uint8_t tag;
    item_key(src, ks, i, &key, &tag, &sw);
    if (tag) {  // the NULL group / empty-marker key: merged straight into the global table by the scatter
      pid[i] = 0xFFFFu;
      continue;
    }
    const uint32_t p = part_of(key, pbits);
    pid[i] = (uint16_t)p;
    atomicAdd(&lh[p], 1u);
  }
  __syncthreads();
  for (int p = threadIdx.x; p < np; p += kPartBlock) {
    const uint32_t c = lh[p];
    hist[(int64_t)blockIdx.x * np + p] = c;
    if (c) atomicAdd(&total[p], c);
  }
}

// Exclusive scan of the partition totals (one workgroup): start[p], and the
// scatter cursors initialised to the starts.  start[np] = number of items.
__global__ __launch_bounds__(kPartBlock) void part_scan_kernel(const uint32_t* __restrict__ total, int np,
                                                               uint64_t* __restrict__ start,
                                                               unsigned long long* __restrict__ cursor) {
  __shared__ uint64_t wsum[kPartBlock / 64];
  __shared__ uint64_t carry;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < np; base += kPartBlock) {
    const int p = base + threadIdx.x;
    const uint64_t v = p < np ? total[p] : 0;
    uint64_t inc = v;
    for (int d = 1; d < 64; d <<= 1) {
      const uint64_t u = __shfl_up(inc, d);
      if (lane >= d) inc += u;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint64_t before = carry;
    for (int w = 0; w < wave; w++) before += wsum[w];
    if (p < np) {
      start[p] = before + inc - v;
      cursor[p] = before + inc - v;
    }
    __syncthreads();
    if (threadIdx.x == kPartBlock - 1) carry = before + inc;
    __syncthreads();
  }
  if (threadIdx.x == 0) start[np] = carry;
}

// The merge record of spilled row j (key, then the words of a one-row
// group), the key from the captured keys when there are any.
__device__ __forceinline__ void spilled_record(const PartSrc& src, const KeySpec& ks, const GPlan& P, int64_t j,
                                               unsigned long long* r) {
  SpillWalk w;
  const int64_t at = spill_at(src, j, &w);
  uint64_t key = 0;
  if (src.keys) key = src.keys[at];
  else (void)row_key(ks, (int64_t)src.rows[at], &key);
  r[0] = key;
  row_words(P, (int64_t)src.rows[at], r + 1, 1);
}

// The first aggregate input the words of a spilled row read (a 4- or 8-byte
// column), whose values a tile loads for all its spilled rows at once; -1 if
// none.
__device__ __forceinline__ int first_value_agg(const GPlan& P) {
  for (int a = 0; a < P.n; a++) {
    const GAgg& g = P.a[a];
    if (g.op == BOLT_AGG_COUNT_STAR || g.kind == K_COUNT || g.kind == K_DSUM) continue;
    if (g.dtype == BOLT_I64 || g.dtype == BOLT_F64 || g.dtype == BOLT_I32 || g.dtype == BOLT_F32 ||
        g.dtype == BOLT_DATE32)
      return a;
  }
  return -1;
}

// Merge records, one per item: the key then the W accumulator words, `stride`
// = 1 + W 64-bit words apart.
struct PartRecs {
  unsigned long long* rec;
  int64_t cap;
  int stride;
};

__global__ __launch_bounds__(kPartBlock) void part_scatter_kernel(PartSrc src, KeySpec ks, GPlan P, int pbits,
                                                                  int64_t chunk, const uint16_t* __restrict__ pid,
                                                                  const uint32_t* __restrict__ hist,
                                                                  unsigned long long* __restrict__ cursor,
                                                                  PartRecs out, GTable t) {
  __shared__ unsigned long long lbase[1 << kPartMaxBits];  // reserved start in each partition
  __shared__ uint32_t lpos[1 << kPartMaxBits];             // items placed so far
  const int np = 1 << pbits;
  // reserve this workgroup's share of each partition it has items in (one
  // atomic each); an item's position is its reservation + an LDS rank
  for (int p = threadIdx.x; p < np; p += kPartBlock) {
    const uint32_t c = hist[(int64_t)blockIdx.x * np + p];
    lbase[p] = c ? atomicAdd(&cursor[p], (unsigned long long)c) : 0ull;
    lpos[p] = 0;
  }
  __syncthreads();
  const int64_t i0 = (int64_t)blockIdx.x * chunk;
  const int64_t i1 = std::min<int64_t>(i0 + chunk, src.n_items);
  const int W = P.words;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += kPartBlock) {
    const uint32_t p = pid[i];
    if (p == 0xFFFFu) {  // tagged partial record (never a spilled row)
      const int64_t g = gslot(t, src.ptag[i], src.pkeys[i]);
      if (g >= 0) merge_slot(P, t.words, t.cap + 2, (uint32_t)g, src.pwords + i, (size_t)src.pcap);
      continue;
    }
    const uint64_t pos = lbase[p] + atomicAdd(&lpos[p], 1u);
    unsigned long long* r = out.rec + pos * (uint64_t)out.stride;
    if (i < src.n_parts) {
      r[0] = src.pkeys[i];
      for (int w = 0; w < W; w++) r[1 + w] = src.pwords[(size_t)w * src.pcap + i];
    } else {
      spilled_record(src, ks, P, i - src.n_parts, r);
    }
  }
}

// ---- LDS-staged scatter -------------------------------------------------------------
//
// A pass moves items into the buckets of a 6-bit digit of their partition
// id: a tile of T items is ranked per digit with LDS atomics, staged in LDS in
// digit order, and written out as runs (one reservation per digit and tile),
// so a wave writes contiguous records instead of 32-byte fragments scattered
// over every partition.  Two passes (the top digit, then the rest) reach up
// to 4096 partitions; with at most 64 partitions one pass writes the final
// layout.
constexpr int kStageBlock = 512;
constexpr int kStageMaxIpt = 4;  // items per thread per tile

template <int PASS>  // 1: items -> buckets of the top digit; 2: records -> final partitions
__global__ __launch_bounds__(kStageBlock) void part_stage_kernel(PartSrc src, KeySpec ks, GPlan P, int pbits, int fb,
                                                                 const uint16_t* __restrict__ pid,
                                                                 const unsigned long long* __restrict__ in,
                                                                 unsigned long long* __restrict__ out, int stride, int T,
                                                                 unsigned long long* __restrict__ cursor,
                                                                 const uint64_t* __restrict__ start, int K2, GTable t) {
  extern __shared__ unsigned long long stage[];  // [T][stride]
  uint8_t* sdig = reinterpret_cast<uint8_t*>(stage + (size_t)T * stride);
  __shared__ uint32_t cnt[64], pre[64];
  __shared__ unsigned long long gbase[64];
  const int ipt = T / kStageBlock;
  const int W = P.words;
  const uint32_t dmask = PASS == 1 ? 63u : ((1u << fb) - 1u);
  // the range of items this workgroup tiles
  int64_t lo, hi, tile0, tstep;
  uint32_t c = 0;  // PASS 2: the top digit whose bucket this workgroup reads
  if (PASS == 1) {
    lo = 0;
    hi = src.n_items;
    tile0 = blockIdx.x;
    tstep = gridDim.x;
  } else {
    c = blockIdx.x / K2;
    lo = (int64_t)start[(uint64_t)c << fb];
    hi = (int64_t)start[(uint64_t)(c + 1) << fb];
    tile0 = blockIdx.x % K2;
    tstep = K2;
  }
  SpillWalk swalk;  // PASS 1: this thread's walk over the spill segments (tiles ascend)
  for (int64_t tile = tile0; lo + tile * T < hi; tile += tstep) {
    const int64_t t0 = lo + tile * T;
    if (threadIdx.x < 64) cnt[threadIdx.x] = 0;
    __syncthreads();
    int rank[kStageMaxIpt];
    uint32_t dig[kStageMaxIpt];
    // every item's partition id (PASS 1) or record key (PASS 2) first, with
    // clamped indices, so the loads are in flight together
    uint64_t raw[kStageMaxIpt], rw1[kStageMaxIpt], rw2[kStageMaxIpt];
    const bool narrow = PASS == 2 && stride <= 3;  // PASS 2: a record of at most three words is loaded whole
#pragma unroll
    for (int u = 0; u < kStageMaxIpt; u++) {
      int64_t i = t0 + (u < ipt ? u : 0) * kStageBlock + threadIdx.x;
      i = i < hi ? i : hi - 1;
      raw[u] = PASS == 1 ? (uint64_t)pid[i] : in[(uint64_t)i * stride];
      rw1[u] = narrow && stride >= 2 ? in[(uint64_t)i * stride + 1] : 0ull;
      rw2[u] = narrow && stride >= 3 ? in[(uint64_t)i * stride + 2] : 0ull;
    }
#pragma unroll
    for (int u = 0; u < kStageMaxIpt; u++) {
      rank[u] = -1;
      dig[u] = 0;
      if (u >= ipt) continue;
      const int64_t i = t0 + u * kStageBlock + threadIdx.x;
      if (i >= hi) continue;
      uint32_t p;
      if (PASS == 1) {
        p = (uint32_t)raw[u];
        if (p == 0xFFFFu) {  // tagged partial record: merged straight into the global table
          const int64_t g = gslot(t, src.ptag[i], src.pkeys[i]);
          if (g >= 0) merge_slot(P, t.words, t.cap + 2, (uint32_t)g, src.pwords + i, (size_t)src.pcap);
          continue;
        }
        dig[u] = fb > 0 ? (p >> fb) : p;
      } else {
        p = part_of(raw[u], pbits);
        dig[u] = p & dmask;
      }
      rank[u] = (int)atomicAdd(&cnt[dig[u]], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 64) {  // one wave: digit offsets within the tile, one reservation per digit
      const uint32_t v = cnt[threadIdx.x];
      uint32_t inc = v;
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(inc, d);
        if ((int)threadIdx.x >= d) inc += o;
      }
      pre[threadIdx.x] = inc - v;
      const uint64_t dest = PASS == 1 ? threadIdx.x : (((uint64_t)c << fb) | threadIdx.x);
      gbase[threadIdx.x] = v ? atomicAdd(&cursor[dest], (unsigned long long)v) : 0ull;
    }
    __syncthreads();
    // PASS 1: the spilled rows' ids, captured keys and first input value for
    // all of this lane's items, loaded before any record is built (indices
    // clamped, so the loads carry no branches and are in flight together)
    int64_t srow[kStageMaxIpt];
    uint64_t skey[kStageMaxIpt], sval[kStageMaxIpt];
    const int va = PASS == 1 ? first_value_agg(P) : -1;
    const int64_t n_spill = src.n_items - src.n_parts;
    if (PASS == 1 && n_spill > 0) {
#pragma unroll
      for (int u = 0; u < kStageMaxIpt; u++) {
        // slots past this tile's items per lane repeat the last one (same
        // lines), so j never decreases along u and the segment walk holds
        int64_t j = t0 + (u < ipt ? u : ipt - 1) * kStageBlock + threadIdx.x - src.n_parts;
        j = j < 0 ? 0 : (j >= n_spill ? n_spill - 1 : j);
        const int64_t at = spill_at(src, j, &swalk);
        srow[u] = (int64_t)src.rows[at];
        skey[u] = src.keys ? src.keys[at] : 0ull;
      }
      if (va >= 0) {
        const GAgg& g = P.a[va];
        const bool w8 = g.dtype == BOLT_I64 || g.dtype == BOLT_F64;
#pragma unroll
        for (int u = 0; u < kStageMaxIpt; u++)
          sval[u] = w8 ? static_cast<const uint64_t*>(g.data)[srow[u]]
                       : (uint64_t) static_cast<const uint32_t*>(g.data)[srow[u]];
      }
    }
#pragma unroll
    for (int u = 0; u < kStageMaxIpt; u++) {
      if (rank[u] < 0) continue;
      const int64_t i = t0 + u * kStageBlock + threadIdx.x;
      const uint32_t sl = pre[dig[u]] + (uint32_t)rank[u];
      sdig[sl] = (uint8_t)dig[u];
      unsigned long long* r = stage + (size_t)sl * stride;
      if (PASS == 2 && narrow) {
        r[0] = raw[u];
        if (stride >= 2) r[1] = rw1[u];
        if (stride >= 3) r[2] = rw2[u];
      } else if (PASS == 2) {
        const unsigned long long* q = in + (uint64_t)i * stride;
        for (int w = 0; w < stride; w++) r[w] = q[w];
      } else if (i < src.n_parts) {
        r[0] = src.pkeys[i];
        for (int w = 0; w < W; w++) r[1 + w] = src.pwords[(size_t)w * src.pcap + i];
      } else {
        uint64_t key = skey[u];
        if (!src.keys) (void)row_key(ks, srow[u], &key);
        r[0] = key;
        if (va >= 0) row_words(P, srow[u], r + 1, 1, P.a[va].data, P.a[va].dtype, sval[u]);
        else row_words(P, srow[u], r + 1, 1);
      }
    }
    __syncthreads();
    // write the staged records out word by word, consecutive lanes on
    // consecutive words, so each digit's run is one contiguous stream of
    // stores whatever the record width
    const uint32_t total = pre[63] + cnt[63];
    const uint32_t nwords = total * (uint32_t)stride;
    for (uint32_t u = threadIdx.x; u < nwords; u += kStageBlock) {
      const uint32_t j = u / (uint32_t)stride, w = u - j * (uint32_t)stride;
      const uint32_t d = sdig[j];
      const uint64_t pos = gbase[d] + (j - pre[d]);
      out[pos * (uint64_t)stride + w] = stage[u];
    }
    __syncthreads();
  }
}

__global__ void part_init_cursors_kernel(const uint64_t* __restrict__ start, unsigned long long* __restrict__ fine,
                                         unsigned long long* __restrict__ coarse, int np, int fb) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < np) fine[p] = start[p];
  if (fb > 0 && p < (np >> fb)) coarse[p] = start[(uint64_t)p << fb];
}

// Global slot of `key`, claiming an empty one; *fresh says whether this call
// claimed it.  0xFFFFFFFF when the probe bound is exceeded.
__device__ __forceinline__ uint32_t global_claim(unsigned long long* keys, uint64_t cap_mask, uint64_t key,
                                                 bool* fresh) {
  uint64_t s = mix64(key) & cap_mask;
  for (uint64_t p = 0; p < (uint64_t)kGlobalProbes && p <= cap_mask; p++) {
    const unsigned long long cur = keys[s];
    if (cur == key) { *fresh = false; return (uint32_t)s; }
    if (cur == kEmpty) {
      const unsigned long long prev = atomicCAS(&keys[s], kEmpty, (unsigned long long)key);
      if (prev == kEmpty) { *fresh = true; return (uint32_t)s; }
      if (prev == key) { *fresh = false; return (uint32_t)s; }
    }
    s = (s + 1) & cap_mask;
  }
  return 0xFFFFFFFFu;
}

__global__ __launch_bounds__(kGbBlock) void part_reduce_kernel(GTable t, GPlan P, PartRecs recs,
                                                               const uint64_t* __restrict__ start, int S) {
  extern __shared__ unsigned long long sm[];
  const int SS = S + 2;
  unsigned long long* skeys = sm;
  unsigned long long* swords = sm + SS;
  uint32_t* sused = reinterpret_cast<uint32_t*>(swords + (size_t)P.words * SS);
  for (int i = threadIdx.x; i < SS; i += kGbBlock) skeys[i] = kEmpty;
  for (int w = 0; w < P.words; w++) {
    const unsigned long long id = word_identity(P, w);
    for (int i = threadIdx.x; i < SS; i += kGbBlock) swords[(size_t)w * SS + i] = id;
  }
  if (threadIdx.x < 3) sused[threadIdx.x] = 0;  // [2]: occupied slots
  __syncthreads();
  const int64_t r0 = (int64_t)start[blockIdx.x], r1 = (int64_t)start[blockIdx.x + 1];
  // four records per lane per iteration: their keys are loaded together, so
  // the lane waits on one HBM round trip per four records (the words then sit
  // in the lines the keys brought in)
  constexpr int kU = 4;
  if (P.words <= 2) {
    // narrow plans: each record's words are loaded with its key and merged
    // from registers, so a lane waits on one round trip per four records
    const uint64_t st = (uint64_t)recs.stride;
    for (int64_t base = r0 + threadIdx.x; base < r1; base += (int64_t)kU * kGbBlock) {
      unsigned long long keys[kU], v0[kU], v1[kU];
#pragma unroll
      for (int u = 0; u < kU; u++) {
        int64_t r = base + (int64_t)u * kGbBlock;
        r = r < r1 ? r : r1 - 1;
        const unsigned long long* rec = recs.rec + (uint64_t)r * st;
        keys[u] = rec[0];
        v0[u] = rec[1];
        v1[u] = P.words == 2 ? rec[2] : 0ull;
      }
#pragma unroll
      for (int u = 0; u < kU; u++) {
        if (base + (int64_t)u * kGbBlock >= r1) continue;
        const int slot = gb_lds_slot(skeys, S, keys[u], &sused[2]);
        if (slot >= 0) {
          merge_slot2(P, swords, SS, (uint32_t)slot, v0[u], v1[u]);
        } else {  // the partition holds more groups than the LDS table: straight to HBM
          const int64_t g = gslot(t, 0, keys[u]);
          if (g >= 0) merge_slot2(P, t.words, t.cap + 2, (uint32_t)g, v0[u], v1[u]);
        }
      }
    }
  } else
  for (int64_t base = r0 + threadIdx.x; base < r1; base += (int64_t)kU * kGbBlock) {
    unsigned long long keys[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const int64_t r = base + (int64_t)u * kGbBlock;
      keys[u] = r < r1 ? recs.rec[r * (uint64_t)recs.stride] : 0ull;
    }
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const int64_t r = base + (int64_t)u * kGbBlock;
      if (r >= r1) continue;
      const unsigned long long* rec = recs.rec + r * (uint64_t)recs.stride;
      const int slot = gb_lds_slot(skeys, S, keys[u], &sused[2]);
      if (slot >= 0) {
        merge_slot(P, swords, SS, (uint32_t)slot, rec + 1, 1);
      } else {  // the partition holds more groups than the LDS table: straight to HBM
        const int64_t g = gslot(t, 0, keys[u]);
        if (g >= 0) merge_slot(P, t.words, t.cap + 2, (uint32_t)g, rec + 1, 1);
      }
    }
  }
  __syncthreads();
  const size_t gstride = t.cap + 2;
  for (int i = threadIdx.x; i < SS; i += kGbBlock) {
    if (i < S) {
      const unsigned long long key = skeys[i];
      if (key == kEmpty) continue;
      bool fresh = false;
      const uint32_t g = global_claim(t.keys, t.cap - 1, key, &fresh);
      if (g == 0xFFFFFFFFu) { atomicOr(&t.err[0], 1u); continue; }
      if (fresh) {  // no other workgroup holds this key: plain stores
        for (int w = 0; w < P.words; w++) t.words[(size_t)w * gstride + g] = swords[(size_t)w * SS + i];
      } else {      // records of this partition that overflowed the LDS table got there first
        merge_slot(P, t.words, gstride, g, swords + i, (size_t)SS);
      }
    }
  }
}

__global__ void part_reset_cursor_kernel(const uint64_t* __restrict__ start, unsigned long long* __restrict__ cursor,
                                         int np) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < np) cursor[p] = start[p];
}

// ---- host ---------------------------------------------------------------------------

// The fan-out: partitions of about S/2 groups each, so one LDS table holds a
// partition, moved by `bits_delta` (the autotuner's choice), in 0 .. 12 bits.
int gb_part_bits(double est, int S, int bits_delta) {
  const double want = std::max(1.0, est / std::max(1.0, 0.5 * S));
  int pbits = 0;
  while ((1 << pbits) < want && pbits < kPartMaxBits) pbits++;
  return std::min(kPartMaxBits, std::max(0, pbits + bits_delta));
}

int gb_part_build(hipStream_t s, const KeySpec& ks, const GPlan& P, const unsigned long long* pkeys,
                  const uint8_t* ptag, const unsigned long long* pwords, int64_t pcap, int64_t n_parts,
                  const uint32_t* spill, const unsigned long long* spill_key, int64_t spill_per,
                  const uint32_t* spill_seg, int nseg, int64_t n_spill, double est, int S, int bits_delta,
                  GbPartitions* out) {
  const int64_t n_items = n_parts + n_spill;
  int pbits = gb_part_bits(est, S, bits_delta);
  const int np = 1 << pbits;
  out->np = np;
  // the spilled rows are read in place in the LDS tier's segments, through
  // the prefix of the segments' counts
  BOLT_TRY(out->spre.alloc((size_t)std::max(nseg, 1) + 1, s));
  if (n_spill > 0) part_spill_prefix_kernel<<<1, kPartBlock, 0, s>>>(spill_seg, nseg, out->spre.ptr);
  PartSrc src{pkeys, ptag,   pwords,   pcap,          n_parts, spill, n_items, spill_key,
              spill_per, out->spre.ptr, nseg};
  const int nb = (int)std::min<int64_t>(512, std::max<int64_t>(1, (n_items + 65535) / 65536));
  const int64_t chunk = (n_items + nb - 1) / nb;
  DevBuf<uint32_t> total;
  BOLT_TRY(out->pid.alloc((size_t)std::max<int64_t>(n_items, 1), s));
  BOLT_TRY(out->hist.alloc((size_t)nb * np, s));
  BOLT_TRY(total.alloc((size_t)np, s));
  BOLT_TRY(out->cursor.alloc((size_t)np, s));
  BOLT_TRY(out->start.alloc((size_t)np + 1, s));
  BOLT_TRY(fill_u32(total.ptr, 0, (size_t)np, s));
  part_hist_kernel<<<nb, kPartBlock, 0, s>>>(src, ks, pbits, chunk, out->pid.ptr, out->hist.ptr, total.ptr);
  part_scan_kernel<<<1, kPartBlock, 0, s>>>(total.ptr, np, out->start.ptr, out->cursor.ptr);
  out->stride = 1 + P.words;  // key + words, unpadded (the staged writes are word streams)
  out->recs_cap = std::max<int64_t>(n_items, 1);
  BOLT_TRY(out->rec.alloc((size_t)out->recs_cap * out->stride, s));
  if (pbits > 6) BOLT_TRY(out->coarse.alloc((size_t)out->recs_cap * out->stride, s));
  BOLT_TRY(out->ccursor.alloc(64, s));
  out->src = src;
  out->nb = nb;
  out->chunk = chunk;
  out->pbits = pbits;
  out->ks = ks;
  BOLT_HIP(hipGetLastError());
  return BOLT_OK;
}

// The scatter needs the global table (tagged records go there directly), so
// it runs per table attempt, with the reduce.
int gb_part_reduce(hipStream_t s, const GTable& t, const GPlan& P, GbPartitions& parts, int S) {
  PartRecs recs{parts.rec.ptr, parts.recs_cap, parts.stride};
  const int fb = parts.pbits > 6 ? parts.pbits - 6 : 0;  // bits of the second pass
  int T = 2048;
  while (T > kStageBlock && (size_t)T * parts.stride * 8 > 65536) T >>= 1;
  if ((size_t)T * parts.stride * 8 <= 65536) {
    // LDS-staged: one pass (<= 64 partitions) or two
    const size_t lds = (size_t)T * parts.stride * 8 + T;
    BOLT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(part_stage_kernel<1>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    BOLT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(part_stage_kernel<2>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    part_init_cursors_kernel<<<(parts.np + 255) / 256, 256, 0, s>>>(parts.start.ptr, parts.cursor.ptr,
                                                                    parts.ccursor.ptr, parts.np, fb);
    const int64_t tiles = (parts.src.n_items + T - 1) / T;
    const unsigned g1 = (unsigned)std::max<int64_t>(1, std::min<int64_t>(1024, tiles));
    part_stage_kernel<1><<<g1, kStageBlock, lds, s>>>(parts.src, parts.ks, P, parts.pbits, fb, parts.pid.ptr, nullptr,
                                                      fb > 0 ? parts.coarse.ptr : parts.rec.ptr, parts.stride, T,
                                                      fb > 0 ? parts.ccursor.ptr : parts.cursor.ptr,
                                                      parts.start.ptr, 1, t);
    if (fb > 0) {
      const int K2 = 8;  // workgroups per top-digit bucket
      part_stage_kernel<2><<<64 * K2, kStageBlock, lds, s>>>(parts.src, parts.ks, P, parts.pbits, fb, nullptr,
                                                             parts.coarse.ptr, parts.rec.ptr, parts.stride, T,
                                                             parts.cursor.ptr, parts.start.ptr, K2, t);
    }
  } else {
    // records too wide to stage: scatter each item to its partition directly
    part_reset_cursor_kernel<<<(parts.np + 255) / 256, 256, 0, s>>>(parts.start.ptr, parts.cursor.ptr, parts.np);
    part_scatter_kernel<<<parts.nb, kPartBlock, 0, s>>>(parts.src, parts.ks, P, parts.pbits, parts.chunk,
                                                        parts.pid.ptr, parts.hist.ptr, parts.cursor.ptr, recs, t);
  }
  const size_t lds = (size_t)(S + 2) * 8 * (1 + P.words) + 16;
  BOLT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(part_reduce_kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  part_reduce_kernel<<<(unsigned)parts.np, kGbBlock, lds, s>>>(t, P, recs, parts.start.ptr, S);
  BOLT_HIP(hipGetLastError());
  return BOLT_OK;
}

}  // namespace bolt
