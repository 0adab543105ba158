// join.hip — INNER hash join: build, probe (pair emission), and the fused
// filter + probe + aggregate form.
//
// Replaces hash_join_indices_on_gpu_resident (src/exec/gpu_join.rs:1583),
// the build / probe kernels (src/jit/hash_join_kernel.rs:484,711) and the
// collision-list variants for duplicate keys (:1183,1368).
//
// Table layout (MI355X-first).  The reference uses an SoA open-addressing
// table (i64 keys with an i64::MIN empty sentinel + a separate u32 row-id
// array, load <= 0.5, src/exec/gpu_join.rs:577), so a hit costs two scattered
// loads.  On MI355X a scattered probe is limited by the number of distinct
// 16-byte requests it issues, not by HBM bytes (a per-lane probe costs the
// same against a 1 MB and a 256 MB table), so here a slot is one 16-byte
// word that one load returns whole:
//   { i64 key, u32 value (build row id or a carried 4-byte column), u32 state }
// with state 0 = empty, so any int64 key is storable (no sentinel).  Linear
// probing at a low load factor (default 0.15, set by BOLT_JOIN_LOAD_FACTOR)
// keeps chains short: a miss rarely reads past its home slot.  The 288 GB of
// HBM makes the larger table (2 GiB for 16M build rows) cheap; on the fused
// 16M x 256M probe, load 0.25 (1 GiB) costs 0.5 ms more per probe than the
// 0.16 ms it saves in the build.
#include <cstdlib>
#include <cstring>

#include "jit.h"
#include "join_dev.h"
#include "reduce_common.h"
#include "scan_common.h"

namespace bolt {

constexpr int kBuildBlock = 256;
constexpr int kProbeBlock = 256;

__device__ __forceinline__ int64_t load_key(const void* p, int dtype, int64_t i) {
  return dtype == BOLT_I64 ? static_cast<const int64_t*>(p)[i] : (int64_t) static_cast<const int32_t*>(p)[i];
}


// ---- build -------------------------------------------------------------------

__global__ __launch_bounds__(kBuildBlock) void join_build_kernel(uint4* __restrict__ tbl, uint64_t mask,
                                                                 const void* __restrict__ key, int kdt,
                                                                 const uint8_t* __restrict__ kvalid,
                                                                 const uint32_t* __restrict__ carry, int64_t n,
                                                                 uint32_t* __restrict__ lost) {
  const int64_t stride = (int64_t)gridDim.x * kBuildBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBuildBlock + threadIdx.x; i < n; i += stride) {
    if (!bit_valid(kvalid, i)) continue;  // NULL build keys never match (src/exec/join.rs:851)
    const int64_t k = load_key(key, kdt, i);
    const uint32_t v = carry ? carry[i] : (uint32_t)i;
    uint64_t s = fmix64((uint64_t)k) & mask;
    bool placed = false;
    for (uint64_t step = 0; step <= mask; step++) {
      uint32_t* st = &tbl[s].w;
      if (*st == 0u && atomicCAS(st, 0u, 1u) == 0u) {
        tbl[s].x = (uint32_t)k;
        tbl[s].y = (uint32_t)((uint64_t)k >> 32);
        tbl[s].z = v;
        placed = true;
        break;
      }
      s = (s + 1) & mask;
    }
    if (!placed) atomicAdd(lost, 1u);
  }
}

__global__ __launch_bounds__(kBuildBlock) void bloom_build_kernel(uint4* __restrict__ bloom, uint64_t bmask,
                                                                  const void* __restrict__ key, int kdt,
                                                                  const uint8_t* __restrict__ kvalid, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * kBuildBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBuildBlock + threadIdx.x; i < n; i += stride) {
    if (!bit_valid(kvalid, i)) continue;
    const uint64_t h = fmix64((uint64_t)load_key(key, kdt, i));
    const uint4 p = bloom_pattern(h);
    uint32_t* b = reinterpret_cast<uint32_t*>(bloom + ((h >> 40) & bmask));
    atomicOr(b + 0, p.x);
    atomicOr(b + 1, p.y);
    atomicOr(b + 2, p.z);
    atomicOr(b + 3, p.w);
  }
}

// ---- region build ---------------------------------------------------------------
//
// Large tables are built region by region instead of by one global claim
// atomic per row: the rows are bucketed by the 4096-slot region of their home
// slot (an LDS histogram per workgroup, offsets, an LDS-cursor scatter), then
// one workgroup per region inserts its rows into an LDS copy of the region,
// checks it for duplicate keys, and writes the whole region out, empty slots
// included, so the table needs no separate fill.  Rows whose probe runs past
// the end of their region (about one in ten thousand at load 0.15) are
// appended to an overflow list and inserted afterwards with global claims.
constexpr int kRegBits = 12;
constexpr int kRegSlots = 1 << kRegBits;
constexpr int kMaxRegions = 32768;  // u32 counters of every region in one workgroup's LDS
constexpr int kRegGrid = 256;       // workgroups of the histogram and scatter passes
constexpr int kRegBlock = 1024;

// counts[b][r] = valid rows of workgroup b's chunk whose home lies in region r
__global__ __launch_bounds__(kRegBlock) void join_region_hist_kernel(const void* __restrict__ key, int kdt,
                                                                     const uint8_t* __restrict__ kvalid, int64_t n,
                                                                     int64_t per, uint64_t mask, int n_regions,
                                                                     uint32_t* __restrict__ counts) {
  extern __shared__ uint32_t sc[];
  for (int r = threadIdx.x; r < n_regions; r += kRegBlock) sc[r] = 0;
  __syncthreads();
  const int64_t lo = (int64_t)blockIdx.x * per;
  const int64_t hi = lo + per < n ? lo + per : n;
  for (int64_t i = lo + threadIdx.x; i < hi; i += kRegBlock) {
    if (!bit_valid(kvalid, i)) continue;
    const uint64_t h = fmix64((uint64_t)load_key(key, kdt, i)) & mask;
    atomicAdd(&sc[h >> kRegBits], 1u);
  }
  __syncthreads();
  for (int r = threadIdx.x; r < n_regions; r += kRegBlock) counts[(int64_t)blockIdx.x * n_regions + r] = sc[r];
}

// counts[b][r] becomes the rows of region r in chunks before b; totals[r] the region's rows
__global__ __launch_bounds__(256) void join_region_offsets_kernel(uint32_t* __restrict__ counts, int blocks,
                                                                  int n_regions, uint32_t* __restrict__ totals) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= n_regions) return;
  uint32_t run = 0;
  for (int b0 = 0; b0 < blocks; b0 += 16) {  // 16 loads in flight per lane
    uint32_t c[16];
#pragma unroll
    for (int j = 0; j < 16; j++) c[j] = b0 + j < blocks ? counts[(int64_t)(b0 + j) * n_regions + r] : 0u;
#pragma unroll
    for (int j = 0; j < 16; j++) {
      if (b0 + j < blocks) counts[(int64_t)(b0 + j) * n_regions + r] = run;
      run += c[j];
    }
  }
  totals[r] = run;
}

// base[r] = rows of regions before r (base[n_regions] = all); one workgroup
__global__ __launch_bounds__(kRegBlock) void join_region_scan_kernel(const uint32_t* __restrict__ totals,
                                                                     int n_regions, uint32_t* __restrict__ base) {
  __shared__ uint32_t part[kRegBlock];
  const int per = (n_regions + kRegBlock - 1) / kRegBlock;
  const int lo = threadIdx.x * per;
  uint32_t sum = 0;
  for (int r = lo; r < lo + per && r < n_regions; r++) sum += totals[r];
  part[threadIdx.x] = sum;
  __syncthreads();
  // inclusive scan of the 1024 partial sums (Hillis-Steele in LDS)
  for (int d = 1; d < kRegBlock; d <<= 1) {
    const uint32_t v = threadIdx.x >= (unsigned)d ? part[threadIdx.x - d] : 0u;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  if (threadIdx.x == kRegBlock - 1) base[n_regions] = part[threadIdx.x];
  uint32_t run = part[threadIdx.x] - sum;  // exclusive
  for (int r = lo; r < lo + per && r < n_regions; r++) {
    base[r] = run;
    run += totals[r];
  }
}

// Scatters the valid rows of workgroup b's chunk into their regions' buckets
// as slot-shaped 16-byte records.
__global__ __launch_bounds__(kRegBlock) void join_region_scatter_kernel(
    const void* __restrict__ key, int kdt, const uint8_t* __restrict__ kvalid, const uint32_t* __restrict__ carry,
    int64_t n, int64_t per, uint64_t mask, int n_regions, const uint32_t* __restrict__ counts,
    const uint32_t* __restrict__ base, uint4* __restrict__ brec) {
  extern __shared__ uint32_t cur[];
  for (int r = threadIdx.x; r < n_regions; r += kRegBlock)
    cur[r] = base[r] + counts[(int64_t)blockIdx.x * n_regions + r];
  __syncthreads();
  const int64_t lo = (int64_t)blockIdx.x * per;
  const int64_t hi = lo + per < n ? lo + per : n;
  for (int64_t i = lo + threadIdx.x; i < hi; i += kRegBlock) {
    if (!bit_valid(kvalid, i)) continue;
    const int64_t k = load_key(key, kdt, i);
    const uint64_t h = fmix64((uint64_t)k) & mask;
    const uint32_t pos = atomicAdd(&cur[h >> kRegBits], 1u);
    // one 16-byte record (a slot's layout, state 1) per row: one scattered write
    brec[pos] = make_uint4((uint32_t)k, (uint32_t)((uint64_t)k >> 32), carry ? carry[i] : (uint32_t)i, 1u);
  }
}

// One workgroup per region: linear probing in LDS from each row's home slot,
// no wrap (a row that reaches the region's end overflows), then the
// duplicate check of the join_dup_scan_kernel on the final LDS copy, then
// the region written out whole.
constexpr int kRegBuildBlock = 1024;

__global__ __launch_bounds__(kRegBuildBlock) void join_region_build_kernel(const uint4* __restrict__ brec,
                                                                const uint32_t* __restrict__ base, uint64_t mask,
                                                                uint4* __restrict__ tbl, int64_t* __restrict__ okey,
                                                                uint32_t* __restrict__ oval,
                                                                unsigned long long* __restrict__ ocount,
                                                                uint32_t* __restrict__ nonunique) {
  __shared__ uint4 sl[kRegSlots];
  const int r = blockIdx.x;
  for (int t = threadIdx.x; t < kRegSlots; t += kRegBuildBlock) sl[t] = make_uint4(0u, 0u, 0u, 0u);
  __syncthreads();
  const uint32_t lo = base[r], hi = base[r + 1];
  for (uint32_t i = lo + threadIdx.x; i < hi; i += kRegBuildBlock) {
    const uint4 rec = brec[i];
    const int64_t k = slot_key(rec);
    const uint32_t v = rec.z;
    bool placed = false;
    for (int t = (int)((fmix64((uint64_t)k) & mask) & (kRegSlots - 1)); t < kRegSlots; t++) {
      if (sl[t].w == 0u && atomicCAS(&sl[t].w, 0u, 1u) == 0u) {
        sl[t].x = (uint32_t)k;
        sl[t].y = (uint32_t)((uint64_t)k >> 32);
        sl[t].z = v;
        placed = true;
        break;
      }
    }
    if (!placed) {
      const unsigned long long p = atomicAdd(ocount, 1ull);  // the list holds every row: p < rows
      okey[p] = k;
      oval[p] = v;
    }
  }
  __syncthreads();
  for (int s = threadIdx.x; s < kRegSlots; s += kRegBuildBlock) {
    const uint4 e = sl[s];
    if (e.w == 0u) continue;
    const int64_t k = slot_key(e);
    for (int t = (int)((fmix64((uint64_t)k) & mask) & (kRegSlots - 1)); t < s; t++)
      if (slot_key(sl[t]) == k) {
        *nonunique = 1u;
        break;
      }
  }
  uint4* out = tbl + (uint64_t)r * kRegSlots;
  typedef unsigned int v4u __attribute__((ext_vector_type(4)));
  for (int t = threadIdx.x; t < kRegSlots; t += kRegBuildBlock) {  // streaming stores: the table is not read back soon
    const uint4 e = sl[t];
    const v4u v = {e.x, e.y, e.z, e.w};
    __builtin_nontemporal_store(v, reinterpret_cast<v4u*>(out + t));
  }
}

// Global claims for the overflow rows (after every region is written).
__global__ __launch_bounds__(kBuildBlock) void join_overflow_insert_kernel(uint4* __restrict__ tbl, uint64_t mask,
                                                                           const int64_t* __restrict__ okey,
                                                                           const uint32_t* __restrict__ oval,
                                                                           const unsigned long long* __restrict__ ocount,
                                                                           uint32_t* __restrict__ lost) {
  const int64_t n = (int64_t)*ocount;
  const int64_t stride = (int64_t)gridDim.x * kBuildBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBuildBlock + threadIdx.x; i < n; i += stride) {
    const int64_t k = okey[i];
    uint64_t s = fmix64((uint64_t)k) & mask;
    bool placed = false;
    for (uint64_t step = 0; step <= mask; step++) {
      uint32_t* st = &tbl[s].w;
      if (*st == 0u && atomicCAS(st, 0u, 1u) == 0u) {
        tbl[s].x = (uint32_t)k;
        tbl[s].y = (uint32_t)((uint64_t)k >> 32);
        tbl[s].z = oval[i];
        placed = true;
        break;
      }
      s = (s + 1) & mask;
    }
    if (!placed) atomicAdd(lost, 1u);
  }
}

// An overflow row's key stored more than once anywhere on its chain (after
// every insert) marks the table non-unique; copies that stayed in one region
// were checked there.
__global__ __launch_bounds__(kBuildBlock) void join_overflow_dup_kernel(const uint4* __restrict__ tbl, uint64_t mask,
                                                                        const int64_t* __restrict__ okey,
                                                                        const unsigned long long* __restrict__ ocount,
                                                                        uint32_t* __restrict__ nonunique) {
  const int64_t n = (int64_t)*ocount;
  const int64_t stride = (int64_t)gridDim.x * kBuildBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBuildBlock + threadIdx.x; i < n; i += stride) {
    int c = 0;
    probe_chain(tbl, mask, okey[i], false, [&](uint32_t) { c++; });
    if (c > 1) *nonunique = 1u;
  }
}

// Marks the table non-unique if any key is stored twice, by one sequential
// pass over the finished table: an occupied slot s holding k was claimed after
// its insert walked every slot from k's home to s, all occupied, so an earlier
// copy of k can only sit in [home(k), s).  That range is usually empty or in
// s's own cache line, so the pass costs about one read of the table instead of
// a random chain walk per build row.
__global__ __launch_bounds__(kBuildBlock) void join_dup_scan_kernel(const uint4* __restrict__ tbl, uint64_t mask,
                                                                    uint32_t* __restrict__ nonunique) {
  const uint64_t stride = (uint64_t)gridDim.x * kBuildBlock;
  for (uint64_t s = (uint64_t)blockIdx.x * kBuildBlock + threadIdx.x; s <= mask; s += stride) {
    const uint4 e = tbl[s];
    if (e.w == 0u) continue;
    const int64_t k = slot_key(e);
    for (uint64_t t = fmix64((uint64_t)k) & mask; t != s; t = (t + 1) & mask) {
      const uint4 f = tbl[t];
      if (slot_key(f) == k) {
        *nonunique = 1u;
        break;
      }
    }
  }
}

// ---- fused filter + probe + aggregate (generic: predicate and aggregates
// are read from kernel arguments; instantiated for 0-2 predicate terms and
// 1-4 aggregates) ---------------------------------------------------------------

template <typename KeyT, int NT, int NA>
__global__ __launch_bounds__(kProbeBlock) void join_probe_agg_kernel(
    const uint4* __restrict__ tbl, uint64_t mask, const uint32_t* __restrict__ nonunique,
    const KeyT* __restrict__ keys, const uint8_t* __restrict__ kvalid, int64_t n, DevFilter f, AggList aggs,
    int carry_dtype, Acc* __restrict__ parts, const uint4* __restrict__ bloom, uint64_t bmask) {
  Acc acc[NA];
#pragma unroll
  for (int a = 0; a < NA; a++) acc_init(acc[a], aggs.a[a].op);
  const bool first_only = (*nonunique == 0u);
  const int64_t stride = (int64_t)gridDim.x * kProbeBlock;
  for (int64_t i = (int64_t)blockIdx.x * kProbeBlock + threadIdx.x; i < n; i += stride) {
    const int64_t k = (int64_t)__builtin_nontemporal_load(keys + i);
    Raw ra[NA];
#pragma unroll
    for (int a = 0; a < NA; a++)
      if (aggs.a[a].src == 0 && aggs.a[a].op != BOLT_AGG_COUNT_STAR) ra[a] = load_raw(aggs.a[a].data, aggs.a[a].dtype, i);
    if (!bit_valid(kvalid, i) || !eval_filter_nt<NT>(f, i) || !bloom_maybe(bloom, bmask, k)) continue;
    probe_chain(tbl, mask, k, first_only, [&](uint32_t w) {
#pragma unroll
      for (int a = 0; a < NA; a++) {
        const AggDesc& d = aggs.a[a];
        if (d.op == BOLT_AGG_COUNT_STAR) { acc[a].n++; continue; }
        int64_t vi, vhi = 0;
        double vf;
        if (d.src == 1) {
          carry_value(w, carry_dtype, vi, vf);
        } else {
          if (!bit_valid(d.validity, i)) continue;
          raw_value(ra[a], d.dtype, vi, vhi, vf);
        }
        acc_update(acc[a], d.op, d.dtype, vi, vhi, vf);
      }
    });
  }
  block_reduce_store<NA>(acc, aggs, parts + (int64_t)blockIdx.x * NA);
}

// Key columns of one side of a join: up to kMaxJoinKeys columns.  `mode` 0:
// one column (key_word); 1: two 32-bit columns packed first-high / second-low
// (TwoI32, src/exec/gpu_join.rs:788); 2: a 64-bit hash of the columns' values,
// equal for equal keys but not only for them, so every hash match is checked
// against the columns themselves (the exact form of the reference's
// candidate + host-verify path for TwoI64 / MultiI32, :840 and
// verify_pairs_on_host).
constexpr int kMaxJoinKeys = 4;
struct KeyCols {
  const void* d[kMaxJoinKeys];
  const uint8_t* v[kMaxJoinKeys];
  int dt[kMaxJoinKeys];
  int n_cols;
  int mode;
  uint64_t hash_mask;  // all ones; fewer bits force hash collisions (tests of the check)
};

__device__ __forceinline__ uint64_t key_word(const void* d, int dt, int64_t i) {
  switch (dt) {
    case BOLT_I64: return (uint64_t) static_cast<const int64_t*>(d)[i];
    case BOLT_F64: {
      const uint64_t b = static_cast<const uint64_t*>(d)[i];
      if ((b & 0x7FFFFFFFFFFFFFFFull) > 0x7FF0000000000000ull) return 0x7FF8000000000000ull;
      return (b & 0x7FFFFFFFFFFFFFFFull) == 0 ? 0ull : b;
    }
    case BOLT_F32: {
      const uint32_t b = static_cast<const uint32_t*>(d)[i];
      if ((b & 0x7FFFFFFFu) > 0x7F800000u) return 0x7FC00000ull;
      return (b & 0x7FFFFFFFu) == 0 ? 0ull : (uint64_t)b;
    }
    case BOLT_BOOL: return static_cast<const uint8_t*>(d)[i] != 0 ? 1ull : 0ull;
    default: return (uint64_t)(int64_t) static_cast<const int32_t*>(d)[i];  // Int32 / Date32
  }
}

// Probe row `i` and build row `b` hold equal (non-NULL) keys.
__device__ __forceinline__ bool same_key(const KeyCols& p, int64_t i, const KeyCols& b, uint32_t r) {
  for (int c = 0; c < p.n_cols; c++)
    if (key_word(p.d[c], p.dt[c], i) != key_word(b.d[c], b.dt[c], r)) return false;
  return true;
}

// ---- probe emitting (probe_row, build_row) pairs -----------------------------
//
// Tiled with a decoupled look-back, like the compaction: a tile of
// kEmitTile probe rows counts its matches, one wave looks back across
// earlier tiles for the tile's output offset, and the pairs are written
// densely.  No output counter is shared (one per wave serialises every wave
// on a single address).  Each lane owns kEmitItems consecutive rows, so the
// pairs come out in probe-row order.  With unique build keys one walk finds each row's
// match and keeps it in a register; with duplicates the chain is walked
// twice (count, then emit).
//
// kPreserve (LEFT / RIGHT / FULL: the probe side is the preserved side) emits
// one (probe_row, kNullRow) pair for a probe row that finds no match, in its
// place in probe-row order, as the reference's host outer join does
// (src/exec/join.rs:396-437).  A non-null `unmatched` (FULL) has one byte per
// build row, 1 on entry; every build row that matches is cleared to 0 with a
// plain byte store (all writers store the same value).
constexpr int kEmitItems = 8;
constexpr int kEmitTile = kProbeBlock * kEmitItems;  // 2048 rows
constexpr uint32_t kNullRow = 0xFFFFFFFFu;

template <typename KeyT, bool kPreserve, bool kVerify>
__global__ __launch_bounds__(kProbeBlock) void join_probe_emit_kernel(
    const uint4* __restrict__ tbl, uint64_t mask, const uint32_t* __restrict__ nonunique,
    const KeyT* __restrict__ keys, const uint8_t* __restrict__ kvalid, int64_t n, DevFilter f,
    uint64_t* __restrict__ status, uint32_t* __restrict__ ticket, uint32_t* __restrict__ err,
    unsigned long long* __restrict__ total_out, int64_t capacity, uint32_t* __restrict__ out_probe,
    uint32_t* __restrict__ out_build, const uint4* __restrict__ bloom, uint64_t bmask,
    uint8_t* __restrict__ unmatched, KeyCols pk, KeyCols bk) {
  __shared__ uint32_t sh_tile;
  __shared__ uint32_t wsum[kProbeBlock / 64];
  __shared__ unsigned long long sh_base;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) sh_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  const int64_t tile = sh_tile;
  const int64_t tbase = tile * kEmitTile;
  const bool first_only = (*nonunique == 0u);
  uint32_t m[kEmitItems], bv[kEmitItems];
  uint32_t mine = 0, pad = 0, live = 0;
  int64_t kk[kEmitItems];
  uint64_t ss[kEmitItems];
  uint4 ee[kEmitItems];
  // keys and the home slots of the lane's rows first, all in flight together,
  // then each row's chain from its already-loaded home slot
#pragma unroll
  for (int j = 0; j < kEmitItems; j++) {
    const int64_t i = tbase + ((int64_t)threadIdx.x * kEmitItems + j);
    const bool lv = i < n && eval_filter(f, i);
    live |= (uint32_t)lv << j;
    const bool ok = lv && bit_valid(kvalid, i);
    kk[j] = ok ? (int64_t)keys[i] : 0;
    ss[j] = ok && bloom_maybe(bloom, bmask, kk[j]) ? fmix64((uint64_t)kk[j]) & mask : ~0ull;
  }
#pragma unroll
  for (int j = 0; j < kEmitItems; j++) ee[j] = ss[j] != ~0ull ? tbl[ss[j]] : make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
  for (int j = 0; j < kEmitItems; j++) {
    const int64_t i = tbase + ((int64_t)threadIdx.x * kEmitItems + j);
    m[j] = 0;
    bv[j] = 0;
    uint4 e = ee[j];
    uint64_t sl = ss[j];
    for (uint64_t step = 0; e.w != 0u && step <= mask; step++) {
      if (slot_key(e) == kk[j] && (!kVerify || same_key(pk, i, bk, e.z))) {
        m[j]++;
        bv[j] = e.z;
        if (unmatched) unmatched[e.z] = 0;
        if (first_only) break;
      }
      sl = (sl + 1) & mask;
      e = tbl[sl];
    }
    if (kPreserve && ((live >> j) & 1u) && m[j] == 0) {
      m[j] = 1;
      bv[j] = kNullRow;
      pad |= 1u << j;
    }
    mine += m[j];
  }
  // lane prefix within the wave, wave prefix within the tile
  uint32_t inc = mine;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t v = __shfl_up(inc, d);
    if (lane >= d) inc += v;
  }
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  if (wave == 0) {
    uint32_t tile_total = 0;
    for (int w = 0; w < kProbeBlock / 64; w++) tile_total += wsum[w];
    const uint64_t excl = lookback_wave(status, tile, tile_total, err);
    if (lane == 0) {
      sh_base = excl;
      if (tile == (int64_t)gridDim.x - 1) *total_out = excl + tile_total;
    }
  }
  __syncthreads();
  unsigned long long pos = sh_base + inc - mine;
  for (int w = 0; w < wave; w++) pos += wsum[w];
#pragma unroll
  for (int j = 0; j < kEmitItems; j++) {
    if (!m[j]) continue;
    const int64_t i = tbase + ((int64_t)threadIdx.x * kEmitItems + j);
    if (first_only || (pad >> j) & 1u) {
      if ((int64_t)pos < capacity) {
        out_probe[pos] = (uint32_t)i;
        out_build[pos] = bv[j];
      }
      pos++;
    } else {
      probe_chain(tbl, mask, (int64_t)keys[i], false, [&](uint32_t w) {
        if (kVerify && !same_key(pk, i, bk, w)) return;
        if ((int64_t)pos < capacity) {
          out_probe[pos] = (uint32_t)i;
          out_build[pos] = w;
        }
        pos++;
      });
    }
  }
}

// ---- join keys that are not a plain Int32 / Int64 column -----------------------
//
// One int64 per row (encode_keys_for_shape, src/exec/gpu_join.rs:719),
// equal exactly when the SQL keys are equal in modes 0 and 1: floats compare
// by their canonical bits, -0.0 as +0.0 and every NaN as one NaN
// (canonical_f32_bits / canonical_f64_bits, :682-705), Bool as 0 / 1, two
// 32-bit columns packed.  Mode 2 hashes the columns (see KeyCols).  A row
// with a NULL in any key column is NULL (src/exec/join.rs:1045).  The validity bitmap is written a
// 64-row word at a time from a ballot (`vout` is 8-byte aligned; NULL when no
// key column has one).
__global__ __launch_bounds__(256) void join_key_kernel(KeyCols kc, int64_t n, int64_t* __restrict__ out,
                                                       uint64_t* __restrict__ vout) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t end = (n + 63) / 64 * 64;  // wave-uniform trip count for the ballot
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < end; i += stride) {
    const bool in = i < n;
    bool valid = in;
    uint64_t k = 0;
    if (in) {
      if (kc.mode == 0) {
        k = key_word(kc.d[0], kc.dt[0], i);
      } else if (kc.mode == 1) {
        k = ((uint64_t) static_cast<const uint32_t*>(kc.d[0])[i] << 32) | static_cast<const uint32_t*>(kc.d[1])[i];
      } else {
        for (int c = 0; c < kc.n_cols; c++) k = fmix64((k * 0x9E3779B97F4A7C15ull) ^ key_word(kc.d[c], kc.dt[c], i));
        k &= kc.hash_mask;
      }
      for (int c = 0; c < kc.n_cols; c++) valid = valid && bit_valid(kc.v[c], i);
      out[i] = (int64_t)k;
    }
    if (vout) {
      const uint64_t b = __ballot(in && valid);
      if ((threadIdx.x & 63) == 0) vout[i >> 6] = b;
    }
  }
}

// FULL OUTER's second part: build rows no probe row matched, ascending, each
// as (kNullRow, build_row) at base + its rank among them
// (src/exec/join.rs:440-449).
__global__ __launch_bounds__(256) void join_unmatched_kernel(const uint8_t* __restrict__ unmatched,
                                                             const uint32_t* __restrict__ rank, int64_t n,
                                                             int64_t base, int64_t capacity,
                                                             uint32_t* __restrict__ out_probe,
                                                             uint32_t* __restrict__ out_build) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if (!unmatched[i]) continue;
    const int64_t pos = base + (int64_t)rank[i];
    if (pos < capacity) {
      out_probe[pos] = kNullRow;
      out_build[pos] = (uint32_t)i;
    }
  }
}

}  // namespace bolt

using namespace bolt;


static size_t join_table_cap_bytes() {
  // BOLT_GPU_JOIN_TABLE_CAP_MB (src/exec/gpu_join.rs:393) overrides; the
  // default is a quarter of device memory (the reference caps at 512 MiB on
  // devices with >= 8 GiB, gpu_join.rs:378-388, sized for consumer cards).
  const char* env = getenv("BOLT_GPU_JOIN_TABLE_CAP_MB");
  if (env && *env) return (size_t)strtoull(env, nullptr, 10) << 20;
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
    (void)hipGetLastError();
    return (size_t)512 << 20;
  }
  return tot / 4;
}

static int check_key_col(const bolt_column_t* c, const char* what) {
  if (c == nullptr) return set_error(BOLT_ERR_PLAN, "%s: key column is NULL", what);
  if (c->dtype != BOLT_I32 && c->dtype != BOLT_I64 && c->dtype != BOLT_DATE32)
    return set_error(BOLT_ERR_TYPE, "%s: join key dtype %s not supported (Int32/Int64/Date32)", what,
                     dtype_name(c->dtype));
  if (c->length < 0 || c->length > 0xFFFFFFFFll)
    return set_error(BOLT_ERR_PLAN, "%s: %lld rows exceeds the u32 row-id limit", what, (long long)c->length);
  return BOLT_OK;
}

extern "C" int bolt_join_build(bolt_stream_t stream, const bolt_column_t* build_key, const bolt_column_t* carry,
                               bolt_join_table_t** out) {
  BOLT_CHECK_DEVICE();
  clear_error();
  *out = nullptr;
  hipStream_t s = (hipStream_t)stream;
  BOLT_TRY(check_key_col(build_key, "bolt_join_build"));
  int32_t carry_dt = 0;
  if (carry) {
    if (dtype_size(carry->dtype) != 4)
      return set_error(BOLT_ERR_TYPE, "bolt_join_build: carried column must be 4 bytes wide, got %s",
                       dtype_name(carry->dtype));
    if (carry->length != build_key->length)
      return set_error(BOLT_ERR_PLAN, "bolt_join_build: carried column length differs from the key");
    if (carry->validity)
      return set_error(BOLT_ERR_UNSUPPORTED, "bolt_join_build: carried column must not have NULLs");
    carry_dt = carry->dtype;
  }
  int64_t n = build_key->length;
  // capacity: next power of two >= n / load factor
  double load = 0.15;
  if (const char* env = getenv("BOLT_JOIN_LOAD_FACTOR")) {
    double v = atof(env);
    if (v > 0.01 && v <= 0.9) load = v;
  }
  uint64_t cap = 64;
  while ((double)cap * load < (double)n) cap <<= 1;
  size_t bytes = (size_t)cap * sizeof(uint4);
  if (bytes > join_table_cap_bytes())
    return set_error(BOLT_DECLINE, "hash join: build side of %lld rows needs a %zu-byte table, over the cap",
                     (long long)n, bytes);
  // Optional Bloom filter, 16 bits per key (BOLT_JOIN_BLOOM=1).  Off by
  // default: a probe is bound by the rate of random 16-byte requests, and a
  // filter lookup costs about as much as a table probe (measured on the 16M x
  // 256M fused probe: 7.2 ms with the filter, 5.9 ms without, half the probes
  // missing), so it only pays where most probes miss.
  uint64_t bblocks = 0;
  {
    const char* env = getenv("BOLT_JOIN_BLOOM");
    if (env && env[0] == '1' && n > 0) {
      bblocks = 64;
      while (bblocks * 8 < (uint64_t)n) bblocks <<= 1;
    }
  }
  auto* t = new bolt_join_table();
  t->mask = cap - 1;
  t->n_build = n;
  t->key_dtype = build_key->dtype == BOLT_I64 ? BOLT_I64 : BOLT_I32;
  t->carry_dtype = carry_dt;
  t->stream = s;
  int st = dev_alloc(reinterpret_cast<void**>(&t->tbl), bytes, s);
  if (st == BOLT_OK) st = dev_alloc(reinterpret_cast<void**>(&t->flags), 64, s);
  if (st == BOLT_OK && bblocks) {
    st = dev_alloc(reinterpret_cast<void**>(&t->bloom), bblocks * sizeof(uint4), s);
    t->bloom_mask = bblocks - 1;
    if (st == BOLT_OK) st = fill_u32(t->bloom, 0, bblocks * 4, s);
  }
  if (st != BOLT_OK) {
    bolt_join_table_release(t);
    return st;
  }
  const uint32_t* cv = carry ? static_cast<const uint32_t*>(carry->data) : nullptr;
  // region build for tables of 2^20 .. kMaxRegions * 4096 slots (BOLT_JOIN_BUILD=atomic: global claims)
  const char* how = getenv("BOLT_JOIN_BUILD");
  const bool region = n > 0 && cap >= (1ull << 20) && (cap >> kRegBits) <= (uint64_t)kMaxRegions &&
                      !(how && strcmp(how, "atomic") == 0);
  // state words must start at 0: the region build writes every slot, the other path fills the table first
  if ((!region && (st = fill_u32(t->tbl, 0, bytes / 4, s)) != BOLT_OK) || (st = fill_u32(t->flags, 0, 16, s)) != BOLT_OK) {
    bolt_join_table_release(t);
    return st;
  }
  if (region) {
    const int n_regions = (int)(cap >> kRegBits);
    const int64_t per = (n + kRegGrid - 1) / kRegGrid;
    DevBuf<uint32_t> counts, totals, rbase, oval;
    DevBuf<uint4> brec;
    DevBuf<int64_t> okey;
    DevBuf<unsigned long long> ocount;
    st = counts.alloc((size_t)kRegGrid * n_regions, s);
    if (st == BOLT_OK) st = totals.alloc((size_t)n_regions, s);
    if (st == BOLT_OK) st = rbase.alloc((size_t)n_regions + 1, s);
    if (st == BOLT_OK) st = brec.alloc((size_t)n, s);
    if (st == BOLT_OK) st = okey.alloc((size_t)n, s);
    if (st == BOLT_OK) st = oval.alloc((size_t)n, s);
    if (st == BOLT_OK) st = ocount.alloc(1, s);
    if (st == BOLT_OK) st = fill_u32(ocount.ptr, 0, 2, s);
    if (st != BOLT_OK) {
      bolt_join_table_release(t);
      return st;
    }
    const size_t lds = (size_t)n_regions * 4;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(join_region_hist_kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess)
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(join_region_scatter_kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess) {
      join_region_hist_kernel<<<kRegGrid, kRegBlock, lds, s>>>(build_key->data, t->key_dtype, build_key->validity, n, per,
                                                               t->mask, n_regions, counts.ptr);
      join_region_offsets_kernel<<<(n_regions + 255) / 256, 256, 0, s>>>(counts.ptr, kRegGrid, n_regions, totals.ptr);
      join_region_scan_kernel<<<1, kRegBlock, 0, s>>>(totals.ptr, n_regions, rbase.ptr);
      join_region_scatter_kernel<<<kRegGrid, kRegBlock, lds, s>>>(build_key->data, t->key_dtype, build_key->validity,
                                                                  cv, n, per, t->mask, n_regions, counts.ptr, rbase.ptr,
                                                                  brec.ptr);
      join_region_build_kernel<<<(unsigned)n_regions, kRegBuildBlock, 0, s>>>(brec.ptr, rbase.ptr, t->mask, t->tbl,
                                                                              okey.ptr, oval.ptr, ocount.ptr, t->flags);
      join_overflow_insert_kernel<<<64, kBuildBlock, 0, s>>>(t->tbl, t->mask, okey.ptr, oval.ptr, ocount.ptr,
                                                             t->flags + 1);
      join_overflow_dup_kernel<<<64, kBuildBlock, 0, s>>>(t->tbl, t->mask, okey.ptr, ocount.ptr, t->flags);
      e = hipGetLastError();
    }
    if (e != hipSuccess) {
      bolt_join_table_release(t);
      return set_error(BOLT_ERR_DEVICE, "hash join region build failed: %s", hipGetErrorString(e));
    }
  }
  if (n > 0) {
    unsigned grid = grid_for(n, kBuildBlock, 4, 8192);
    if (!region) {
      join_build_kernel<<<grid, kBuildBlock, 0, s>>>(t->tbl, t->mask, build_key->data, t->key_dtype,
                                                     build_key->validity, cv, n, t->flags + 1);
      join_dup_scan_kernel<<<grid_for((int64_t)t->mask + 1, kBuildBlock, 4, 8192), kBuildBlock, 0, s>>>(
          t->tbl, t->mask, t->flags);
    }
    if (t->bloom)
      bloom_build_kernel<<<grid, kBuildBlock, 0, s>>>(t->bloom, t->bloom_mask, build_key->data, t->key_dtype,
                                                      build_key->validity, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      bolt_join_table_release(t);
      return set_error(BOLT_ERR_DEVICE, "hash join build launch failed: %s", hipGetErrorString(e));
    }
  }
  if (hipEventCreateWithFlags(&t->built, hipEventDisableTiming) != hipSuccess ||
      hipEventRecord(t->built, s) != hipSuccess) {
    bolt_join_table_release(t);
    return set_error(BOLT_ERR_DEVICE, "hash join build: completion event");
  }
  *out = t;
  return BOLT_OK;
}

// A probe on a stream other than the build's waits for the build to finish
// (on the device: the host does not block).
static int wait_built(const bolt_join_table_t* t, hipStream_t s) {
  if (t->built && s != t->stream) BOLT_HIP(hipStreamWaitEvent(s, t->built, 0));
  return BOLT_OK;
}

extern "C" int bolt_join_table_unique(const bolt_join_table_t* t, int32_t* unique) {
  BOLT_CHECK_DEVICE();
  uint32_t h[2] = {0, 0};
  BOLT_HIP(hipMemcpyAsync(h, t->flags, 8, hipMemcpyDeviceToHost, t->stream));
  BOLT_HIP(hipStreamSynchronize(t->stream));
  if (h[1] != 0) return set_error(BOLT_ERR_OTHER, "hash join build lost %u rows (table full)", h[1]);
  *unique = h[0] == 0 ? 1 : 0;
  return BOLT_OK;
}

extern "C" void bolt_join_table_release(bolt_join_table_t* t) {
  if (!t) return;
  if (t->tbl) dev_free(t->tbl, t->stream);
  if (t->flags) dev_free(t->flags, t->stream);
  if (t->bloom) dev_free(t->bloom, t->stream);
  if (t->built) (void)hipEventDestroy(t->built);
  delete t;
}

static int check_probe_args(const bolt_join_table_t* t, const bolt_column_t* probe_cols, int32_t n_probe_cols,
                            int32_t key_col, const char* what) {
  if (t == nullptr) return set_error(BOLT_ERR_PLAN, "%s: table is NULL", what);
  if (probe_cols == nullptr || key_col < 0 || key_col >= n_probe_cols)
    return set_error(BOLT_ERR_PLAN, "%s: probe key column %d out of range (%d columns)", what, key_col,
                     n_probe_cols);
  BOLT_TRY(check_key_col(&probe_cols[key_col], what));
  int64_t n = probe_cols[key_col].length;
  for (int i = 0; i < n_probe_cols; i++)
    if (probe_cols[i].length != n)
      return set_error(BOLT_ERR_PLAN, "%s: probe column %d has %lld rows, key has %lld", what, i,
                       (long long)probe_cols[i].length, (long long)n);
  return BOLT_OK;
}

template <typename KeyT, int NT>
static int launch_probe_agg_nt(int na, const bolt_join_table_t* t, const bolt_column_t& key, const DevFilter& f,
                               const AggList& chunk, unsigned grid, Acc* parts, hipStream_t s) {
  const KeyT* kp = static_cast<const KeyT*>(key.data);
#define PROBE_AGG_CASE(NA)                                                                              \
  case NA:                                                                                              \
    join_probe_agg_kernel<KeyT, NT, NA><<<grid, kProbeBlock, 0, s>>>(t->tbl, t->mask, t->flags, kp,        \
                                                                     key.validity, key.length, f, chunk, \
                                                                     t->carry_dtype, parts, t->bloom,   \
                                                                     t->bloom_mask);                    \
    break;
  switch (na) {
    PROBE_AGG_CASE(1)
    PROBE_AGG_CASE(2)
    PROBE_AGG_CASE(3)
    PROBE_AGG_CASE(4)
    default: return set_error(BOLT_ERR_OTHER, "join aggregate: bad chunk size %d", na);
  }
#undef PROBE_AGG_CASE
  BOLT_HIP(hipGetLastError());
  return BOLT_OK;
}

template <typename KeyT>
static int launch_probe_agg(int na, const bolt_join_table_t* t, const bolt_column_t& key, const DevFilter& f,
                            const AggList& chunk, unsigned grid, Acc* parts, hipStream_t s) {
  switch (f.n) {
    case 0: return launch_probe_agg_nt<KeyT, 0>(na, t, key, f, chunk, grid, parts, s);
    case 1: return launch_probe_agg_nt<KeyT, 1>(na, t, key, f, chunk, grid, parts, s);
    case 2: return launch_probe_agg_nt<KeyT, 2>(na, t, key, f, chunk, grid, parts, s);
    default: return launch_probe_agg_nt<KeyT, -1>(na, t, key, f, chunk, grid, parts, s);
  }
}

// The radix-partitioned probe + aggregate (join_part.hip).
namespace bolt {
bool join_partitioned_wanted(const bolt_join_table_t* t, int64_t n_probe, const AggList& aggs);
int join_probe_aggregate_partitioned(hipStream_t s, const bolt_join_table_t* t, const bolt_column_t& key,
                                     const DevFilter& f, const AggList& all, Acc* outs);
}  // namespace bolt

// Grid for the probe kernels; BOLT_JOIN_PROBE_GRID overrides (tuning).
static unsigned probe_grid(int64_t n) {
  const char* env = getenv("BOLT_JOIN_PROBE_GRID");
  if (env && *env) return (unsigned)strtoul(env, nullptr, 10);
  return grid_for(n, kProbeBlock, 4, 256 * 32);
}

extern "C" int bolt_join_probe_aggregate(bolt_stream_t stream, const bolt_join_table_t* t,
                                         const bolt_column_t* probe_cols, int32_t n_probe_cols, int32_t key_col,
                                         const bolt_filter_t* filter, const bolt_agg_t* aggs, int32_t n_aggs,
                                         bolt_scalar_t* results) {
  BOLT_CHECK_DEVICE();
  clear_error();
  hipStream_t s = (hipStream_t)stream;
  BOLT_TRY(check_probe_args(t, probe_cols, n_probe_cols, key_col, "bolt_join_probe_aggregate"));
  BOLT_TRY(wait_built(t, s));
  const bolt_column_t& key = probe_cols[key_col];
  DevFilter f;
  BOLT_TRY(make_dev_filter(filter, probe_cols, n_probe_cols, key.length, &f));
  AggList all;
  BOLT_TRY(make_agg_list(aggs, n_aggs, probe_cols, n_probe_cols, t->carry_dtype, &all));
  if (n_aggs == 0) return BOLT_OK;
  unsigned grid = probe_grid(key.length);
  DevBuf<Acc> parts, outs;
  BOLT_TRY(outs.alloc((size_t)n_aggs, s));
  bool done = false;
  if (join_partitioned_wanted(t, key.length, all)) {
    BOLT_TRY(join_probe_aggregate_partitioned(s, t, key, f, all, outs.ptr));
    done = true;
  }
  if (!done && jit_enabled() && key.length > 0) {
    JitSpec sp;
    JitLits lits;
    JitCols cols;
    hipFunction_t fn;
    const bool spec_ok = jit_spec_from(f, all, true, key.dtype == BOLT_I64 ? BOLT_I64 : BOLT_I32,
                                       key.validity != nullptr, t->carry_dtype, &sp, &lits, &cols) == BOLT_OK;
    sp.bloom = t->bloom != nullptr;
    if (spec_ok && jit_get_keyed(jit_row_agg_key(sp), [&] { return jit_row_agg_source(sp); }, &fn) == BOLT_OK) {
      BOLT_TRY(parts.alloc((size_t)grid * n_aggs, s));
      prof_begin(s);
      BOLT_TRY(jit_launch_row_agg(fn, grid, s, t->tbl, t->mask, t->flags, key.length, key.data, key.validity, lits, cols,
                                  parts.ptr, t->bloom, t->bloom_mask));
      prof_end(s);
      for (int first = 0; first < n_aggs; first += 4) {
        int na = n_aggs - first < 4 ? n_aggs - first : 4;
        BOLT_TRY(launch_finish(na, parts.ptr, (int)grid, n_aggs, first, agg_chunk(all, first, na), outs.ptr + first, s));
      }
      done = true;
    }
    clear_error();
  }
  if (!done) {
    int n_chunks = (n_aggs + 3) / 4;
    BOLT_TRY(parts.alloc((size_t)grid * 4 * n_chunks, s));
    prof_begin(s);
    for (int c = 0; c < n_chunks; c++) {
      int first = c * 4, na = n_aggs - first < 4 ? n_aggs - first : 4;
      AggList chunk = agg_chunk(all, first, na);
      Acc* p = parts.ptr + (size_t)c * grid * 4;
      if (key.dtype == BOLT_I64) BOLT_TRY(launch_probe_agg<int64_t>(na, t, key, f, chunk, grid, p, s));
      else BOLT_TRY(launch_probe_agg<int32_t>(na, t, key, f, chunk, grid, p, s));
      BOLT_TRY(launch_finish(na, p, (int)grid, na, 0, chunk, outs.ptr + first, s));
    }
    prof_end(s);
  }
  Acc host[kMaxAggs];
  uint32_t fl[2];
  BOLT_HIP(hipMemcpyAsync(host, outs.ptr, sizeof(Acc) * n_aggs, hipMemcpyDeviceToHost, s));
  BOLT_HIP(hipMemcpyAsync(fl, t->flags, 8, hipMemcpyDeviceToHost, s));
  BOLT_HIP(hipStreamSynchronize(s));
  if (fl[1] != 0) return set_error(BOLT_ERR_OTHER, "hash join build lost %u rows (table full)", fl[1]);
  for (int i = 0; i < n_aggs; i++) BOLT_TRY(acc_finalize(host[i], all.a[i].op, all.a[i].dtype, &results[i]));
  return BOLT_OK;
}

// Pair-emitting probe of `key` (rows passing `f`) into (out_probe, out_build);
// *total receives the full pair count, also past `capacity`.  `preserve` and
// `unmatched` as for join_probe_emit_kernel.  Synchronises.
static int probe_emit(hipStream_t s, const bolt_join_table_t* t, const bolt_column_t& key, const DevFilter& f,
                      bool preserve, uint8_t* unmatched, uint32_t* out_probe, uint32_t* out_build,
                      int64_t capacity, int64_t* total_out, const KeyCols* verify_probe = nullptr,
                      const KeyCols* verify_build = nullptr) {
  KeyCols pk, bk;
  memset(&pk, 0, sizeof pk);
  memset(&bk, 0, sizeof bk);
  const bool verify = verify_probe != nullptr;
  if (verify) {
    pk = *verify_probe;
    bk = *verify_build;
  }
  DevBuf<unsigned long long> counter;
  DevBuf<uint64_t> status;
  DevBuf<uint32_t> misc;  // [0] ticket, [1] look-back error
  BOLT_TRY(counter.alloc(1, s));
  BOLT_TRY(fill_u32(counter.ptr, 0, 2, s));
  BOLT_TRY(misc.alloc(4, s));
  BOLT_TRY(fill_u32(misc.ptr, 0, 4, s));
  if (key.length > 0) {
    const int64_t tiles = (key.length + kEmitTile - 1) / kEmitTile;
    BOLT_TRY(status.alloc((size_t)tiles, s));
    BOLT_TRY(fill_u32(status.ptr, 0, (size_t)tiles * 2, s));
#define EMIT_LAUNCH(KT, PRESERVE, VERIFY)                                                                        \
  join_probe_emit_kernel<KT, PRESERVE, VERIFY><<<(unsigned)tiles, kProbeBlock, 0, s>>>(                          \
      t->tbl, t->mask, t->flags, static_cast<const KT*>(key.data), key.validity, key.length, f, status.ptr,      \
      misc.ptr, misc.ptr + 1, counter.ptr, capacity, out_probe, out_build, t->bloom, t->bloom_mask, unmatched,  \
      pk, bk)
    if (verify && key.dtype != BOLT_I64) return set_error(BOLT_ERR_OTHER, "hash join: hashed keys must be Int64");
    if (verify) {
      if (preserve) EMIT_LAUNCH(int64_t, true, true);
      else EMIT_LAUNCH(int64_t, false, true);
    } else if (key.dtype == BOLT_I64) {
      if (preserve) EMIT_LAUNCH(int64_t, true, false);
      else EMIT_LAUNCH(int64_t, false, false);
    } else {
      if (preserve) EMIT_LAUNCH(int32_t, true, false);
      else EMIT_LAUNCH(int32_t, false, false);
    }
#undef EMIT_LAUNCH
    BOLT_HIP(hipGetLastError());
  }
  unsigned long long total = 0;
  uint32_t fl[2], hm[2];
  BOLT_HIP(hipMemcpyAsync(&total, counter.ptr, 8, hipMemcpyDeviceToHost, s));
  BOLT_HIP(hipMemcpyAsync(fl, t->flags, 8, hipMemcpyDeviceToHost, s));
  BOLT_HIP(hipMemcpyAsync(hm, misc.ptr, 8, hipMemcpyDeviceToHost, s));
  BOLT_HIP(hipStreamSynchronize(s));
  if (fl[1] != 0) return set_error(BOLT_ERR_OTHER, "hash join build lost %u rows (table full)", fl[1]);
  if (hm[1] != 0) return set_error(BOLT_ERR_OTHER, "hash join probe: output offset look-back timed out");
  *total_out = (int64_t)total;
  return BOLT_OK;
}

extern "C" int bolt_join_probe(bolt_stream_t stream, const bolt_join_table_t* t, const bolt_column_t* probe_cols,
                               int32_t n_probe_cols, int32_t key_col, const bolt_filter_t* filter,
                               uint32_t* out_probe_rows, uint32_t* out_build_rows, int64_t capacity,
                               int64_t* n_out) {
  BOLT_CHECK_DEVICE();
  clear_error();
  hipStream_t s = (hipStream_t)stream;
  *n_out = 0;
  BOLT_TRY(check_probe_args(t, probe_cols, n_probe_cols, key_col, "bolt_join_probe"));
  BOLT_TRY(wait_built(t, s));
  if (t->carry_dtype != 0)
    return set_error(BOLT_ERR_PLAN, "bolt_join_probe: table stores a carried column, not build row ids");
  const bolt_column_t& key = probe_cols[key_col];
  DevFilter f;
  BOLT_TRY(make_dev_filter(filter, probe_cols, n_probe_cols, key.length, &f));
  int64_t total = 0;
  BOLT_TRY(probe_emit(s, t, key, f, false, nullptr, out_probe_rows, out_build_rows, capacity, &total));
  *n_out = total;
  if (total > capacity)
    return set_error(BOLT_DECLINE, "hash join: %lld matches exceed the output capacity %lld", (long long)total,
                     (long long)capacity);
  return BOLT_OK;
}

extern "C" int bolt_inner_join(bolt_stream_t stream, const bolt_column_t* left_key, const bolt_column_t* right_key,
                               uint32_t* out_left_rows, uint32_t* out_right_rows, int64_t capacity, int64_t* n_out) {
  BOLT_CHECK_DEVICE();
  clear_error();
  *n_out = 0;
  BOLT_TRY(check_key_col(left_key, "bolt_inner_join"));
  BOLT_TRY(check_key_col(right_key, "bolt_inner_join"));
  if (left_key->length == 0 || right_key->length == 0) return BOLT_OK;
  // the smaller side builds (src/exec/join.rs:192)
  bool build_is_left = left_key->length <= right_key->length;
  const bolt_column_t* b = build_is_left ? left_key : right_key;
  const bolt_column_t* p = build_is_left ? right_key : left_key;
  bolt_join_table_t* t = nullptr;
  BOLT_TRY(bolt_join_build(stream, b, nullptr, &t));
  int st = build_is_left ? bolt_join_probe(stream, t, p, 1, 0, nullptr, out_right_rows, out_left_rows, capacity, n_out)
                         : bolt_join_probe(stream, t, p, 1, 0, nullptr, out_left_rows, out_right_rows, capacity, n_out);
  bolt_join_table_release(t);
  return st;
}

// Key columns of one side as a probe / build key: a plain Int32 / Int64 /
// Date32 column is used in place; anything else is encoded into `enc` /
// `enc_valid` by join_key_kernel.
static bool is_int32_key(int dt) { return dt == BOLT_I32 || dt == BOLT_DATE32; }

// Encoding mode of a key shape (KeyCols::mode).
static int key_mode(const bolt_column_t* cols, int32_t n_keys) {
  if (n_keys == 1) return 0;
  if (n_keys == 2 && is_int32_key(cols[0].dtype) && is_int32_key(cols[1].dtype)) return 1;
  return 2;
}

static void fill_key_cols(const bolt_column_t* cols, int32_t n_keys, KeyCols* kc) {
  memset(kc, 0, sizeof *kc);
  kc->n_cols = n_keys;
  kc->mode = key_mode(cols, n_keys);
  kc->hash_mask = ~0ull;
  if (const char* env = getenv("BOLT_JOIN_HASH_BITS")) {  // tests: force collisions of hashed keys
    const int bits = atoi(env);
    if (bits > 0 && bits < 64) kc->hash_mask = (1ull << bits) - 1;
  }
  for (int c = 0; c < n_keys; c++) {
    kc->d[c] = cols[c].data;
    kc->v[c] = cols[c].validity;
    kc->dt[c] = cols[c].dtype;
  }
}

static int side_key(hipStream_t s, const bolt_column_t* cols, int32_t n_keys, DevBuf<int64_t>& enc,
                    DevBuf<uint64_t>& enc_valid, bolt_column_t* key) {
  const int64_t n = cols[0].length;
  const int dt = cols[0].dtype;
  if (n_keys == 1 && (dt == BOLT_I32 || dt == BOLT_I64 || dt == BOLT_DATE32)) {
    *key = cols[0];
    return BOLT_OK;
  }
  KeyCols kc;
  fill_key_cols(cols, n_keys, &kc);
  bool any_nulls = false;
  for (int c = 0; c < n_keys; c++) any_nulls |= cols[c].validity != nullptr;
  BOLT_TRY(enc.alloc((size_t)(n > 0 ? n : 1), s));
  if (any_nulls) BOLT_TRY(enc_valid.alloc((size_t)((n + 63) / 64 + 1), s));
  if (n > 0) {
    join_key_kernel<<<grid_for(n, 256, 4), 256, 0, s>>>(kc, n, enc.ptr, any_nulls ? enc_valid.ptr : nullptr);
    BOLT_HIP(hipGetLastError());
  }
  memset(key, 0, sizeof *key);
  key->data = enc.ptr;
  key->validity = any_nulls ? reinterpret_cast<const uint8_t*>(enc_valid.ptr) : nullptr;
  key->length = n;
  key->dtype = BOLT_I64;
  return BOLT_OK;
}

static int check_join_side(const bolt_column_t* cols, int32_t n_keys, const char* side) {
  if (cols == nullptr) return set_error(BOLT_ERR_PLAN, "bolt_join: %s key columns are NULL", side);
  for (int c = 0; c < n_keys; c++) {
    const bolt_column_t& k = cols[c];
    if (k.length != cols[0].length)
      return set_error(BOLT_ERR_PLAN, "bolt_join: %s key column %d has %lld rows, column 0 has %lld", side, c,
                       (long long)k.length, (long long)cols[0].length);
    if (k.length < 0 || k.length >= (int64_t)kNullRow)
      return set_error(BOLT_ERR_PLAN, "bolt_join: %s side of %lld rows exceeds the u32 row-id limit", side,
                       (long long)k.length);
    if (k.length > 0 && k.data == nullptr) return set_error(BOLT_ERR_PLAN, "bolt_join: %s key column %d has no data", side, c);
  }
  return BOLT_OK;
}

extern "C" int bolt_join(bolt_stream_t stream, const bolt_column_t* left_keys, const bolt_column_t* right_keys,
                         int32_t n_keys, int32_t join_type, uint32_t* out_left_rows, uint32_t* out_right_rows,
                         int64_t capacity, int64_t* n_out) {
  BOLT_CHECK_DEVICE();
  clear_error();
  hipStream_t s = (hipStream_t)stream;
  *n_out = 0;
  if (join_type < BOLT_JOIN_INNER || join_type > BOLT_JOIN_FULL)
    return set_error(BOLT_ERR_PLAN, "bolt_join: unknown join type %d", join_type);
  if (n_keys < 1)
    return set_error(BOLT_ERR_PLAN, "bolt_join: at least one equi-join key is required (src/exec/join.rs:290)");
  BOLT_TRY(check_join_side(left_keys, n_keys, "left"));
  BOLT_TRY(check_join_side(right_keys, n_keys, "right"));
  for (int c = 0; c < n_keys; c++) {
    const int dl = left_keys[c].dtype, dr = right_keys[c].dtype;
    if (dl != dr)
      return set_error(BOLT_ERR_TYPE, "bolt_join: key %d is %s on the left and %s on the right", c, dtype_name(dl),
                       dtype_name(dr));
    const bool ok1 = dl == BOLT_I32 || dl == BOLT_I64 || dl == BOLT_DATE32 || dl == BOLT_F32 || dl == BOLT_F64 ||
                     dl == BOLT_BOOL;
    const bool ok_multi = dl == BOLT_I32 || dl == BOLT_I64 || dl == BOLT_DATE32;
    if (n_keys == 1 && !ok1)
      return set_error(BOLT_ERR_TYPE, "bolt_join: join key dtype %s not supported", dtype_name(dl));
    if (n_keys > 1 && !ok_multi)
      return set_error(BOLT_ERR_UNSUPPORTED,
                       "bolt_join: multi-column keys are Int32 / Int64 / Date32 columns (key %d is %s)", c,
                       dtype_name(dl));
  }
  if (n_keys > kMaxJoinKeys)
    return set_error(BOLT_ERR_UNSUPPORTED, "bolt_join: %d key columns (at most %d)", n_keys, kMaxJoinKeys);
  const int64_t nl = left_keys[0].length, nr = right_keys[0].length;
  // Build side: the smaller for INNER (src/exec/join.rs:192); the side that is
  // not preserved for LEFT / RIGHT, the right for FULL (join.rs:305).
  bool build_is_left;
  switch (join_type) {
    case BOLT_JOIN_INNER: build_is_left = nl <= nr; break;
    case BOLT_JOIN_RIGHT: build_is_left = true; break;
    default: build_is_left = false; break;
  }
  const bool preserve = join_type != BOLT_JOIN_INNER;
  const bool full = join_type == BOLT_JOIN_FULL;
  if (join_type == BOLT_JOIN_INNER && (nl == 0 || nr == 0)) return BOLT_OK;
  const bolt_column_t* bcols = build_is_left ? left_keys : right_keys;
  const bolt_column_t* pcols = build_is_left ? right_keys : left_keys;
  uint32_t* out_build = build_is_left ? out_left_rows : out_right_rows;
  uint32_t* out_probe = build_is_left ? out_right_rows : out_left_rows;
  if ((out_left_rows == nullptr || out_right_rows == nullptr) && capacity > 0)
    return set_error(BOLT_ERR_PLAN, "bolt_join: NULL output arrays");

  DevBuf<int64_t> benc, penc;
  DevBuf<uint64_t> bval, pval;
  bolt_column_t bkey, pkey;
  BOLT_TRY(side_key(s, bcols, n_keys, benc, bval, &bkey));
  BOLT_TRY(side_key(s, pcols, n_keys, penc, pval, &pkey));
  bolt_join_table_t* t = nullptr;
  BOLT_TRY(bolt_join_build(stream, &bkey, nullptr, &t));
  const int64_t n_build = bkey.length;
  DevBuf<uint8_t> unmatched;
  int st = BOLT_OK;
  if (full && n_build > 0) {
    st = unmatched.alloc((size_t)((n_build + 3) / 4 * 4), s);
    if (st == BOLT_OK) st = fill_u32(unmatched.ptr, 0x01010101u, (size_t)((n_build + 3) / 4), s);
  }
  DevFilter f;
  memset(&f, 0, sizeof f);
  int64_t total = 0;
  KeyCols pk, bk;
  const bool verify = key_mode(bcols, n_keys) == 2;
  if (verify) {
    fill_key_cols(pcols, n_keys, &pk);
    fill_key_cols(bcols, n_keys, &bk);
  }
  if (st == BOLT_OK)
    st = probe_emit(s, t, pkey, f, preserve, full ? unmatched.ptr : nullptr, out_probe, out_build, capacity, &total,
                    verify ? &pk : nullptr, verify ? &bk : nullptr);
  bolt_join_table_release(t);
  BOLT_TRY(st);
  if (full && n_build > 0) {
    DevBuf<uint32_t> rank;
    BOLT_TRY(rank.alloc((size_t)n_build, s));
    int64_t n_unmatched = 0;
    BOLT_TRY(bolt_prefix_scan_mask(stream, unmatched.ptr, n_build, rank.ptr, &n_unmatched));
    if (n_unmatched > 0 && total < capacity) {
      join_unmatched_kernel<<<grid_for(n_build, 256, 4), 256, 0, s>>>(unmatched.ptr, rank.ptr, n_build, total,
                                                                      capacity, out_probe, out_build);
      BOLT_HIP(hipGetLastError());
      BOLT_HIP(hipStreamSynchronize(s));
    }
    total += n_unmatched;
  }
  *n_out = total;
  if (total > capacity)
    return set_error(BOLT_DECLINE, "bolt_join: %lld output rows exceed the capacity %lld", (long long)total,
                     (long long)capacity);
  return BOLT_OK;
}
