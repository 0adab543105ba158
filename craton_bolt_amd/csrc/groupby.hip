// groupby.hip — hash GROUP BY.
//
// Replaces execute_groupby (src/exec/groupby.rs:295) with its cascade of
// Tier-1 LDS kernels (src/jit/shmem_*_kernel.rs, groupby_shmem_*_exec.rs),
// Tier-2 partitioned kernels (src/jit/partition_*.rs, groupby_tier2_*.rs) and
// the classic global-atomic path (src/jit/hash_kernels.rs:246,764,
// src/jit/float_atomics.rs:139), which launch one kernel per aggregate and
// round-trip to the host between passes.  Here every aggregate of the query
// is updated by the same pass:
//
//   pass A  LDS tier.  Each workgroup keeps an open-addressing table of up to
//           a few thousand groups in LDS (keys + one 64-bit word per
//           accumulator field) and updates it with LDS atomics.  Rows whose
//           key finds no LDS slot within a short probe are spilled (row ids
//           appended with one atomic per wave).  At the end the workgroup
//           appends its groups as partial records.  A HyperLogLog sketch of
//           every partial key and spilled key estimates the group count.
//   pass B  HBM tier.  A global open-addressing table sized from the sketch
//           (retried 4x larger if it fills) absorbs the partial records and
//           the spilled rows with global atomics.
//   pass C  Extraction: occupied slots are compacted with decoupled look-back
//           into the output columns, finalising each aggregate on the way.
//
// Accumulators are exact where the reference's are: integer SUM keeps a
// 128-bit total per group through a carry-propagating pair of 64-bit
// atomics (overflow of i64 is reported at finalize, like checked_group_sum,
// src/exec/groupby.rs:1775); Decimal128 SUM keeps 192 bits; MIN / MAX of
// every type use native unsigned 64-bit atomic min / max on an
// order-preserving image (floats in the NaN-largest order of
// src/jit/float_atomics.rs:27-62), so no CAS loop is needed.  Float SUM uses
// hardware f64 atomic adds (order-dependent, like the reference's atomics).
//
// Keys (src/exec/groupby_common.rs:352): one Int32 / Int64 / Float32 /
// Float64 / Date32 column, or two 32-bit columns packed (a << 32) | b; float
// keys canonicalise -0.0 and NaN (groupby_common.rs:302).  A single-column
// key with NULLs forms one NULL group (groupby.rs:528-606); two-column keys
// drop rows with a NULL key, as the reference does.  There is no sentinel
// restriction: a key equal to the empty marker and the NULL group live in two
// dedicated slots after the table.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <map>
#include <mutex>
#include <vector>

#include "groupby_dev.h"
#include "jit.h"
#include "scan_common.h"

namespace bolt {

int radix_sort_u64_pairs(hipStream_t s, const uint64_t* keys_in, const uint32_t* vals_in, int64_t n,
                         uint64_t* keys_out, uint32_t* vals_out, int key_bits, uint64_t in_xor = 0,
                         uint64_t out_xor = 0);

// ---- pass A: LDS tier ------------------------------------------------------------

__global__ __launch_bounds__(kGbBlock) void gb_lds_kernel(KeySpec ks, DevFilter f, GPlan P, int64_t n, int S,
                                                         PartOut out) {
  extern __shared__ unsigned long long sm[];
  const int SS = S + 2;  // + NULL group + empty-marker key
  unsigned long long* skeys = sm;
  unsigned long long* swords = sm + SS;
  uint32_t* shll = reinterpret_cast<uint32_t*>(swords + (size_t)P.words * SS);
  uint32_t* sused = shll + kHllRegs;  // [0..1] specials used, [2] spill count, [3] occupied slots
  for (int i = threadIdx.x; i < SS; i += kGbBlock) skeys[i] = kEmpty;
  for (int w = 0; w < P.words; w++) {
    unsigned long long id = word_identity(P, w);
    for (int i = threadIdx.x; i < SS; i += kGbBlock) swords[(size_t)w * SS + i] = id;
  }
  for (int i = threadIdx.x; i < kHllRegs; i += kGbBlock) shll[i] = 0;
  if (threadIdx.x < 4) sused[threadIdx.x] = 0;
  __syncthreads();

  const int64_t stride = (int64_t)gridDim.x * kGbBlock;
  const int64_t end = ((n + kGbBlock - 1) / kGbBlock) * kGbBlock;  // wave-uniform trip count
  for (int64_t row = (int64_t)blockIdx.x * kGbBlock + threadIdx.x; row < end; row += stride) {
    bool spill = false;
    uint64_t key = 0;
    if (row < n && eval_filter(f, row)) {
      int kr = row_key(ks, row, &key);
      int slot = -1;
      if (kr == 2) {
        slot = S;
        sused[0] = 1;
      } else if (kr == 1) {
        if (key == kEmpty) {
          slot = S + 1;
          sused[1] = 1;
        } else {
          slot = gb_lds_slot(skeys, S, key, &sused[3]);
          spill = slot < 0;
        }
      }
      if (slot >= 0) update_slot(P, swords, SS, (uint32_t)slot, row);
    }
    gb_spill(spill, row, key, sused, shll, out);
  }
  gb_lds_flush(skeys, swords, S, P.words, shll, sused, out);
}

// ---- pass B: global table ------------------------------------------------------------


__global__ void gb_init_kernel(GTable t, GPlan P) {
  const size_t total = t.cap + 2;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    if (i < t.cap) t.keys[i] = kEmpty;
    for (int w = 0; w < P.words; w++) t.words[(size_t)w * total + i] = word_identity(P, w);
  }
  if (blockIdx.x == 0 && threadIdx.x < 2) t.used_special[threadIdx.x] = 0;
}


__global__ __launch_bounds__(256) void gb_merge_parts_kernel(GTable t, GPlan P, const unsigned long long* pkeys,
                                                            const uint8_t* ptag, const unsigned long long* pwords,
                                                            int64_t pcap, int64_t np) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < np; i += stride) {
    int64_t s = gslot(t, ptag[i], pkeys[i]);
    if (s < 0) continue;
    merge_slot(P, t.words, t.cap + 2, (uint32_t)s, pwords + i, (size_t)pcap);
  }
}

// Partial records merged in two levels when the groups are few: each
// workgroup folds its share of the records into an LDS table (LDS atomics),
// then merges that table into the global one, so a group receives one
// global update per merging workgroup instead of one per LDS-tier workgroup
// (the latter all hit the same few words).  Records that find no LDS slot go
// to the global table directly.
__global__ __launch_bounds__(kGbBlock) void gb_merge_lds_kernel(GTable t, GPlan P, const unsigned long long* pkeys,
                                                               const uint8_t* ptag, const unsigned long long* pwords,
                                                               int64_t pcap, int64_t np, int S) {
  extern __shared__ unsigned long long sm[];
  const int SS = S + 2;
  unsigned long long* skeys = sm;
  unsigned long long* swords = sm + SS;
  uint32_t* sused = reinterpret_cast<uint32_t*>(swords + (size_t)P.words * SS);  // [0..1] specials, [2] occupied
  for (int i = threadIdx.x; i < SS; i += kGbBlock) skeys[i] = kEmpty;
  for (int w = 0; w < P.words; w++) {
    const unsigned long long id = word_identity(P, w);
    for (int i = threadIdx.x; i < SS; i += kGbBlock) swords[(size_t)w * SS + i] = id;
  }
  if (threadIdx.x < 3) sused[threadIdx.x] = 0;
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * kGbBlock;
  for (int64_t r = (int64_t)blockIdx.x * kGbBlock + threadIdx.x; r < np; r += stride) {
    const uint8_t tag = ptag[r];
    const unsigned long long key = pkeys[r];
    int slot;
    if (tag == 1) { slot = S; sused[0] = 1; }
    else if (tag == 2) { slot = S + 1; sused[1] = 1; }
    else slot = gb_lds_slot(skeys, S, key, &sused[2]);
    if (slot >= 0) {
      merge_slot(P, swords, SS, (uint32_t)slot, pwords + r, (size_t)pcap);
    } else {
      const int64_t g = gslot(t, tag, key);
      if (g >= 0) merge_slot(P, t.words, t.cap + 2, (uint32_t)g, pwords + r, (size_t)pcap);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < SS; i += kGbBlock) {
    int64_t g = -1;
    if (i < S) {
      if (skeys[i] != kEmpty) g = gslot(t, 0, skeys[i]);
    } else if (sused[i - S]) {
      g = gslot(t, i - S + 1, 0);
    }
    if (g >= 0) merge_slot(P, t.words, t.cap + 2, (uint32_t)g, swords + i, (size_t)SS);
  }
}

// Rows processed straight against the global table: the spilled rows
// (rows != nullptr) or every row of the input (rows == nullptr).
__global__ __launch_bounds__(256) void gb_rows_kernel(GTable t, GPlan P, KeySpec ks, DevFilter f,
                                                     const uint32_t* rows, int64_t nrows) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nrows; i += stride) {
    int64_t row = rows ? (int64_t)rows[i] : i;
    if (!rows && !eval_filter(f, row)) continue;
    uint64_t key = 0;
    int kr = row_key(ks, row, &key);
    if (kr == 0) continue;
    int64_t s = gslot(t, kr == 2 ? 1 : 0, key);
    if (s < 0) continue;
    update_slot(P, t.words, t.cap + 2, (uint32_t)s, row);
  }
}

// The spilled rows, segment by segment: blocks (b, k) for k < K share
// workgroup b's segment.
__global__ __launch_bounds__(256) void gb_rows_seg_kernel(GTable t, GPlan P, KeySpec ks, const uint32_t* spill,
                                                         int64_t per, const uint32_t* seg_count, int K) {
  const int b = blockIdx.x / K, k = blockIdx.x % K;
  const uint32_t cnt = seg_count[b];
  for (uint32_t i = (uint32_t)k * blockDim.x + threadIdx.x; i < cnt; i += (uint32_t)K * blockDim.x) {
    const int64_t row = (int64_t)spill[(int64_t)b * per + i];
    uint64_t key = 0;
    const int kr = row_key(ks, row, &key);
    if (kr == 0) continue;
    const int64_t sl = gslot(t, kr == 2 ? 1 : 0, key);
    if (sl < 0) continue;
    update_slot(P, t.words, t.cap + 2, (uint32_t)sl, row);
  }
}

// Second pass for VAR / STDDEV: M2 = sum over the group of (x - mean)^2, with
// the mean taken from the finished first pass (two-pass variance).
__global__ __launch_bounds__(256) void gb_m2_kernel(GTable t, GPlan P, KeySpec ks, DevFilter f, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const size_t total = t.cap + 2;
  for (int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; row < n; row += stride) {
    if (!eval_filter(f, row)) continue;
    uint64_t key = 0;
    int kr = row_key(ks, row, &key);
    if (kr == 0) continue;
    int64_t s;
    if (kr == 2) s = (int64_t)t.cap;
    else if (key == kEmpty) s = (int64_t)t.cap + 1;
    else {
      uint32_t q = global_find(t.keys, t.cap - 1, key);
      if (q == 0xFFFFFFFFu) continue;
      s = q;
    }
    for (int a = 0; a < P.n; a++) {
      const GAgg& g = P.a[a];
      if (g.kind != K_VAR || !bit_valid(g.validity, row)) continue;
      double v = g.dtype == BOLT_F64 ? static_cast<const double*>(g.data)[row]
               : g.dtype == BOLT_F32 ? (double)static_cast<const float*>(g.data)[row]
               : g.dtype == BOLT_I64 ? (double)static_cast<const int64_t*>(g.data)[row]
                                     : (double)static_cast<const int32_t*>(g.data)[row];
      double sum = as_f64(t.words[(size_t)g.w * total + s]);
      double cnt = (double)t.words[(size_t)g.nw * total + s];
      double d = v - sum / cnt;
      unsafeAtomicAdd(reinterpret_cast<double*>(&t.words[(size_t)(g.w + 1) * total + s]), d * d);
    }
  }
}

// ---- pass C: extraction ---------------------------------------------------------------


constexpr int kExBlock = 256;
constexpr int kExItems = 16;

// ITEMS slots per thread: 16 for large tables, 1 for small ones (more tiles,
// so a table of a few thousand slots is not left to one or two workgroups).
template <int ITEMS>
__global__ __launch_bounds__(kExBlock) void gb_extract_kernel(GTable t, GPlan P, KeySpec ks, OutSpec o,
                                                             uint64_t* __restrict__ status,
                                                             uint32_t* __restrict__ ticket,
                                                             unsigned long long* __restrict__ total_out) {
  __shared__ uint32_t sh_tile;
  __shared__ uint32_t sh_cnt[ITEMS * 4];
  __shared__ uint64_t sh_prefix;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) sh_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  const int64_t tile = sh_tile;
  const uint64_t total = t.cap + 2;
  const uint64_t base = (uint64_t)tile * kExBlock * ITEMS;
  uint32_t keep = 0;
#pragma unroll
  for (int j = 0; j < ITEMS; j++) {
    uint64_t s = base + (uint64_t)j * kExBlock + threadIdx.x;
    bool k = false;
    if (s < t.cap) k = t.keys[s] != kEmpty;
    else if (s < total) k = t.used_special[s - t.cap] != 0;
    keep |= (uint32_t)k << j;
    uint64_t b = __ballot(k);
    if (lane == 0) sh_cnt[j * 4 + wave] = (uint32_t)__popcll(b);
  }
  __syncthreads();
  if (wave == 0) {
    uint32_t c = lane < ITEMS * 4 ? sh_cnt[lane] : 0u;
    uint32_t incl = c;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      uint32_t v = __shfl_up(incl, d);
      if (lane >= d) incl += v;
    }
    uint32_t tile_total = __shfl(incl, 63);
    if (lane < ITEMS * 4) sh_cnt[lane] = incl - c;
    uint64_t excl = lookback_wave(status, tile, tile_total, &t.err[1]);
    if (lane == 0) {
      sh_prefix = excl;
      if (tile == (int64_t)gridDim.x - 1) *total_out = excl + tile_total;
    }
  }
  __syncthreads();
  const uint64_t prefix = sh_prefix;
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  for (int j = 0; j < ITEMS; j++) {
    bool k = (keep >> j) & 1u;
    uint64_t b = __ballot(k);
    if (!k) continue;
    uint64_t s = base + (uint64_t)j * kExBlock + threadIdx.x;
    uint64_t pos = prefix + sh_cnt[j * 4 + wave] + (uint64_t)__popcll(b & below);
    bool null_group = s == t.cap;
    uint64_t packed = s < t.cap ? t.keys[s] : (null_group ? 0x7FFFFFFFFFFFFFFFull : kEmpty);
    o.packed[pos] = packed;
    o.slot_of[pos] = (uint32_t)s;
    if (ks.n_keys == 1) {
      write_key(o.key[0], pos, null_group ? 0 : packed);
      if (o.key[0].valid && !null_group) set_valid(o.key[0].valid, pos);
    } else {
      write_key(o.key[0], pos, packed >> 32);
      write_key(o.key[1], pos, packed & 0xFFFFFFFFull);
    }
    for (int a = 0; a < P.n; a++) finalize_to(t, P.a[a], o.agg[a], total, s, pos);
  }
}

// err[0] = 1 when a valid Decimal128 key does not fit i64 (its high word is
// not the sign extension of its low word).
__global__ void dec_key_fit_kernel(const uint64_t* __restrict__ d, const uint8_t* __restrict__ valid, int64_t n,
                                   uint32_t* __restrict__ err) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    if (bit_valid(valid, i) && d[2 * i + 1] != (uint64_t)((int64_t)d[2 * i] >> 63)) *err = 1u;
}

// The reference synthesises the NULL group of a single nullable key column
// under the packed key i64::MAX; when a genuine key equal to i64::MAX exists
// it runs its valid-flag executor instead, which drops the NULL-key rows
// (src/exec/groupby.rs:553-583).  res[0] = output row of the NULL group
// (preset -1), res[1] = 1 when a genuine i64::MAX key is among the groups.
__global__ void gb_null_clash_kernel(const unsigned long long* __restrict__ packed, const uint32_t* __restrict__ slot_of,
                                     int64_t n, uint64_t cap, int64_t* __restrict__ res) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if (slot_of[i] == (uint32_t)cap) res[0] = i;
    else if (packed[i] == 0x7FFFFFFFFFFFFFFFull) res[1] = 1;
  }
}

struct MoveCols {
  uint8_t* d[kMaxAggs + 2];
  uint32_t* v[kMaxAggs + 2];
  int32_t w[kMaxAggs + 2];
  int32_t n;
};

// Output row `src` (the last) moves to row `dst`, in every column, its
// validity bit and the packed key / slot arrays the ordering reads.
__global__ void gb_move_row_kernel(MoveCols mc, unsigned long long* __restrict__ packed, uint32_t* __restrict__ slot_of,
                                   int64_t dst, int64_t src) {
  const int c = threadIdx.x;
  if (c == 0) {
    packed[dst] = packed[src];
    slot_of[dst] = slot_of[src];
  }
  if (c >= mc.n) return;
  for (int b = 0; b < mc.w[c]; b++) mc.d[c][dst * mc.w[c] + b] = mc.d[c][src * mc.w[c] + b];
  if (uint32_t* v = mc.v[c]) {
    const uint32_t bit = 1u << (dst & 31);
    if ((v[src >> 5] >> (src & 31)) & 1u) atomicOr(&v[dst >> 5], bit);
    else atomicAnd(&v[dst >> 5], ~bit);
  }
}

// The calling thread's last bolt_groupby plan (bolt_groupby_last_plan).
struct GbLastPlan {
  int32_t tier;  // 0 LDS tier only, 1 global table, 2 partitioned tier
  int64_t spilled, partials;
  int32_t partitions;
  int32_t est_bucket, words;  // the fan-out tuning key
};
static thread_local GbLastPlan t_last_plan{0, 0, 0, 0, 0, 0};

// ---- the partitioned tier's fan-out ---------------------------------------------
// bolt_groupby_autotune_partitions times the partition counts around the
// default (gb_part_bits) and keeps the fastest per (log2 of the estimated
// group count, accumulator words) for the rest of the process;
// BOLT_GB_PART_BITS_DELTA overrides.  While it times a candidate,
// t_force_delta holds it.
std::mutex g_part_mu;
std::map<std::pair<int, int>, int> g_part_delta;
static thread_local int t_force_delta = 0;
static thread_local bool t_forcing = false;

int est_bucket(double est) { return est < 1 ? 0 : (int)std::floor(std::log2(est)); }

int part_bits_delta(double est, int words) {
  if (t_forcing) return t_force_delta;
  if (const char* e = getenv("BOLT_GB_PART_BITS_DELTA")) return std::max(-12, std::min(12, atoi(e)));
  std::lock_guard<std::mutex> lk(g_part_mu);
  const auto it = g_part_delta.find({est_bucket(est), words});
  return it == g_part_delta.end() ? 0 : it->second;
}

// Gather for the sorted output order: out[i] = in[perm[i]].
__global__ void gather_kernel(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, int width,
                              const uint32_t* __restrict__ perm, int64_t n, const uint32_t* __restrict__ vsrc,
                              uint32_t* __restrict__ vdst) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    int64_t p = perm[i];
    switch (width) {
      case 4: reinterpret_cast<uint32_t*>(dst)[i] = reinterpret_cast<const uint32_t*>(src)[p]; break;
      case 8: reinterpret_cast<uint64_t*>(dst)[i] = reinterpret_cast<const uint64_t*>(src)[p]; break;
      case 16: reinterpret_cast<uint4*>(dst)[i] = reinterpret_cast<const uint4*>(src)[p]; break;
    }
    if (vsrc && ((vsrc[p >> 5] >> (p & 31)) & 1u)) atomicOr(&vdst[i >> 5], 1u << (i & 31));
  }
}

// i64 order of the packed key as an unsigned radix key
__global__ void flip_sign_kernel(const unsigned long long* __restrict__ in, uint64_t* __restrict__ out,
                                 uint32_t* __restrict__ idx, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    out[i] = in[i] ^ 0x8000000000000000ull;
    idx[i] = (uint32_t)i;
  }
}

// Ascending order of a few signed 64-bit keys (ties by position): the rank
// of a key is the number of keys that precede it.  A workgroup ranks 64 keys;
// its four waves count over interleaved quarters of all keys (staged through
// LDS, read as broadcasts) and add their counts.  One launch and no host
// round trip, where the radix sort needs a histogram read-back; O(n^2) work,
// so only for small n.
constexpr int kRankChunk = 2048;
constexpr int64_t kRankSortMax = 4096;
__global__ __launch_bounds__(256) void rank_sort_kernel(const unsigned long long* __restrict__ keys, int64_t n,
                                                        uint32_t* __restrict__ perm) {
  __shared__ long long sk[kRankChunk];
  __shared__ uint32_t part[4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t i = (int64_t)blockIdx.x * 64 + lane;
  const long long ki = i < n ? (long long)keys[i] : 0;
  uint32_t r0 = 0, r1 = 0;
  for (int64_t base = 0; base < n; base += kRankChunk) {
    const int m = (int)(n - base < kRankChunk ? n - base : kRankChunk);
    __syncthreads();
    for (int j = threadIdx.x; j < m; j += 256) sk[j] = (long long)keys[base + j];
    __syncthreads();
    const int64_t li = i - base;  // position of key i relative to this chunk
    int j = wave;
    for (; j + 4 < m; j += 8) {
      const long long a = sk[j], b = sk[j + 4];
      r0 += (a < ki || (a == ki && j < li)) ? 1u : 0u;
      r1 += (b < ki || (b == ki && j + 4 < li)) ? 1u : 0u;
    }
    if (j < m) {
      const long long a = sk[j];
      r0 += (a < ki || (a == ki && j < li)) ? 1u : 0u;
    }
  }
  part[wave][lane] = r0 + r1;
  __syncthreads();
  if (wave == 0 && i < n) perm[part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane]] = (uint32_t)i;
}

}  // namespace bolt

using namespace bolt;

// ---- result tables ------------------------------------------------------------------------

struct bolt_table {
  std::vector<bolt_column_t> cols;
  std::vector<void*> owned;
  int64_t n_rows = 0;
  hipStream_t stream = nullptr;
};

extern "C" int bolt_table_num_rows(const bolt_table_t* t, int64_t* n) {
  if (!t) return set_error(BOLT_ERR_PLAN, "table is NULL");
  *n = t->n_rows;
  return BOLT_OK;
}
extern "C" int bolt_table_num_columns(const bolt_table_t* t, int32_t* n) {
  if (!t) return set_error(BOLT_ERR_PLAN, "table is NULL");
  *n = (int32_t)t->cols.size();
  return BOLT_OK;
}
extern "C" int bolt_table_column(const bolt_table_t* t, int32_t i, bolt_column_t* col) {
  if (!t || i < 0 || i >= (int32_t)t->cols.size()) return set_error(BOLT_ERR_PLAN, "column index out of range");
  *col = t->cols[i];
  return BOLT_OK;
}
extern "C" void bolt_table_release(bolt_table_t* t) {
  if (!t) return;
  for (void* p : t->owned) dev_free(p, t->stream);
  delete t;
}

static int table_alloc_col(bolt_table_t* t, int32_t dtype, int64_t cap, bool with_valid, hipStream_t s,
                           bolt_column_t* out) {
  void* d = nullptr;
  BOLT_TRY(dev_alloc(&d, (size_t)(cap > 0 ? cap : 1) * dtype_size(dtype), s));
  t->owned.push_back(d);
  void* v = nullptr;
  if (with_valid) {
    size_t words = (size_t)(cap + 31) / 32 + 1;
    BOLT_TRY(dev_alloc(&v, words * 4, s));
    t->owned.push_back(v);
    BOLT_TRY(fill_u32(v, 0, words, s));
  }
  out->data = d;
  out->validity = static_cast<const uint8_t*>(v);
  out->length = 0;
  out->dtype = dtype;
  out->scale = 0;
  return BOLT_OK;
}

// ---- plan construction ---------------------------------------------------------------------

int bolt::make_gplan(const bolt_agg_t* aggs, int32_t n_aggs, const bolt_column_t* inputs, int32_t n_inputs,
                      int64_t n, GPlan* P, int32_t* out_dtype) {
  memset(P, 0, sizeof *P);
  if (n_aggs < 0 || n_aggs > kMaxAggs) return set_error(BOLT_ERR_PLAN, "%d aggregates (max %d)", n_aggs, kMaxAggs);
  int w = 0;
  for (int i = 0; i < n_aggs; i++) {
    GAgg& g = P->a[i];
    g.op = aggs[i].op;
    g.nw = -1;
    if (g.op < BOLT_AGG_SUM || g.op > BOLT_AGG_STDDEV_SAMP) return set_error(BOLT_ERR_PLAN, "aggregate %d: unknown op", i);
    if (g.op == BOLT_AGG_COUNT_STAR) {
      g.kind = K_COUNT;
      g.dtype = BOLT_I64;
      g.w = w++;
      out_dtype[i] = BOLT_I64;
      continue;
    }
    if (aggs[i].col < 0 || aggs[i].col >= n_inputs)
      return set_error(BOLT_ERR_PLAN, "aggregate %d: input column %d out of range (%d inputs)", i, aggs[i].col, n_inputs);
    const bolt_column_t& c = inputs[aggs[i].col];
    if (c.length != n)
      return set_error(BOLT_ERR_PLAN, "aggregate %d: input has %lld rows, keys have %lld", i, (long long)c.length,
                       (long long)n);
    g.data = c.data;
    g.validity = c.validity;
    g.dtype = c.dtype;
    bool nullable = c.validity != nullptr;
    int dt = c.dtype;
    if (g.op == BOLT_AGG_COUNT) {
      g.kind = K_COUNT;
      g.w = w++;
      out_dtype[i] = BOLT_I64;
      continue;
    }
    if (dt == BOLT_BOOL) return set_error(BOLT_ERR_TYPE, "aggregate %d: Bool inputs are not supported", i);
    if (dt == BOLT_TIMESTAMP)
      return set_error(BOLT_ERR_TYPE, "aggregate %d: a Timestamp enters through an expression program (EXTRACT, "
                       "DATE_TRUNC, comparisons)", i);
    switch (g.op) {
      case BOLT_AGG_SUM:
        if (dt == BOLT_DATE32) return set_error(BOLT_ERR_TYPE, "SUM over a temporal type is undefined in SQL (Date32)");
        if (dt == BOLT_DEC128) { g.kind = K_DSUM; g.w = w; w += 3; out_dtype[i] = BOLT_DEC128; }
        else if (dtype_is_float(dt)) { g.kind = K_FSUM; g.w = w++; out_dtype[i] = dt; }
        else { g.kind = K_ISUM; g.w = w; w += 2; out_dtype[i] = BOLT_I64; }
        if (nullable) g.nw = w++;
        break;
      case BOLT_AGG_MIN:
      case BOLT_AGG_MAX:
        if (dt == BOLT_DEC128) return set_error(BOLT_ERR_UNSUPPORTED, "GROUP BY MIN/MAX over Decimal128");
        g.kind = K_MINMAX;
        g.w = w++;
        out_dtype[i] = dt;
        if (nullable) g.nw = w++;
        break;
      case BOLT_AGG_AVG:
        if (dt == BOLT_DATE32 || dt == BOLT_DEC128) return set_error(BOLT_ERR_TYPE, "AVG over %s is not supported", dtype_name(dt));
        g.kind = K_AVG;
        g.w = w++;
        g.nw = w++;
        out_dtype[i] = BOLT_F64;
        break;
      default:
        if (dt == BOLT_DATE32 || dt == BOLT_DEC128) return set_error(BOLT_ERR_TYPE, "VAR/STDDEV over %s is not supported", dtype_name(dt));
        g.kind = K_VAR;
        g.w = w;
        w += 2;
        g.nw = w++;
        out_dtype[i] = BOLT_F64;
        break;
    }
  }
  P->n = n_aggs;
  P->words = w;
  return BOLT_OK;
}

double bolt::hll_estimate(const uint32_t* regs) {
  const double m = kHllRegs;
  double sum = 0;
  int zeros = 0;
  for (int i = 0; i < kHllRegs; i++) {
    sum += std::ldexp(1.0, -(int)regs[i]);
    zeros += regs[i] == 0;
  }
  double alpha = 0.7213 / (1.0 + 1.079 / m);
  double e = alpha * m * m / sum;
  if (e <= 2.5 * m && zeros > 0) e = m * std::log(m / zeros);
  return e;
}

uint64_t bolt::next_pow2(uint64_t x) {
  uint64_t p = 1;
  while (p < x) p <<= 1;
  return p;
}

// The generated LDS-tier kernel's shape for this query (filter, keys, plan).
static int gb_jit_spec(const KeySpec& ks, const GPlan& P, const DevFilter& f, GbJitSpec* sp, JitLits* lits,
                       JitCols* cols) {
  *sp = GbJitSpec();
  memset(lits, 0, sizeof *lits);
  memset(cols, 0, sizeof *cols);
  sp->n_keys = ks.n_keys;
  sp->t0 = ks.t0;
  sp->t1 = ks.t1;
  sp->kv0 = ks.v0 != nullptr;
  sp->kv1 = ks.n_keys == 2 && ks.v1 != nullptr;
  sp->words = P.words;
  sp->ident.assign(P.words, 0ull);
  for (int a = 0; a < P.n; a++)
    if (P.a[a].kind == K_MINMAX) sp->ident[P.a[a].w] = P.a[a].op == BOLT_AGG_MIN ? ~0ull : 0ull;
  for (int t = 0; t < f.n; t++) {
    const int c = jit_row_agg_col(f.t[t].data, f.t[t].validity, f.t[t].dtype, &sp->cols, cols);
    if (c < 0) return set_error(BOLT_ERR_UNSUPPORTED, "jit: too many columns");
    sp->terms.push_back({c, f.t[t].op});
    lits->i[t] = f.t[t].lit_i64;
    lits->f[t] = f.t[t].lit_f64;
    lits->hi[t] = f.t[t].lit_hi;
  }
  for (int a = 0; a < P.n; a++) {
    const GAgg& g = P.a[a];
    GbJitSpec::Agg ja{g.op, g.kind, g.dtype, -1, g.w, g.nw};
    if (g.op != BOLT_AGG_COUNT_STAR) {
      ja.col = jit_row_agg_col(g.data, g.validity, g.dtype, &sp->cols, cols);
      if (ja.col < 0) return set_error(BOLT_ERR_UNSUPPORTED, "jit: too many columns");
    }
    sp->aggs.push_back(ja);
  }
  // rows in flight per lane: as many as the registers hold at 1024 threads
  // (32-bit words loaded per row: keys and every column)
  int words = ks.n_keys;
  for (const JitSpec::Col& c : sp->cols) words += c.dtype == BOLT_DEC128 ? 4 : (dtype_size(c.dtype) + 3) / 4;
  sp->unroll = words <= 6 ? 8 : words <= 12 ? 4 : 2;
  if (const char* u = getenv("BOLT_GB_UNROLL")) {
    const int v = atoi(u);
    if (v >= 1 && v <= 8) sp->unroll = v;
  }
  return BOLT_OK;
}

static int lds_budget_bytes() {
  static int cached = -1;
  if (cached < 0) {
    int dev = 0, v = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) {
      (void)hipGetLastError();
      v = 65536;
    }
    int kb = 96;  // one 1024-thread workgroup per CU either way; BOLT_GB_LDS_KB overrides
    if (const char* e = getenv("BOLT_GB_LDS_KB")) {
      const int x = atoi(e);
      if (x >= 16 && x <= 160) kb = x;
    }
    cached = std::min(v, kb * 1024);
  }
  return cached;
}

extern "C" int bolt_groupby(bolt_stream_t stream, const bolt_column_t* keys, int32_t n_keys,
                            const bolt_column_t* inputs, int32_t n_inputs, const bolt_agg_t* aggs, int32_t n_aggs,
                            const bolt_filter_t* filter, uint32_t flags, bolt_table_t** out) {
  BOLT_CHECK_DEVICE();
  clear_error();
  *out = nullptr;
  hipStream_t s = (hipStream_t)stream;
  if (n_keys < 1 || n_keys > 2 || keys == nullptr)
    return set_error(BOLT_ERR_PLAN, "bolt_groupby: 1 or 2 key columns required, got %d", n_keys);
  const int64_t n = keys[0].length;
  if (n > 0xFFFFFFFFll) return set_error(BOLT_ERR_PLAN, "bolt_groupby: %lld rows exceeds the u32 row-id limit", (long long)n);
  KeySpec ks;
  memset(&ks, 0, sizeof ks);
  ks.n_keys = n_keys;
  for (int k = 0; k < n_keys; k++) {
    int dt = keys[k].dtype;
    // src/exec/groupby_common.rs:117 key_bit_width accepts Int32 / Int64 /
    // floats; Date32 packs like Int32, and a single Decimal128 key whose
    // values fit i64 packs like Int64 (wider values decline)
    if (!(dt == BOLT_I32 || dt == BOLT_I64 || dt == BOLT_F32 || dt == BOLT_F64 || dt == BOLT_DATE32 ||
          (dt == BOLT_DEC128 && n_keys == 1)))
      return set_error(BOLT_ERR_TYPE, "GROUP BY key dtype %s not supported", dtype_name(dt));
    if (keys[k].length != n) return set_error(BOLT_ERR_PLAN, "bolt_groupby: key columns differ in length");
    if (n_keys == 2 && dtype_size(dt) != 4)
      return set_error(BOLT_ERR_UNSUPPORTED,
                       "multi-column GROUP BY with > 64 bits of key width not yet supported");
  }
  ks.d0 = keys[0].data;
  ks.v0 = keys[0].validity;
  ks.t0 = keys[0].dtype;
  if (n_keys == 2) {
    ks.d1 = keys[1].data;
    ks.v1 = keys[1].validity;
    ks.t1 = keys[1].dtype;
  }
  ks.null_group = (n_keys == 1 && keys[0].validity != nullptr) ? 1 : 0;
  GPlan P;
  int32_t out_dt[kMaxAggs];
  BOLT_TRY(make_gplan(aggs, n_aggs, inputs, n_inputs, n, &P, out_dt));
  DevFilter f;
  BOLT_TRY(make_dev_filter(filter, inputs, n_inputs, n, &f));
  if (keys[0].dtype == BOLT_DEC128 && n > 0) {
    DevBuf<uint32_t> err;
    uint32_t h = 0;
    BOLT_TRY(err.alloc(1, s));
    BOLT_TRY(fill_u32(err.ptr, 0, 1, s));
    dec_key_fit_kernel<<<grid_for(n, 256, 4), 256, 0, s>>>(static_cast<const uint64_t*>(keys[0].data),
                                                          keys[0].validity, n, err.ptr);
    BOLT_HIP(hipGetLastError());
    BOLT_HIP(hipMemcpyAsync(&h, err.ptr, 4, hipMemcpyDeviceToHost, s));
    BOLT_HIP(hipStreamSynchronize(s));
    if (h)
      return set_error(BOLT_DECLINE, "GROUP BY on a Decimal128 key with values beyond 64 bits (the host path takes it)");
  }

  // ---- pass A ----
  const int W = P.words;
  const int budget = lds_budget_bytes();
  int S = (int)((budget - (kHllRegs + 4) * 4) / (8 * (1 + W))) - 2;
  if (S > 4096) S = 4096;
  const bool use_lds = !(flags & BOLT_GROUPBY_FORCE_GLOBAL) && S >= 256 && n > 0;
  int64_t n_parts = 0, n_spill = 0;
  double est = 0;
  DevBuf<unsigned long long> pkeys, pwords;
  DevBuf<uint8_t> ptag;
  DevBuf<uint32_t> spill, hll, spill_seg;
  DevBuf<unsigned long long> spill_key;
  unsigned grid = 0;
  int64_t spill_per = 0;
  if (use_lds) {
    const size_t lds = (size_t)(S + 2) * 8 * (1 + W) + (kHllRegs + 4) * 4;
    // the per-query generated kernel when hipRTC is available, else the
    // precompiled one; rows_per_iter is what one workgroup iteration covers
    hipFunction_t gfn = nullptr;
    GbJitSpec gs;
    JitLits glits;
    JitCols gcols;
    if (jit_enabled() && gb_jit_spec(ks, P, f, &gs, &glits, &gcols) == BOLT_OK &&
        jit_get_keyed(gb_jit_key(gs), [&] { return jit_gb_lds_source(gs); }, &gfn) != BOLT_OK)
      gfn = nullptr;
    clear_error();
    const int64_t rows_per_iter = (int64_t)kGbBlock * (gfn ? gs.unroll : 1);
    if (!gfn)
      BOLT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(gb_lds_kernel),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    grid = grid_for(n, kGbBlock, 16, 256);
    const int64_t pcap = (int64_t)grid * (S + 2);
    // a workgroup handles at most ceil(n / (grid * rows_per_iter)) iterations
    spill_per = (n + (int64_t)grid * rows_per_iter - 1) / ((int64_t)grid * rows_per_iter) * rows_per_iter;
    BOLT_TRY(pkeys.alloc((size_t)pcap, s));
    BOLT_TRY(ptag.alloc((size_t)pcap, s));
    BOLT_TRY(pwords.alloc((size_t)pcap * (W > 0 ? W : 1), s));
    BOLT_TRY(spill.alloc((size_t)grid * spill_per, s));
    BOLT_TRY(spill_seg.alloc(grid, s));
    // inputs large enough for the partitioned tier: the spilled rows' keys are
    // captured beside the row ids (8 bytes more per spilled row, from a
    // register), so that tier's histogram and first scatter read them in
    // order instead of gathering them by row id
    if (n > ((int64_t)1 << 25)) BOLT_TRY(spill_key.alloc((size_t)grid * spill_per, s));
    // sketch registers and the two counters in one buffer: one read-back
    BOLT_TRY(hll.alloc(kHllRegs + 4, s));
    BOLT_TRY(fill_u32(hll.ptr, 0, kHllRegs + 4, s));
    unsigned long long* counters = reinterpret_cast<unsigned long long*>(hll.ptr + kHllRegs);
    PartOut po{pkeys.ptr, ptag.ptr, pwords.ptr, pcap, counters, spill.ptr, counters + 1, spill_seg.ptr,
               spill_per, hll.ptr, spill_key.ptr};
    prof_begin(s);
    if (gfn) {
      long long nn = n;
      int SSz = S;
      void* args[] = {&nn, &SSz, (void*)&ks.d0, (void*)&ks.v0, (void*)&ks.d1, (void*)&ks.v1, &glits, &gcols, &po};
      BOLT_HIP(hipModuleLaunchKernel(gfn, grid, 1, 1, kGbBlock, 1, 1, (unsigned)lds, s, args, nullptr));
    } else {
      gb_lds_kernel<<<grid, kGbBlock, lds, s>>>(ks, f, P, n, S, po);
    }
    prof_end(s);
    BOLT_HIP(hipGetLastError());
    std::vector<uint32_t> regs(kHllRegs + 4);
    BOLT_HIP(hipMemcpyAsync(regs.data(), hll.ptr, (kHllRegs + 4) * 4, hipMemcpyDeviceToHost, s));
    BOLT_HIP(hipStreamSynchronize(s));
    unsigned long long hc[2];
    memcpy(hc, regs.data() + kHllRegs, 16);
    n_parts = (int64_t)hc[0];
    n_spill = (int64_t)hc[1];
    est = hll_estimate(regs.data());
  } else {
    n_spill = n;  // every row goes to the global table
    est = (double)n;
  }

  // ---- pass B (global table) and pass C (extraction) ----
  // One host round trip for both: a table that overflowed is detected with
  // the extraction's results and the two passes rerun 4x larger.
  const int64_t upper = n_parts + n_spill;  // exact bound on the group count
  uint64_t want = (uint64_t)std::min<double>((double)upper, est * 1.25 + 64.0);
  uint64_t cap = next_pow2(std::max<uint64_t>(1024, 2 * want));
  if (const char* e = getenv("BOLT_GB_INITIAL_CAP")) {  // testing: force the overflow-and-rerun path
    const long long c = atoll(e);
    if (c >= 1024) cap = next_pow2((uint64_t)c);
  }
  bool has_var = false;
  for (int a = 0; a < P.n; a++) has_var |= P.a[a].kind == K_VAR;
  // many spilled rows over many groups: the radix-partitioned tier instead of
  // global atomics per spilled row (its records are built once, whatever the
  // table size).  Its passes cost a fixed ~2 ms more than the atomics save
  // below a few tens of millions of spilled rows (h2o q2 / q3 / q5 at 10M
  // rows: 5.0 / 3.8 / 5.6 ms partitioned, 3.0 / 3.0 / 4.4 ms with the global
  // table; 256M Zipf rows: 10.4 ms partitioned, 14.7 ms global), so it takes
  // over from 2^25 spilled rows.
  const bool partitioned = use_lds && n > 0 &&
                           ((flags & BOLT_GROUPBY_FORCE_PARTITIONED) ||
                            (n_spill > (int64_t)(1 << 25) && est > 0.75 * (double)S));
  // the partitioned tier writes each group once, with one claim, so its
  // table is sized for at most half load (1.6x the padded estimate) rather than
  // 2x: half the slots to initialise and for the extraction to scan
  if (partitioned && !getenv("BOLT_GB_INITIAL_CAP"))
    cap = next_pow2(std::max<uint64_t>(1024, (uint64_t)((double)want * 1.6)));
  GbPartitions parts;
  if (partitioned)
    BOLT_TRY(gb_part_build(s, ks, P, pkeys.ptr, ptag.ptr, pwords.ptr, (int64_t)pkeys.count, n_parts, spill.ptr,
                           spill_key.ptr, spill_per, spill_seg.ptr, (int)grid, n_spill, est, S,
                           part_bits_delta(est, W), &parts));
  t_last_plan = GbLastPlan{!use_lds ? 1 : (partitioned ? 2 : (n_spill > 0 ? 1 : 0)), n_spill, n_parts,
                           partitioned ? parts.np : 0, est_bucket(est), W};
  GTable t;
  DevBuf<unsigned long long> tkeys, twords;
  DevBuf<uint32_t> tmisc;
  BOLT_TRY(tmisc.alloc(8, s));
  bolt_table* tbl = nullptr;
  std::vector<bolt_column_t> cols(n_keys + n_aggs);
  DevBuf<unsigned long long> packed;
  DevBuf<uint32_t> slot_of;
  int64_t n_groups = 0;
  int st = BOLT_OK;
  for (int attempt = 0;; attempt++) {
    BOLT_TRY(tkeys.alloc(cap, s));
    BOLT_TRY(twords.alloc((cap + 2) * (W > 0 ? W : 1), s));
    BOLT_TRY(fill_u32(tmisc.ptr, 0, 8, s));
    t.keys = tkeys.ptr;
    t.words = twords.ptr;
    t.used_special = tmisc.ptr;
    t.err = tmisc.ptr + 4;
    t.cap = cap;
    gb_init_kernel<<<grid_for((int64_t)cap + 2, 256, 1), 256, 0, s>>>(t, P);
    if (partitioned) {
      BOLT_TRY(gb_part_reduce(s, t, P, parts, S));
    } else if (n_parts > 0) {
      if (est < 0.75 * (double)S && n_parts > 4096) {
        // few groups: two-level merge through LDS tables of the same size
        const unsigned mgrid = (unsigned)std::min<int64_t>(64, (n_parts + 4095) / 4096);
        const size_t mlds = (size_t)(S + 2) * 8 * (1 + W) + 16;
        BOLT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(gb_merge_lds_kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)mlds));
        gb_merge_lds_kernel<<<mgrid, kGbBlock, mlds, s>>>(t, P, pkeys.ptr, ptag.ptr, pwords.ptr,
                                                          (int64_t)pkeys.count, n_parts, S);
      } else {
        gb_merge_parts_kernel<<<grid_for(n_parts, 256, 4), 256, 0, s>>>(t, P, pkeys.ptr, ptag.ptr, pwords.ptr,
                                                                        (int64_t)pkeys.count, n_parts);
      }
    }
    if (n_spill > 0 && !partitioned) {
      if (use_lds) {
        const int K = 16;  // blocks per spill segment
        gb_rows_seg_kernel<<<grid * K, 256, 0, s>>>(t, P, ks, spill.ptr, spill_per, spill_seg.ptr, K);
      } else {
        if (attempt == 0) prof_begin(s);
        gb_rows_kernel<<<grid_for(n_spill, 256, 4), 256, 0, s>>>(t, P, ks, f, nullptr, n_spill);
        if (attempt == 0) prof_end(s);
      }
    }
    if (has_var && n > 0) gb_m2_kernel<<<grid_for(n, 256, 4), 256, 0, s>>>(t, P, ks, f, n);
    BOLT_HIP(hipGetLastError());

    const uint64_t total = cap + 2;
    const int64_t out_cap = (int64_t)std::min<uint64_t>(total, (uint64_t)std::max<int64_t>(upper, 0) + 2);
    tbl = new bolt_table();
    tbl->stream = s;
    OutSpec o;
    memset(&o, 0, sizeof o);
    for (int k = 0; k < n_keys && st == BOLT_OK; k++) {
      st = table_alloc_col(tbl, keys[k].dtype, out_cap, k == 0 && ks.null_group, s, &cols[k]);
      cols[k].scale = keys[k].scale;
      o.key[k] = OutCol{static_cast<uint8_t*>(cols[k].data), (uint32_t*)cols[k].validity, keys[k].dtype, 0};
    }
    for (int a = 0; a < n_aggs && st == BOLT_OK; a++) {
      const GAgg& g = P.a[a];
      bool can_null = (g.nw >= 0 && g.kind != K_COUNT) ||
                      g.op == BOLT_AGG_VAR_SAMP || g.op == BOLT_AGG_STDDEV_SAMP;
      st = table_alloc_col(tbl, out_dt[a], out_cap, can_null, s, &cols[n_keys + a]);
      o.agg[a] = OutCol{static_cast<uint8_t*>(cols[n_keys + a].data), (uint32_t*)cols[n_keys + a].validity, out_dt[a], 0};
    }
    if (st == BOLT_OK) st = packed.alloc((size_t)out_cap, s);
    if (st == BOLT_OK) st = slot_of.alloc((size_t)out_cap, s);
    if (st != BOLT_OK) {
      bolt_table_release(tbl);
      return st;
    }
    o.packed = packed.ptr;
    o.slot_of = slot_of.ptr;
    const bool small = total <= (1u << 20);
    const int64_t per_tile = (int64_t)kExBlock * (small ? 1 : kExItems);
    const int64_t tiles = (int64_t)((total + per_tile - 1) / per_tile);
    // look-back status words and the ticket / total in one buffer, one fill
    DevBuf<uint64_t> status;
    if ((st = status.alloc((size_t)tiles + 2, s)) != BOLT_OK ||
        (st = fill_u32(status.ptr, 0, ((size_t)tiles + 2) * 2, s)) != BOLT_OK) {
      bolt_table_release(tbl);
      return st;
    }
    uint32_t* emisc = reinterpret_cast<uint32_t*>(status.ptr + tiles);
    if (small)
      gb_extract_kernel<1><<<(unsigned)tiles, kExBlock, 0, s>>>(t, P, ks, o, status.ptr, emisc,
                                                               reinterpret_cast<unsigned long long*>(emisc + 2));
    else
      gb_extract_kernel<kExItems><<<(unsigned)tiles, kExBlock, 0, s>>>(
          t, P, ks, o, status.ptr, emisc, reinterpret_cast<unsigned long long*>(emisc + 2));
    hipError_t le = hipGetLastError();
    uint32_t eh[4] = {0, 0, 0, 0}, terr[4] = {0, 0, 0, 0};
    if (le == hipSuccess) le = hipMemcpyAsync(eh, emisc, 16, hipMemcpyDeviceToHost, s);
    if (le == hipSuccess) le = hipMemcpyAsync(terr, t.err, 16, hipMemcpyDeviceToHost, s);
    if (le == hipSuccess) le = hipStreamSynchronize(s);
    if (le != hipSuccess) {
      bolt_table_release(tbl);
      return set_error(BOLT_ERR_DEVICE, "GROUP BY extraction failed: %s", hipGetErrorString(le));
    }
    if (terr[0]) {  // the global table overflowed: discard this result and rerun larger
      bolt_table_release(tbl);
      tbl = nullptr;
      if (attempt >= 3 || cap >= (1ull << 33))
        return set_error(BOLT_ERR_MEMORY, "GROUP BY hash table overflow at %llu slots", (unsigned long long)cap);
      cap *= 4;
      continue;
    }
    n_groups = (int64_t)(((uint64_t)eh[3] << 32) | eh[2]);
    if (terr[1]) {
      bolt_table_release(tbl);
      return set_error(BOLT_ERR_OTHER, "GROUP BY extraction look-back timed out");
    }
    if (terr[2] & 1u) {
      bolt_table_release(tbl);
      return set_error(BOLT_ERR_TYPE, "SUM(integer) overflow: accumulator exceeds i64 range");
    }
    if (terr[2] & 2u) {
      bolt_table_release(tbl);
      return set_error(BOLT_ERR_TYPE, "SUM(Decimal128) precision overflow: accumulator exceeds i128 range");
    }
    break;
  }

  if (ks.null_group && (keys[0].dtype == BOLT_I64 || keys[0].dtype == BOLT_DEC128) && n_groups >= 2) {
    DevBuf<int64_t> clash;
    int64_t hc[2] = {-1, 0};
    st = clash.alloc(2, s);
    if (st == BOLT_OK) st = fill_u32(clash.ptr, 0xFFFFFFFFu, 2, s);
    if (st == BOLT_OK) st = fill_u32(reinterpret_cast<uint32_t*>(clash.ptr + 1), 0, 2, s);
    if (st == BOLT_OK) {
      gb_null_clash_kernel<<<grid_for(n_groups, 256, 4), 256, 0, s>>>(packed.ptr, slot_of.ptr, n_groups, cap,
                                                                     clash.ptr);
      if (hipGetLastError() != hipSuccess || hipMemcpyAsync(hc, clash.ptr, 16, hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess)
        st = set_error(BOLT_ERR_DEVICE, "GROUP BY NULL-group check failed");
    }
    if (st == BOLT_OK && hc[0] >= 0 && hc[1]) {
      MoveCols mc;
      memset(&mc, 0, sizeof mc);
      mc.n = (int32_t)cols.size();
      for (int c = 0; c < mc.n; c++) {
        mc.d[c] = static_cast<uint8_t*>(const_cast<void*>(cols[c].data));
        mc.v[c] = reinterpret_cast<uint32_t*>(const_cast<uint8_t*>(cols[c].validity));
        mc.w[c] = dtype_size(cols[c].dtype);
      }
      if (hc[0] != n_groups - 1)
        gb_move_row_kernel<<<1, 64, 0, s>>>(mc, packed.ptr, slot_of.ptr, hc[0], n_groups - 1);
      if (hipGetLastError() != hipSuccess) st = set_error(BOLT_ERR_DEVICE, "GROUP BY NULL-group removal failed");
      n_groups--;
    }
    if (st != BOLT_OK) {
      bolt_table_release(tbl);
      return st;
    }
  }

  // ---- ordering by packed key (src/exec/groupby.rs:800) ----
  if ((flags & BOLT_GROUPBY_SORTED) && n_groups > 1) {
    DevBuf<uint64_t> kin, kout;
    DevBuf<uint32_t> vin, perm;
    st = perm.alloc((size_t)n_groups, s);
    if (st == BOLT_OK && n_groups > kRankSortMax) {
      st = kin.alloc((size_t)n_groups, s);
      if (st == BOLT_OK) st = kout.alloc((size_t)n_groups, s);
      if (st == BOLT_OK) st = vin.alloc((size_t)n_groups, s);
    }
    if (st == BOLT_OK && n_groups <= kRankSortMax) {
      rank_sort_kernel<<<(unsigned)((n_groups + 63) / 64), 256, 0, s>>>(packed.ptr, n_groups, perm.ptr);
    } else if (st == BOLT_OK) {
      flip_sign_kernel<<<grid_for(n_groups, 256, 4), 256, 0, s>>>(packed.ptr, kin.ptr, vin.ptr, n_groups);
      st = radix_sort_u64_pairs(s, kin.ptr, vin.ptr, n_groups, kout.ptr, perm.ptr, 64);
    }
    for (size_t c = 0; c < cols.size() && st == BOLT_OK; c++) {
      bolt_column_t sorted;
      st = table_alloc_col(tbl, cols[c].dtype, n_groups, cols[c].validity != nullptr, s, &sorted);
      if (st != BOLT_OK) break;
      gather_kernel<<<grid_for(n_groups, 256, 4), 256, 0, s>>>(
          static_cast<const uint8_t*>(cols[c].data), static_cast<uint8_t*>(sorted.data), dtype_size(cols[c].dtype),
          perm.ptr, n_groups, reinterpret_cast<const uint32_t*>(cols[c].validity),
          reinterpret_cast<uint32_t*>(const_cast<uint8_t*>(sorted.validity)));
      cols[c] = sorted;
    }
    if (st == BOLT_OK && hipGetLastError() != hipSuccess) st = set_error(BOLT_ERR_DEVICE, "GROUP BY gather failed");
    if (st == BOLT_OK && hipStreamSynchronize(s) != hipSuccess) st = set_error(BOLT_ERR_DEVICE, "GROUP BY sort failed");
    if (st != BOLT_OK) {
      bolt_table_release(tbl);
      return st;
    }
  }
  for (auto& c : cols) c.length = n_groups;
  tbl->cols = cols;
  tbl->n_rows = n_groups;
  *out = tbl;
  return BOLT_OK;
}

extern "C" int bolt_groupby_last_plan(int32_t* tier, int64_t* spilled_rows, int64_t* partial_records,
                                      int32_t* partitions) {
  if (tier) *tier = bolt::t_last_plan.tier;
  if (spilled_rows) *spilled_rows = bolt::t_last_plan.spilled;
  if (partial_records) *partial_records = bolt::t_last_plan.partials;
  if (partitions) *partitions = bolt::t_last_plan.partitions;
  return BOLT_OK;
}

extern "C" int bolt_groupby_autotune_partitions(bolt_stream_t stream, const bolt_column_t* keys, int32_t n_keys,
                                                const bolt_column_t* inputs, int32_t n_inputs, const bolt_agg_t* aggs,
                                                int32_t n_aggs, int32_t reps, int32_t* best_partitions,
                                                float* best_ms) {
  using namespace bolt;
  if (reps < 1) reps = 1;
  double best = 1e300;
  int best_delta = 0, best_np = 0, key_bucket = 0, key_words = 0;
  int seen_np[5] = {-1, -1, -1, -1, -1};
  int st = BOLT_OK;
  for (int d = -2; d <= 2 && st == BOLT_OK; d++) {
    t_forcing = true;
    t_force_delta = d;
    double total = 0;
    for (int r = 0; r <= reps && st == BOLT_OK; r++) {  // one untimed call first (pool, code objects)
      bolt_table_t* t = nullptr;
      const auto t0 = std::chrono::steady_clock::now();
      st = bolt_groupby(stream, keys, n_keys, inputs, n_inputs, aggs, n_aggs, nullptr,
                        BOLT_GROUPBY_FORCE_PARTITIONED, &t);
      const auto t1 = std::chrono::steady_clock::now();
      if (t) bolt_table_release(t);
      if (r > 0) total += std::chrono::duration<double, std::milli>(t1 - t0).count();
    }
    t_forcing = false;
    if (st != BOLT_OK) break;
    const int np = t_last_plan.partitions;
    bool dup = false;  // a delta that clamps to a partition count already timed
    for (int k = 0; k < 5; k++) dup |= seen_np[k] == np;
    seen_np[d + 2] = np;
    key_bucket = t_last_plan.est_bucket;
    key_words = t_last_plan.words;
    if (!dup && total / reps < best) {
      best = total / reps;
      best_delta = d;
      best_np = np;
    }
  }
  t_forcing = false;
  if (st != BOLT_OK) return st;
  {
    std::lock_guard<std::mutex> lk(g_part_mu);
    g_part_delta[{key_bucket, key_words}] = best_delta;
  }
  if (best_partitions) *best_partitions = best_np;
  if (best_ms) *best_ms = (float)best;
  return BOLT_OK;
}
