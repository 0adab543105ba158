// sort.hip — stable LSD radix sort of (64-bit key, u32 value) pairs, and
// the multi-column ORDER BY permutation built on it.
//
// Replaces run_radix_pipeline_i64 (src/exec/gpu_sort.rs:1493) and the
// histogram / scatter kernels (src/jit/sort_kernel_radix.rs:489,885), which
// use 4-bit digits (16 passes for int64) with a device-to-host histogram
// copy, a host scan and a host-to-device offset copy between passes.  Here:
//   * one up-front pass histograms all eight 8-bit digits at once (LDS
//     histograms, one global atomic per bin per block);
//   * each digit is then one "onesweep" pass (Adinets & Merrill, 2022): a
//     4096-key tile ranks its keys with wave64 ballot matching (8 ballots
//     per key give each key its set of same-digit peers), looks back across
//     earlier tiles per digit for its global offsets (decoupled look-back,
//     one digit per thread), reorders the tile by digit in LDS and writes
//     runs of equal digits with coalesced stores;
//   * passes whose digit is constant over the whole input are skipped;
//   * for large inputs with three or more varying digits the sort is hybrid
//     (MSD-grouped, LSD-finished): only the two most significant varying
//     digits get global onesweep passes; they leave up to 65536 contiguous
//     buckets of ~n/65536 keys, which one workgroup each sorts by the
//     remaining bytes entirely in LDS and writes back once (Stehle &
//     Jacobsen's hybrid radix sort, 2017).  Three HBM round trips replace
//     eight for 64-bit keys.  Skewed inputs whose largest bucket exceeds the
//     LDS capacity fall back to plain LSD.
// Key transforms follow src/exec/gpu_sort.rs:1070-1180: signed integers flip
// the sign bit, floats use the IEEE-monotonic flip after canonicalising NaN
// (sorted last) and -0.0 (== +0.0), DESC is the bitwise NOT; for int64
// pairs the transform is an xor applied as the first pass loads and undone
// as the last one stores.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "bolt_internal.h"

namespace bolt {

constexpr int kRadixBits = 8;
constexpr int kBins = 1 << kRadixBits;
constexpr int kSortBlock = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortBlock * kSortItems;  // 4096 keys
constexpr int kSortWaves = kSortBlock / 64;
constexpr int kMaxPasses = 9;  // up to 72 key bits

constexpr uint32_t kSFlagAgg = 1u << 30;
constexpr uint32_t kSFlagInc = 2u << 30;
constexpr uint32_t kSValue = (1u << 30) - 1;

// ---- histogram of every digit ---------------------------------------------------

__global__ __launch_bounds__(kSortBlock) void radix_hist_kernel(const uint64_t* __restrict__ keys, int64_t n,
                                                                int passes, uint32_t* __restrict__ hist,
                                                                uint64_t in_xor) {
  __shared__ uint32_t h[kMaxPasses][kBins];
  for (int i = threadIdx.x; i < passes * kBins; i += kSortBlock) (&h[0][0])[i] = 0;
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * kSortBlock;
  for (int64_t i = (int64_t)blockIdx.x * kSortBlock + threadIdx.x; i < n; i += stride) {
    uint64_t k = __builtin_nontemporal_load(keys + i) ^ in_xor;
    for (int p = 0; p < passes; p++) atomicAdd(&h[p][(k >> (p * kRadixBits)) & (kBins - 1)], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < passes * kBins; i += kSortBlock) {
    uint32_t v = (&h[0][0])[i];
    if (v) atomicAdd(&hist[i], v);
  }
}

// exclusive scan of each pass's 256 bins (one block per pass)
__global__ __launch_bounds__(kBins) void radix_scan_kernel(const uint32_t* __restrict__ hist,
                                                           uint32_t* __restrict__ offs) {
  __shared__ uint32_t s[kBins];
  const int p = blockIdx.x, t = threadIdx.x;
  s[t] = hist[p * kBins + t];
  __syncthreads();
  for (int d = 1; d < kBins; d <<= 1) {
    uint32_t v = t >= d ? s[t - d] : 0;
    __syncthreads();
    s[t] += v;
    __syncthreads();
  }
  offs[p * kBins + t] = s[t] - hist[p * kBins + t];
}

// ---- one onesweep pass ------------------------------------------------------------

__global__ __launch_bounds__(kSortBlock) void radix_pass_kernel(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint64_t* __restrict__ kout,
    uint32_t* __restrict__ vout, int64_t n, int shift, const uint32_t* __restrict__ digit_base,
    uint32_t* __restrict__ status, uint32_t* __restrict__ ticket, uint32_t* __restrict__ err, uint64_t in_xor,
    uint64_t out_xor) {
  __shared__ uint64_t sk[kSortTile];
  __shared__ uint32_t sv[kSortTile];
  __shared__ uint32_t wh[kSortWaves][kBins];  // per-wave digit counts, then per-wave prefixes
  __shared__ uint32_t lex[kBins];             // tile-local exclusive digit offsets
  __shared__ uint32_t gb[kBins];              // global base of each digit's run in this tile
  __shared__ uint32_t wtot[kSortWaves];
  __shared__ uint32_t sh_tile;

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) sh_tile = atomicAdd(ticket, 1u);
  for (int i = threadIdx.x; i < kSortWaves * kBins; i += kSortBlock) (&wh[0][0])[i] = 0;
  __syncthreads();
  const int64_t tile = sh_tile;
  const int64_t tbase = tile * kSortTile;
  // wave w owns tile keys [w*64*ITEMS, (w+1)*64*ITEMS), item j covers 64
  // consecutive keys: input order = (wave, item, lane)
  uint64_t key[kSortItems];
  uint32_t val[kSortItems];
  uint16_t rank[kSortItems];
  // branch-free loads (clamped index) so all 2 x ITEMS loads are in flight at
  // once; out-of-range lanes become ~0 (last digit) and are never stored
#pragma unroll
  for (int j = 0; j < kSortItems; j++) {
    const int64_t i = tbase + (int64_t)wave * 64 * kSortItems + j * 64 + lane;
    const int64_t ic = i < n ? i : n - 1;
    key[j] = __builtin_nontemporal_load(kin + ic);
    val[j] = __builtin_nontemporal_load(vin + ic);
  }
#pragma unroll
  for (int j = 0; j < kSortItems; j++) {
    const int64_t i = tbase + (int64_t)wave * 64 * kSortItems + j * 64 + lane;
    key[j] = i < n ? key[j] ^ in_xor : ~0ull;
  }
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
  for (int j = 0; j < kSortItems; j++) {
    const uint32_t d = (uint32_t)(key[j] >> shift) & (kBins - 1);
    uint64_t peers = ~0ull;
#pragma unroll
    for (int b = 0; b < kRadixBits; b++) {
      uint64_t m = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? m : ~m;
    }
    const int leader = __builtin_ctzll(peers);
    uint32_t old = 0;
    if (lane == leader) {
      old = wh[wave][d];
      wh[wave][d] = old + (uint32_t)__popcll(peers);
    }
    old = __shfl(old, leader);
    rank[j] = (uint16_t)(old + (uint32_t)__popcll(peers & lt));
  }
  __syncthreads();
  // thread t owns digit t: wave prefixes and the tile's count, published at
  // once so that later tiles' look-backs find it early
  const int t = threadIdx.x;
  uint32_t cnt = 0;
#pragma unroll
  for (int w = 0; w < kSortWaves; w++) {
    uint32_t c = wh[w][t];
    wh[w][t] = cnt;
    cnt += c;
  }
  // The status word holds its whole payload, so relaxed device-scope atomics
  // suffice (acquire / release at agent scope would invalidate / write back
  // the per-XCD L2).
  uint32_t* st = status + (size_t)tile * kBins + t;
  __hip_atomic_store(st, (tile == 0 ? kSFlagInc : kSFlagAgg) | cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  // exclusive scan of the 256 digit counts (4 waves x 64 digits, shuffles)
  {
    uint32_t inc = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t v = __shfl_up(inc, d);
      if (lane >= d) inc += v;
    }
    if (lane == 63) wtot[wave] = inc;
    lex[t] = inc - cnt;
  }
  __syncthreads();
  {
    uint32_t add = 0;
    for (int k = 0; k < wave; k++) add += wtot[k];
    lex[t] += add;
  }
  // decoupled look-back for digit t, four predecessors per round trip
  uint32_t prefix = 0;
  if (tile > 0) {
    int64_t pt = tile - 1;
    uint32_t spins = 0;
    bool done = false;
    while (!done) {
      uint32_t v[4];
#pragma unroll
      for (int q = 0; q < 4; q++)
        v[q] = pt - q >= 0 ? __hip_atomic_load(status + (size_t)(pt - q) * kBins + t, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT)
                           : kSFlagInc;  // before tile 0: an inclusive zero
      int q = 0;
      for (; q < 4; q++) {
        const uint32_t fl = v[q] & ~kSValue;
        if (fl == 0) break;  // not published yet: wait on this one
        prefix += v[q] & kSValue;
        if (fl == kSFlagInc) {
          done = true;
          break;
        }
      }
      if (done) break;
      pt -= q;
      if (q < 4) {
        if (++spins > (1u << 24)) {
          atomicOr(err, 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __hip_atomic_store(st, kSFlagInc | (prefix + cnt), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  gb[t] = digit_base[t] + prefix;
  __syncthreads();
  // reorder the tile by digit in LDS (stable: rank follows input order)
#pragma unroll
  for (int j = 0; j < kSortItems; j++) {
    int64_t i = tbase + (int64_t)wave * 64 * kSortItems + j * 64 + lane;
    if (i < n) {
      const uint32_t d = (uint32_t)(key[j] >> shift) & (kBins - 1);
      const uint32_t pos = lex[d] + wh[wave][d] + rank[j];
      sk[pos] = key[j];
      sv[pos] = val[j];
    }
  }
  __syncthreads();
  const int64_t valid = n - tbase < kSortTile ? n - tbase : kSortTile;
#pragma unroll
  for (int j = 0; j < kSortItems; j++) {
    const int i = j * kSortBlock + threadIdx.x;
    if (i < valid) {
      const uint64_t k = sk[i];
      const uint32_t d = (uint32_t)(k >> shift) & (kBins - 1);
      const uint64_t dst = (uint64_t)gb[d] + (uint32_t)i - lex[d];
      kout[dst] = k ^ out_xor;
      vout[dst] = sv[i];
    }
  }
}

__global__ void copy_pairs_kernel(const uint64_t* __restrict__ ki, const uint32_t* __restrict__ vi,
                                  uint64_t* __restrict__ ko, uint32_t* __restrict__ vo, int64_t n, uint64_t x) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    ko[i] = ki[i] ^ x;
    vo[i] = vi[i];
  }
}

// ---- hybrid tail: per-bucket LDS sort ------------------------------------------------
//
// After two stable global passes on the two most significant varying digits
// (p1 < p2), the keys form up to 65536 contiguous buckets, one per (digit p2,
// digit p1) pair, and every bit above byte p1 other than those two digits is
// constant.  One workgroup then sorts each bucket entirely in LDS by the
// remaining low bytes (stable LSD, skipping bytes that are constant within the
// bucket) and writes it once: three HBM round trips instead of one per digit.

constexpr int kLocBlock = 512;
constexpr int kLocItems = 11;
constexpr int kLocMax = kLocBlock * kLocItems;  // 5632 keys per bucket
constexpr int kLocWaves = kLocBlock / 64;
constexpr int kBuckets = 1 << 16;

__device__ __forceinline__ uint32_t bucket_of(uint64_t k, int s1, int s2) {
  return (uint32_t)((((k >> s2) & 255) << 8) | ((k >> s1) & 255));
}

// start[b] = first index whose bucket is >= b (b = 0 .. 65536); *max_size =
// the largest bucket.  keys are sorted by bucket.
__global__ void bucket_bounds_kernel(const uint64_t* __restrict__ keys, int64_t n, int s1, int s2,
                                     uint32_t* __restrict__ start, uint32_t* __restrict__ max_size) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > kBuckets) return;
  auto lower = [&](uint32_t target) -> int64_t {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (bucket_of(keys[mid], s1, s2) < target) lo = mid + 1;
      else hi = mid;
    }
    return lo;
  };
  const int64_t s0 = b < kBuckets ? lower((uint32_t)b) : n;
  start[b] = (uint32_t)s0;
  if (b < kBuckets) {
    const int64_t e = b + 1 < kBuckets ? lower((uint32_t)(b + 1)) : n;
    atomicMax(max_size, (uint32_t)(e - s0));
  }
}

__global__ __launch_bounds__(kLocBlock) void local_sort_kernel(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint64_t* __restrict__ kout,
    uint32_t* __restrict__ vout, const uint32_t* __restrict__ start, int low_bits, uint32_t byte_mask,
    uint64_t out_xor, uint32_t* __restrict__ err) {
  __shared__ uint64_t sk[kLocMax];
  __shared__ uint32_t sv[kLocMax];
  __shared__ uint32_t wh[kLocWaves][kBins];
  __shared__ uint32_t lex[kBins];
  __shared__ uint32_t wtot[4];
  __shared__ uint32_t wtot8[kLocWaves];
  __shared__ uint32_t sh_max;
  __shared__ uint64_t red[2][kLocWaves];
  const uint32_t s0 = start[blockIdx.x];
  const int size = (int)(start[blockIdx.x + 1] - s0);
  if (size <= 0) return;
  if (size > kLocMax) {  // the host checks the largest bucket first; never write out of range
    if (threadIdx.x == 0) atomicOr(err, 2u);
    return;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t low_mask = (1ull << low_bits) - 1;  // low_bits <= 48
  const uint64_t top = kin[s0] & ~low_mask;           // constant over the bucket
  // element order: (wave, item, lane); w = (low bits << 16) | element index
  uint64_t w[kLocItems];
  uint32_t vv[kLocItems];
  uint64_t a = ~0ull, o = 0;
  // branch-free (clamped) loads first, so they are all in flight together
#pragma unroll
  for (int j = 0; j < kLocItems; j++) {
    const int i = wave * 64 * kLocItems + j * 64 + lane;
    const int ic = i < size ? i : size - 1;
    w[j] = kin[s0 + ic];
    vv[j] = vin[s0 + ic];
  }
#pragma unroll
  for (int j = 0; j < kLocItems; j++) {
    const int i = wave * 64 * kLocItems + j * 64 + lane;
    const uint64_t low = w[j] & low_mask;
    const bool in = i < size;
    a &= in ? low : ~0ull;
    o |= in ? low : 0ull;
    w[j] = in ? ((low << 16) | (uint64_t)i) : ~0ull;  // out-of-range: after every element of the bucket
    if (in) sv[i] = vv[j];
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    a &= __shfl_xor(a, m);
    o |= __shfl_xor(o, m);
  }
  if (lane == 0) {
    red[0][wave] = a;
    red[1][wave] = o;
  }
  __syncthreads();
  uint64_t varying = 0;
  {
    uint64_t ra = ~0ull, ro = 0;
    for (int k = 0; k < kLocWaves; k++) {
      ra &= red[0][k];
      ro |= red[1][k];
    }
    varying = ra ^ ro;
  }
  const int t = threadIdx.x;
  // ---- fast path: one counting pass on the 11 bits below the bucket's top
  // varying bit (sub-buckets of ~size/2048 keys), then an insertion sort of
  // each sub-bucket by the whole packed word; its low 16 bits are the
  // element's position, so equal keys keep their input order (stable) even
  // though the atomic scatter is not.  Buckets with a sub-bucket above 64
  // keys (heavy duplicates) take the LSD passes below instead.
  if (varying != 0) {
    const int h = 63 - __clzll(varying);
    const int dsh = 16 + (h >= 10 ? h - 10 : 0);
    uint32_t* cnt = &wh[0][0];  // 2048 counters
    for (int i = t; i < kLocWaves * kBins; i += kLocBlock) cnt[i] = 0;
    if (t == 0) sh_max = 0;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kLocItems; j++)
      if (w[j] != ~0ull) atomicAdd(&cnt[(uint32_t)(w[j] >> dsh) & 2047u], 1u);
    __syncthreads();
    {
      const uint32_t c0 = cnt[4 * t], c1 = cnt[4 * t + 1], c2 = cnt[4 * t + 2], c3 = cnt[4 * t + 3];
      const uint32_t local = c0 + c1 + c2 + c3;
      uint32_t inc = local;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t v = __shfl_up(inc, d);
        if (lane >= d) inc += v;
      }
      if (lane == 63) wtot8[wave] = inc;
      const uint32_t mx = max(max(c0, c1), max(c2, c3));
      if (mx) atomicMax(&sh_max, mx);
      __syncthreads();
      uint32_t carry = 0;
      for (int k = 0; k < wave; k++) carry += wtot8[k];
      const uint32_t ex = carry + inc - local;
      cnt[4 * t] = ex;
      cnt[4 * t + 1] = ex + c0;
      cnt[4 * t + 2] = ex + c0 + c1;
      cnt[4 * t + 3] = ex + c0 + c1 + c2;
    }
    __syncthreads();
    if (sh_max <= 64u) {  // block-uniform
#pragma unroll
      for (int j = 0; j < kLocItems; j++)
        if (w[j] != ~0ull) sk[atomicAdd(&cnt[(uint32_t)(w[j] >> dsh) & 2047u], 1u)] = w[j];
      __syncthreads();
      // cnt[d] is now the end of sub-bucket d (= the start of d + 1)
      for (int d = t; d < 2048; d += kLocBlock) {
        const int b = d ? (int)cnt[d - 1] : 0, e = (int)cnt[d];
        for (int x = b + 1; x < e; x++) {
          const uint64_t v = sk[x];
          int y = x - 1;
          while (y >= b && sk[y] > v) {
            sk[y + 1] = sk[y];
            y--;
          }
          sk[y + 1] = v;
        }
      }
      __syncthreads();
      for (int i = t; i < size; i += kLocBlock) {
        const uint64_t ww = sk[i];
        kout[s0 + i] = (top | (ww >> 16)) ^ out_xor;
        vout[s0 + i] = sv[(uint32_t)(ww & 0xFFFF)];
      }
      return;
    }
  }
  // ---- general path: stable LSD over the bytes that vary in this bucket
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  for (int q = 0; q * 8 < low_bits; q++) {
    if (!((byte_mask >> q) & 1u) || ((varying >> (8 * q)) & 255) == 0) continue;  // block-uniform
    const int shift = 16 + 8 * q;
    for (int i = t; i < kLocWaves * kBins; i += kLocBlock) (&wh[0][0])[i] = 0;
    __syncthreads();
    // rank within the wave: peers by ballot matching; the per-wave digit
    // counters are bumped with returning LDS atomics issued for all items
    // back to back (LDS keeps a wave's operations in order, so ranks stay
    // stable) and read once afterwards
    uint32_t rank[kLocItems], oldv[kLocItems];
    int lead[kLocItems];
#pragma unroll
    for (int j = 0; j < kLocItems; j++) {
      const uint32_t d = (uint32_t)(w[j] >> shift) & (kBins - 1);
      uint64_t peers = ~0ull;
#pragma unroll
      for (int bit = 0; bit < kRadixBits; bit++) {
        const uint64_t m = __ballot((d >> bit) & 1u);
        peers &= ((d >> bit) & 1u) ? m : ~m;
      }
      lead[j] = __builtin_ctzll(peers);
      rank[j] = (uint32_t)__popcll(peers & lt);
      oldv[j] = 0;
      if (lane == lead[j]) oldv[j] = atomicAdd(&wh[wave][d], (uint32_t)__popcll(peers));
    }
#pragma unroll
    for (int j = 0; j < kLocItems; j++) rank[j] += __shfl(oldv[j], lead[j]);
    __syncthreads();
    // digit t: exclusive offsets per wave, then an exclusive scan over the
    // 256 digits (waves 0..3 scan 64 digits each with shuffles)
    if (t < kBins) {
      uint32_t cnt = 0;
#pragma unroll
      for (int k = 0; k < kLocWaves; k++) {
        const uint32_t c = wh[k][t];
        wh[k][t] = cnt;
        cnt += c;
      }
      uint32_t inc = cnt;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t v = __shfl_up(inc, d);
        if (lane >= d) inc += v;
      }
      if (lane == 63) wtot[wave] = inc;
      lex[t] = inc - cnt;
    }
    __syncthreads();
    if (t < kBins) {
      uint32_t add = 0;
      for (int k = 0; k < (t >> 6); k++) add += wtot[k];
      lex[t] += add;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kLocItems; j++) {
      const uint32_t d = (uint32_t)(w[j] >> shift) & (kBins - 1);
      sk[lex[d] + wh[wave][d] + rank[j]] = w[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kLocItems; j++) w[j] = sk[wave * 64 * kLocItems + j * 64 + lane];
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < kLocItems; j++) {
    const int i = wave * 64 * kLocItems + j * 64 + lane;
    if (i < size) {
      kout[s0 + i] = (top | (w[j] >> 16)) ^ out_xor;
      vout[s0 + i] = sv[(uint32_t)(w[j] & 0xFFFF)];
    }
  }
}

// Sort (keys_in, vals_in) ascending by the low `key_bits` bits of
// (key ^ in_xor), stable, writing (key image ^ out_xor, value) to keys_out /
// vals_out.  Inputs are not modified and must not alias the outputs.
int radix_sort_u64_pairs(hipStream_t s, const uint64_t* keys_in, const uint32_t* vals_in, int64_t n,
                         uint64_t* keys_out, uint32_t* vals_out, int key_bits, uint64_t in_xor, uint64_t out_xor) {
  if (n <= 0) return BOLT_OK;
  if (n >= (1ll << 30)) return set_error(BOLT_ERR_PLAN, "radix sort: %lld keys exceeds the 2^30 limit", (long long)n);
  const int passes = (key_bits + kRadixBits - 1) / kRadixBits;
  if (passes < 1 || passes > kMaxPasses) return set_error(BOLT_ERR_PLAN, "radix sort: bad key width %d", key_bits);
  DevBuf<uint32_t> hist, offs, misc;
  BOLT_TRY(hist.alloc((size_t)passes * kBins, s));
  BOLT_TRY(offs.alloc((size_t)passes * kBins, s));
  BOLT_TRY(misc.alloc(4, s));
  BOLT_TRY(fill_u32(hist.ptr, 0, (size_t)passes * kBins, s));
  radix_hist_kernel<<<grid_for(n, kSortBlock, 16, 2048), kSortBlock, 0, s>>>(keys_in, n, passes, hist.ptr, in_xor);
  radix_scan_kernel<<<passes, kBins, 0, s>>>(hist.ptr, offs.ptr);
  BOLT_HIP(hipGetLastError());
  std::vector<uint32_t> h((size_t)passes * kBins);
  BOLT_HIP(hipMemcpyAsync(h.data(), hist.ptr, h.size() * 4, hipMemcpyDeviceToHost, s));
  BOLT_HIP(hipStreamSynchronize(s));
  std::vector<int> todo;
  for (int p = 0; p < passes; p++) {
    bool trivial = false;
    for (int b = 0; b < kBins; b++)
      if (h[(size_t)p * kBins + b] == (uint32_t)n) trivial = true;
    if (!trivial) todo.push_back(p);
  }
  if (todo.empty()) {
    copy_pairs_kernel<<<grid_for(n, 256, 4), 256, 0, s>>>(keys_in, vals_in, keys_out, vals_out, n, in_xor ^ out_xor);
    BOLT_HIP(hipGetLastError());
    return BOLT_OK;
  }
  const int64_t tiles = (n + kSortTile - 1) / kSortTile;
  DevBuf<uint64_t> ktmp;
  DevBuf<uint32_t> vtmp, status;
  BOLT_TRY(ktmp.alloc((size_t)n, s));
  BOLT_TRY(vtmp.alloc((size_t)n, s));
  BOLT_TRY(status.alloc((size_t)tiles * kBins, s));
  auto pass = [&](const uint64_t* ksrc, const uint32_t* vsrc, uint64_t* kdst, uint32_t* vdst, int p, uint64_t ix,
                  uint64_t ox) -> int {
    BOLT_TRY(fill_u32(status.ptr, 0, (size_t)tiles * kBins, s));
    BOLT_TRY(fill_u32(misc.ptr, 0, 4, s));
    radix_pass_kernel<<<(unsigned)tiles, kSortBlock, 0, s>>>(ksrc, vsrc, kdst, vdst, n, p * kRadixBits,
                                                             offs.ptr + (size_t)p * kBins, status.ptr, misc.ptr,
                                                             misc.ptr + 1, ix, ox);
    BOLT_HIP(hipGetLastError());
    return BOLT_OK;
  };
  auto check_lookback = [&]() -> int {
    uint32_t e = 0;
    BOLT_HIP(hipMemcpyAsync(&e, misc.ptr + 1, 4, hipMemcpyDeviceToHost, s));
    BOLT_HIP(hipStreamSynchronize(s));
    if (e) return set_error(BOLT_ERR_OTHER, "radix sort look-back timed out");
    return BOLT_OK;
  };
  const int np = (int)todo.size();
  // hybrid: two global passes on the top varying digits, then LDS bucket sorts
  if (np >= 3 && n >= (1 << 20) && getenv("BOLT_SORT_LSD") == nullptr) {
    const int p1 = todo[np - 2], p2 = todo[np - 1];
    BOLT_TRY(pass(keys_in, vals_in, keys_out, vals_out, p1, in_xor, 0));
    BOLT_TRY(pass(keys_out, vals_out, ktmp.ptr, vtmp.ptr, p2, 0, 0));
    DevBuf<uint32_t> starts;
    BOLT_TRY(starts.alloc(kBuckets + 1, s));
    BOLT_TRY(fill_u32(misc.ptr + 2, 0, 2, s));
    bucket_bounds_kernel<<<(kBuckets + 1 + 255) / 256, 256, 0, s>>>(ktmp.ptr, n, p1 * kRadixBits, p2 * kRadixBits,
                                                                  starts.ptr, misc.ptr + 2);
    BOLT_HIP(hipGetLastError());
    uint32_t mx[2] = {0, 0};
    BOLT_HIP(hipMemcpyAsync(mx, misc.ptr + 1, 8, hipMemcpyDeviceToHost, s));
    BOLT_HIP(hipStreamSynchronize(s));
    if (mx[0]) return set_error(BOLT_ERR_OTHER, "radix sort look-back timed out");
    if (mx[1] <= (uint32_t)kLocMax) {
      uint32_t byte_mask = 0;
      for (int q = 0; q < np - 2; q++) byte_mask |= 1u << todo[q];
      BOLT_TRY(fill_u32(misc.ptr + 2, 0, 1, s));
      local_sort_kernel<<<kBuckets, kLocBlock, 0, s>>>(ktmp.ptr, vtmp.ptr, keys_out, vals_out, starts.ptr,
                                                       p1 * kRadixBits, byte_mask, out_xor, misc.ptr + 2);
      BOLT_HIP(hipGetLastError());
      uint32_t e2 = 0;
      BOLT_HIP(hipMemcpyAsync(&e2, misc.ptr + 2, 4, hipMemcpyDeviceToHost, s));
      BOLT_HIP(hipStreamSynchronize(s));
      if (e2) return set_error(BOLT_ERR_OTHER, "radix sort: bucket larger than the LDS sort");
      return BOLT_OK;
    }
    // a bucket too large for LDS (skewed keys): plain LSD from the input
  }
  // LSD: one pass per varying digit, ping-pong so the last pass writes the output
  const uint64_t* ksrc = keys_in;
  const uint32_t* vsrc = vals_in;
  for (int q = 0; q < np; q++) {
    const bool to_out = ((np - 1 - q) % 2) == 0;
    uint64_t* kdst = to_out ? keys_out : ktmp.ptr;
    uint32_t* vdst = to_out ? vals_out : vtmp.ptr;
    BOLT_TRY(pass(ksrc, vsrc, kdst, vdst, todo[q], q == 0 ? in_xor : 0, q == np - 1 ? out_xor : 0));
    ksrc = kdst;
    vsrc = vdst;
  }
  return check_lookback();
}

// ---- key transforms ------------------------------------------------------------------

// Ascending radix image of column value `row` (bits used: 32 or 64).
__device__ __forceinline__ uint64_t radix_image(const void* d, int dt, int64_t row) {
  switch (dt) {
    case BOLT_I32: case BOLT_DATE32: return (uint64_t)(static_cast<const uint32_t*>(d)[row] ^ 0x80000000u);
    case BOLT_I64: return static_cast<const uint64_t*>(d)[row] ^ 0x8000000000000000ull;
    case BOLT_F32: {
      uint32_t b = static_cast<const uint32_t*>(d)[row];
      if ((b & 0x7FFFFFFFu) > 0x7F800000u) b = 0x7FC00000u;  // canonical NaN
      else if ((b & 0x7FFFFFFFu) == 0) b = 0;
      uint32_t m = (uint32_t)((int32_t)b >> 31) | 0x80000000u;  // gpu_sort.rs:1150
      return (uint64_t)(b ^ m);
    }
    case BOLT_F64: {
      uint64_t b = static_cast<const uint64_t*>(d)[row];
      if ((b & 0x7FFFFFFFFFFFFFFFull) > 0x7FF0000000000000ull) b = 0x7FF8000000000000ull;
      else if ((b & 0x7FFFFFFFFFFFFFFFull) == 0) b = 0;
      uint64_t m = (uint64_t)((int64_t)b >> 63) | 0x8000000000000000ull;  // gpu_sort.rs:1169
      return b ^ m;
    }
    case BOLT_BOOL: return static_cast<const uint8_t*>(d)[row] ? 1 : 0;
    case BOLT_DEC128: return static_cast<const uint64_t*>(d)[2 * row];  // low word, unsigned (mode 2: high word)
  }
  return 0;
}

// key[i] = image of column value at perm[i]; mode 0 = value (DESC = NOT),
// mode 1 = null rank only (0 first / 1 last per nulls_first), mode 2 = the
// signed high word of a Decimal128 (its low word, unsigned, is mode 0).
__global__ void order_key_kernel(const void* __restrict__ d, const uint8_t* __restrict__ valid, int dt, int bits,
                                 int desc, int nulls_first, int mode, const uint32_t* __restrict__ perm,
                                 uint64_t* __restrict__ key, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const uint64_t mask = bits >= 64 ? ~0ull : ((1ull << bits) - 1);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    int64_t row = perm[i];
    if (mode == 1) {
      bool isnull = !bit_valid(valid, row);
      key[i] = isnull ? (nulls_first ? 0 : 1) : (nulls_first ? 1 : 0);
    } else {
      uint64_t k = 0;
      if (bit_valid(valid, row))
        k = mode == 2 ? static_cast<const uint64_t*>(d)[2 * row + 1] ^ 0x8000000000000000ull : radix_image(d, dt, row);
      key[i] = (desc ? ~k : k) & mask;
    }
  }
}

// ---- ORDER BY + LIMIT: radix select on the leading key ---------------------------------
//
// `coarse_of` maps the leading key of a row to a u64 that never decreases
// along the full sort order (NULL rank, direction and value in one word; a
// Decimal128 by its signed high word).  The K-th smallest image is T; every
// row among the first K of the order has an image <= T, so sorting only
// those rows by all keys (and by row id, for the stability of the full sort)
// gives the same first K rows as sorting everything.
struct TopKey {
  const void* d;
  const uint8_t* v;
  int dt, bits, desc, nulls_first;
};

// Raw bits of row i's key value: W = byte width (16 = Decimal128, whose
// signed high word is the part of the value the select looks at).
template <int W>
__device__ __forceinline__ uint64_t load_raw(const void* d, int64_t i) {
  if constexpr (W == 1) return static_cast<const uint8_t*>(d)[i];
  else if constexpr (W == 4) return static_cast<const uint32_t*>(d)[i];
  else if constexpr (W == 8) return static_cast<const uint64_t*>(d)[i];
  else return static_cast<const uint64_t*>(d)[2 * i + 1];
}

// Ascending image of raw key bits (the transforms of radix_image).
__device__ __forceinline__ uint64_t raw_image(uint64_t b, int dt) {
  switch (dt) {
    case BOLT_I32: case BOLT_DATE32: return (uint64_t)((uint32_t)b ^ 0x80000000u);
    case BOLT_F32: {
      uint32_t x = (uint32_t)b;
      if ((x & 0x7FFFFFFFu) > 0x7F800000u) x = 0x7FC00000u;
      else if ((x & 0x7FFFFFFFu) == 0) x = 0;
      return (uint64_t)(x ^ ((uint32_t)((int32_t)x >> 31) | 0x80000000u));
    }
    case BOLT_F64: {
      if ((b & 0x7FFFFFFFFFFFFFFFull) > 0x7FF0000000000000ull) b = 0x7FF8000000000000ull;
      else if ((b & 0x7FFFFFFFFFFFFFFFull) == 0) b = 0;
      return b ^ ((uint64_t)((int64_t)b >> 63) | 0x8000000000000000ull);
    }
    case BOLT_BOOL: return b ? 1 : 0;
    default: return b ^ 0x8000000000000000ull;  // Int64; Decimal128 high word
  }
}

__device__ __forceinline__ uint64_t coarse_of(const TopKey& k, uint64_t raw, bool valid) {
  uint64_t x = 0;
  if (valid) {
    x = raw_image(raw, k.dt);
    if (k.desc) x = ~x & (k.bits >= 64 ? ~0ull : ((1ull << k.bits) - 1));
  }
  if (k.v == nullptr) return x;
  // every image stays below 2^bits, so the select can start at bit `bits`
  if (!valid) return k.nulls_first ? 0ull : (k.bits >= 64 ? ~0ull : (1ull << k.bits) - 1);
  return k.nulls_first ? (x >> 1) + 1 : x >> 1;
}

constexpr int kSelBits = 14;
constexpr int kSelBins = 1 << kSelBits;  // 64 KB of LDS counters
constexpr int kSelBlock = 1024;
constexpr int kSelUnroll = 4;  // rows in flight per lane

// Per-workgroup histogram of digit (image >> shift) & (2^width - 1) over the
// rows whose image >> (shift + width) equals `prefix` (every row when
// shift + width = 64), written to part[block][bin] with plain stores (a
// global atomic per bin per workgroup would cost more than the row pass).
// With `spec` non-null (the first pass) the rows whose digit is at most
// spec_bound are also appended to spec (unordered, up to spec_cap): a guess
// at the candidates that saves the collect pass when the K-th row's digit
// turns out to be within the bound.
template <int W>
__global__ __launch_bounds__(kSelBlock) void topk_hist_kernel(TopKey k, int64_t n, int shift, int width,
                                                              uint64_t prefix, uint32_t* __restrict__ part,
                                                              uint64_t spec_bound, uint32_t* __restrict__ spec,
                                                              unsigned long long* __restrict__ spec_count,
                                                              int64_t spec_cap) {
  __shared__ uint32_t sh[kSelBins];
  __shared__ int spec_full;  // set once this workgroup has seen the guess buffer full
  const int nb = 1 << width;
  for (int b = threadIdx.x; b < nb; b += kSelBlock) sh[b] = 0;
  if (threadIdx.x == 0) spec_full = 0;
  __syncthreads();
  const int hi = shift + width;
  const uint64_t dmask = (uint64_t)nb - 1;
  const int64_t stride = (int64_t)gridDim.x * kSelBlock;
  for (int64_t i0 = (int64_t)blockIdx.x * kSelBlock + threadIdx.x; i0 < n; i0 += stride * kSelUnroll) {
    uint64_t x[kSelUnroll];
#pragma unroll
    for (int u = 0; u < kSelUnroll; u++) {
      const int64_t i = i0 + u * stride;
      x[u] = i < n ? load_raw<W>(k.d, i) : 0;
    }
#pragma unroll
    for (int u = 0; u < kSelUnroll; u++) {
      const int64_t i = i0 + u * stride;
      bool take = false;
      if (i < n) {
        const uint64_t c = coarse_of(k, x[u], bit_valid(k.v, i));
        if (hi >= 64 || (c >> hi) == prefix) atomicAdd(&sh[(c >> shift) & dmask], 1u);
        // a guess that has already overflowed is abandoned: the collect pass
        // runs instead, so further appends (one atomic per wave) are waste
        take = spec != nullptr && (c >> shift) <= spec_bound && !*(volatile int*)&spec_full;
      }
      const uint64_t m = __ballot(take);  // lanes still in the loop
      if (m) {
        const int lane = threadIdx.x & 63;
        const int leader = __builtin_ctzll(m);
        unsigned long long b = 0;
        if (lane == leader) {
          b = atomicAdd(spec_count, (unsigned long long)__popcll(m));
          if ((int64_t)b + __popcll(m) > spec_cap) *(volatile int*)&spec_full = 1;
        }
        b = __shfl(b, leader);
        if (take) {
          const int64_t pos = (int64_t)b + __popcll(m & ((1ull << lane) - 1));
          if (pos < spec_cap) spec[pos] = (uint32_t)i;
        }
      }
    }
  }
  __syncthreads();
  uint32_t* out = part + (int64_t)blockIdx.x * nb;
  for (int b = threadIdx.x; b < nb; b += kSelBlock) out[b] = sh[b];
}

// hist[b] += sum of part[w][b] over this launch row's share of workgroups
// (hist starts at zero; gridDim.y shares split the workgroups)
__global__ __launch_bounds__(256) void topk_hist_sum_kernel(const uint32_t* __restrict__ part, int blocks, int nb,
                                                            uint32_t* __restrict__ hist) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= nb) return;
  const int per = (blocks + (int)gridDim.y - 1) / (int)gridDim.y;
  const int w0 = (int)blockIdx.y * per;
  const int w1 = w0 + per < blocks ? w0 + per : blocks;
  uint32_t t = 0;
  for (int w = w0; w < w1; w++) t += part[(int64_t)w * nb + b];
  if (t) atomicAdd(&hist[b], t);
}

// Appends (unordered) the rows whose image >> shift <= bound.
template <int W>
__global__ __launch_bounds__(256) void topk_collect_kernel(TopKey k, int64_t n, int shift, uint64_t bound,
                                                           uint32_t* __restrict__ out,
                                                           unsigned long long* __restrict__ count, int64_t cap) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  const int64_t end = (n + 255) / 256 * 256;  // block-uniform trip count for the ballot
  const int lane = threadIdx.x & 63;
  for (int64_t i0 = (int64_t)blockIdx.x * 256 + threadIdx.x; i0 < end; i0 += stride * kSelUnroll) {
    uint64_t x[kSelUnroll];
#pragma unroll
    for (int u = 0; u < kSelUnroll; u++) {
      const int64_t i = i0 + u * stride;
      x[u] = i < n ? load_raw<W>(k.d, i) : 0;
    }
    bool take[kSelUnroll];
#pragma unroll
    for (int u = 0; u < kSelUnroll; u++) {
      const int64_t i = i0 + u * stride;
      take[u] = i < n && (shift >= 64 ? 0ull : coarse_of(k, x[u], bit_valid(k.v, i)) >> shift) <= bound;
    }
#pragma unroll
    for (int u = 0; u < kSelUnroll; u++) {
      const uint64_t m = __ballot(take[u]);
      if (m == 0) continue;
      const int leader = __builtin_ctzll(m);
      unsigned long long base = 0;
      if (lane == leader) base = atomicAdd(count, (unsigned long long)__popcll(m));
      base = __shfl(base, leader);
      if (take[u]) {
        const int64_t pos = (int64_t)base + __popcll(m & ((1ull << lane) - 1));
        if (pos < cap) out[pos] = (uint32_t)(i0 + u * stride);
      }
    }
  }
}

// ---- small final sets: narrowed over the candidates, sorted in one workgroup ----------
constexpr int kSmallSort = 4096;  // (image, row) pairs sorted in LDS by one workgroup

// hist[digit] over the candidates whose image >> (shift + width) == prefix
template <int W>
__global__ __launch_bounds__(256) void topk_cand_hist_kernel(TopKey k, const uint32_t* __restrict__ cand, int64_t c,
                                                             int shift, int width, uint64_t prefix,
                                                             uint32_t* __restrict__ hist) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= c) return;
  const uint64_t x = coarse_of(k, load_raw<W>(k.d, cand[j]), true);
  const int hi = shift + width;
  if (hi < 64 && (x >> hi) != prefix) return;
  atomicAdd(&hist[(x >> shift) & ((1ull << width) - 1)], 1u);
}

// Candidates tied on the whole image `img`: histogram of their row ids'
// bits [shift, shift + width) among those whose higher row-id bits are
// `rprefix` (the order among equal keys is by row id).
template <int W>
__global__ __launch_bounds__(256) void topk_tie_hist_kernel(TopKey k, const uint32_t* __restrict__ cand, int64_t c,
                                                            uint64_t img, int shift, int width, uint32_t rprefix,
                                                            uint32_t* __restrict__ hist) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= c) return;
  const uint32_t r = cand[j];
  if (coarse_of(k, load_raw<W>(k.d, r), true) != img) return;
  const int hi = shift + width;
  if (hi < 32 && (r >> hi) != rprefix) return;
  atomicAdd(&hist[(r >> shift) & ((1u << width) - 1)], 1u);
}

// appends the candidates whose image >> shift <= bound; with shift 0 and
// rmax < 0xFFFFFFFF, of the rows whose image equals `bound` only those with
// row id <= rmax
template <int W>
__global__ __launch_bounds__(256) void topk_cand_collect_kernel(TopKey k, const uint32_t* __restrict__ cand,
                                                                int64_t c, int shift, uint64_t bound,
                                                                uint32_t* __restrict__ out,
                                                                unsigned long long* __restrict__ count, int64_t cap,
                                                                uint32_t rmax = 0xFFFFFFFFu) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int lane = threadIdx.x & 63;
  bool take = false;
  uint32_t r = 0;
  if (j < c) {
    r = cand[j];
    const uint64_t x = shift >= 64 ? 0ull : coarse_of(k, load_raw<W>(k.d, r), true) >> shift;
    take = x < bound || (x == bound && r <= rmax);
  }
  const uint64_t m = __ballot(take);
  if (m == 0) return;
  const int leader = __builtin_ctzll(m);
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(count, (unsigned long long)__popcll(m));
  base = __shfl(base, leader);
  if (take) {
    const int64_t pos = (int64_t)base + __popcll(m & ((1ull << lane) - 1));
    if (pos < cap) out[pos] = r;
  }
}

// Sorts c <= kSmallSort rows by (image, row id) with a bitonic network in
// LDS (padded to a power of two with pairs that sort last) and writes rows
// [first, first + m) of the order.
template <int W>
__global__ __launch_bounds__(1024) void topk_small_sort_kernel(TopKey k, const uint32_t* __restrict__ rows, int c,
                                                               int npow, int first, int m,
                                                               uint32_t* __restrict__ out) {
  __shared__ uint64_t key[kSmallSort];
  __shared__ uint32_t row[kSmallSort];
  for (int i = threadIdx.x; i < npow; i += 1024) {
    if (i < c) {
      const uint32_t r = rows[i];
      key[i] = coarse_of(k, load_raw<W>(k.d, r), true);
      row[i] = r;
    } else {
      key[i] = ~0ull;
      row[i] = 0xFFFFFFFFu;
    }
  }
  __syncthreads();
  for (int size = 2; size <= npow; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = threadIdx.x; i < npow; i += 1024) {
        const int j = i ^ stride;
        if (j > i) {
          const bool up = (i & size) == 0;
          const bool gt = key[i] > key[j] || (key[i] == key[j] && row[i] > row[j]);
          if (gt == up) {
            const uint64_t tk = key[i];
            key[i] = key[j];
            key[j] = tk;
            const uint32_t tr = row[i];
            row[i] = row[j];
            row[j] = tr;
          }
        }
      }
      __syncthreads();
    }
  for (int j = threadIdx.x; j < m; j += 1024) out[j] = row[first + j];
}

// out[j] = rows[perm[first + j]]
__global__ void topk_emit_kernel(const uint32_t* __restrict__ rows, const uint32_t* __restrict__ perm, int64_t first,
                                 int64_t m, uint32_t* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m) out[j] = rows[perm[first + j]];
}

__global__ void iota_kernel(uint32_t* __restrict__ out, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = (uint32_t)i;
}

}  // namespace bolt

using namespace bolt;

extern "C" int bolt_sort_pairs_i64(bolt_stream_t stream, const int64_t* keys_in, const uint32_t* vals_in, int64_t n,
                                   int32_t descending, int64_t* keys_out, uint32_t* vals_out) {
  BOLT_CHECK_DEVICE();
  clear_error();
  hipStream_t s = (hipStream_t)stream;
  if (n < 0) return set_error(BOLT_ERR_PLAN, "bolt_sort_pairs_i64: negative length");
  if (n == 0) return BOLT_OK;
  // radix image: sign bit flipped (src/exec/gpu_sort.rs:1091), all bits
  // inverted for DESC; applied on load by the first pass and undone on the
  // final store, so no separate transform kernels run
  const uint64_t x = descending ? 0x7FFFFFFFFFFFFFFFull : 0x8000000000000000ull;
  const uint64_t* kin = reinterpret_cast<const uint64_t*>(keys_in);
  const uint32_t* vin = vals_in;
  DevBuf<uint64_t> kcopy;
  DevBuf<uint32_t> vcopy;
  if ((const void*)keys_in == (const void*)keys_out || (const void*)vals_in == (const void*)vals_out) {
    BOLT_TRY(kcopy.alloc((size_t)n, s));
    BOLT_TRY(vcopy.alloc((size_t)n, s));
    BOLT_HIP(hipMemcpyAsync(kcopy.ptr, keys_in, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
    BOLT_HIP(hipMemcpyAsync(vcopy.ptr, vals_in, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
    kin = kcopy.ptr;
    vin = vcopy.ptr;
  }
  prof_begin(s);
  BOLT_TRY(radix_sort_u64_pairs(s, kin, vin, n, reinterpret_cast<uint64_t*>(keys_out), vals_out, 64, x, x));
  prof_end(s);
  return BOLT_OK;
}

extern "C" int bolt_sort_indices(bolt_stream_t stream, const bolt_column_t* keys, int32_t n_keys,
                                 const int32_t* descending, const int32_t* nulls_first, uint32_t* out_rows) {
  BOLT_CHECK_DEVICE();
  clear_error();
  hipStream_t s = (hipStream_t)stream;
  if (n_keys < 1 || keys == nullptr) return set_error(BOLT_ERR_PLAN, "bolt_sort_indices: no sort keys");
  const int64_t n = keys[0].length;
  for (int k = 0; k < n_keys; k++) {
    if (keys[k].length != n) return set_error(BOLT_ERR_PLAN, "bolt_sort_indices: key columns differ in length");
    int dt = keys[k].dtype;
    if (!(dt == BOLT_I32 || dt == BOLT_I64 || dt == BOLT_F32 || dt == BOLT_F64 || dt == BOLT_DATE32 || dt == BOLT_BOOL ||
          dt == BOLT_DEC128))
      return set_error(BOLT_ERR_UNSUPPORTED, "bolt_sort_indices: sort key dtype %s not supported", dtype_name(dt));
  }
  if (n == 0) return BOLT_OK;
  if (n >= (1ll << 30)) return set_error(BOLT_ERR_PLAN, "bolt_sort_indices: %lld rows exceeds 2^30", (long long)n);
  DevBuf<uint64_t> k0, k1;
  DevBuf<uint32_t> p0;
  BOLT_TRY(k0.alloc((size_t)n, s));
  BOLT_TRY(k1.alloc((size_t)n, s));
  BOLT_TRY(p0.alloc((size_t)n, s));
  iota_kernel<<<grid_for(n, 256, 4), 256, 0, s>>>(out_rows, n);
  BOLT_HIP(hipGetLastError());
  // least significant key first; each step is a stable sort of the current
  // permutation, so earlier (more significant) keys win ties last
  for (int k = n_keys - 1; k >= 0; k--) {
    const bolt_column_t& c = keys[k];
    const int bits = c.dtype == BOLT_BOOL ? 1 : c.dtype == BOLT_DEC128 ? 64 : dtype_size(c.dtype) * 8;
    const int desc = descending ? descending[k] : 0;
    const int nf = nulls_first ? nulls_first[k] : 1;
    // value (a Decimal128 as its unsigned low word, then its signed high
    // word), then the NULL rank: least significant first
    std::vector<int> modes{0};
    if (c.dtype == BOLT_DEC128) modes.push_back(2);
    if (c.validity) modes.push_back(1);
    for (int mode : modes) {
      order_key_kernel<<<grid_for(n, 256, 4), 256, 0, s>>>(c.data, c.validity, c.dtype, bits, desc, nf, mode, out_rows,
                                                          k0.ptr, n);
      BOLT_HIP(hipGetLastError());
      BOLT_TRY(radix_sort_u64_pairs(s, k0.ptr, out_rows, n, k1.ptr, p0.ptr, mode == 1 ? 1 : bits, 0, 0));
      BOLT_HIP(hipMemcpyAsync(out_rows, p0.ptr, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
    }
  }
  BOLT_HIP(hipStreamSynchronize(s));
  return BOLT_OK;
}

extern "C" int bolt_sort_limit(bolt_stream_t stream, const bolt_column_t* keys, int32_t n_keys,
                               const int32_t* descending, const int32_t* nulls_first, int64_t offset, int64_t limit,
                               uint32_t* out_rows, int64_t* n_out) {
  BOLT_CHECK_DEVICE();
  clear_error();
  hipStream_t s = (hipStream_t)stream;
  *n_out = 0;
  if (n_keys < 1 || keys == nullptr) return set_error(BOLT_ERR_PLAN, "bolt_sort_limit: no sort keys");
  if (offset < 0 || limit < 0) return set_error(BOLT_ERR_PLAN, "bolt_sort_limit: negative offset or limit");
  const int64_t n = keys[0].length;
  if (n >= (1ll << 31)) return set_error(BOLT_ERR_PLAN, "bolt_sort_limit: %lld rows exceeds 2^31", (long long)n);
  // execute_limit (src/exec/limit.rs:18): min(limit, n - offset) rows, none past the end
  const int64_t m = offset >= n ? 0 : (limit < n - offset ? limit : n - offset);
  if (m == 0) {
    for (int k = 0; k < n_keys; k++)
      if (keys[k].length != n) return set_error(BOLT_ERR_PLAN, "bolt_sort_limit: key columns differ in length");
    return BOLT_OK;
  }
  const int64_t K = offset + m;  // rows of the order that must be exact
  std::vector<int32_t> desc(n_keys + 1, 0), nf(n_keys + 1, 1);
  for (int k = 0; k < n_keys; k++) {
    desc[k] = descending ? descending[k] : 0;
    nf[k] = nulls_first ? nulls_first[k] : 1;
  }
  // The full sort, then the slice.
  auto full_sort = [&]() -> int {
    DevBuf<uint32_t> perm;
    BOLT_TRY(perm.alloc((size_t)n, s));
    BOLT_TRY(bolt_sort_indices(stream, keys, n_keys, desc.data(), nf.data(), perm.ptr));
    BOLT_HIP(hipMemcpyAsync(out_rows, perm.ptr + offset, (size_t)m * 4, hipMemcpyDeviceToDevice, s));
    BOLT_HIP(hipStreamSynchronize(s));
    *n_out = m;
    return BOLT_OK;
  };
  // inputs of up to 2^20 rows (where the sort is launch-bound and a select
  // saves nothing), or a limit that reaches a quarter of the rows
  if (n <= (1 << 20) || K * 4 >= n) return full_sort();
  for (int k = 0; k < n_keys; k++) {
    if (keys[k].length != n) return set_error(BOLT_ERR_PLAN, "bolt_sort_limit: key columns differ in length");
    const int dt = keys[k].dtype;
    if (!(dt == BOLT_I32 || dt == BOLT_I64 || dt == BOLT_F32 || dt == BOLT_F64 || dt == BOLT_DATE32 || dt == BOLT_BOOL ||
          dt == BOLT_DEC128))
      return set_error(BOLT_ERR_UNSUPPORTED, "bolt_sort_limit: sort key dtype %s not supported", dtype_name(dt));
  }
  const bolt_column_t& c0 = keys[0];
  TopKey tk;
  tk.d = c0.data;
  tk.v = c0.validity;
  tk.dt = c0.dtype;
  tk.bits = c0.dtype == BOLT_BOOL ? 1 : (c0.dtype == BOLT_DEC128 ? 64 : dtype_size(c0.dtype) * 8);
  tk.desc = desc[0];
  tk.nulls_first = nf[0];
  const int kw = c0.dtype == BOLT_BOOL ? 1 : dtype_size(c0.dtype);  // 16 for Decimal128
  // Radix select, most significant digit first: after each pass `prefix`
  // holds the top 64 - hi bits of T, `below` the rows with a smaller prefix
  // and `eq` those with the same one; stop once the tie bucket is small.
  const unsigned hgrid = grid_for(n, kSelBlock, kSelUnroll, 512);
  DevBuf<uint32_t> hist, part;
  BOLT_TRY(hist.alloc(kSelBins, s));
  BOLT_TRY(part.alloc((size_t)hgrid * kSelBins, s));
  std::vector<uint32_t> h(kSelBins);
  uint64_t prefix = 0;
  int hi = tk.bits;  // every image is below 2^bits (coarse_of)
  int64_t below = 0, need = K, eq = n;
  const unsigned grid = grid_for(n, 256, kSelUnroll, 4096);
  // Candidates guessed in the first pass: the rows whose first digit is at
  // most spec_bound, where the K-th row's digit lies when the keys are spread
  // (about K * 2^14 / n); capacity for twice the expected count.
  const uint64_t spec_bound = (uint64_t)((double)K * (double)kSelBins / (double)n) + 1;
  const int64_t spec_cap = (int64_t)std::min<double>(
      (double)n, 2.0 * (double)(spec_bound + 1) * (double)n / (double)kSelBins + 65536.0);
  DevBuf<uint32_t> spec;
  DevBuf<unsigned long long> spec_cnt;
  BOLT_TRY(spec.alloc((size_t)spec_cap, s));
  BOLT_TRY(spec_cnt.alloc(1, s));
  BOLT_TRY(fill_u32(spec_cnt.ptr, 0, 2, s));
  unsigned long long spec_got = 0;
  int first_digit = -1;
  while (hi > 0) {
    const int width = hi >= kSelBits ? kSelBits : hi;
    const int shift = hi - width;
    uint32_t* sp = first_digit < 0 ? spec.ptr : nullptr;
#define HIST_LAUNCH(WW)                                                                                         \
  topk_hist_kernel<WW><<<hgrid, kSelBlock, 0, s>>>(tk, n, shift, width, prefix, part.ptr, spec_bound, sp,       \
                                                   spec_cnt.ptr, spec_cap)
    switch (kw) {
      case 1: HIST_LAUNCH(1); break;
      case 4: HIST_LAUNCH(4); break;
      case 8: HIST_LAUNCH(8); break;
      default: HIST_LAUNCH(16); break;
    }
#undef HIST_LAUNCH
    BOLT_TRY(fill_u32(hist.ptr, 0, (size_t)1 << width, s));
    topk_hist_sum_kernel<<<dim3((unsigned)(((1 << width) + 255) / 256), 32), 256, 0, s>>>(part.ptr, (int)hgrid,
                                                                                        1 << width, hist.ptr);
    BOLT_HIP(hipGetLastError());
    BOLT_HIP(hipMemcpyAsync(h.data(), hist.ptr, (size_t)(1 << width) * 4, hipMemcpyDeviceToHost, s));
    if (first_digit < 0) BOLT_HIP(hipMemcpyAsync(&spec_got, spec_cnt.ptr, 8, hipMemcpyDeviceToHost, s));
    BOLT_HIP(hipStreamSynchronize(s));
    int64_t cum = 0;
    int d = 0;
    for (; d < (1 << width); d++) {
      if (cum + (int64_t)h[d] >= need) break;
      cum += h[d];
    }
    if (d == (1 << width)) return set_error(BOLT_ERR_OTHER, "bolt_sort_limit: radix select lost rows");
    if (first_digit < 0) first_digit = d;
    below += cum;
    need -= cum;
    eq = h[d];
    prefix = (prefix << width) | (uint64_t)d;
    hi = shift;
    if (eq <= (1 << 16)) break;
  }
  const int64_t Cx = below + eq;  // rows whose image >> hi <= prefix
  if (Cx * 2 > n) return full_sort();  // the leading key ties too often to narrow the rows
  DevBuf<uint32_t> cand;
  DevBuf<unsigned long long> cnt;
  BOLT_TRY(cnt.alloc(1, s));
  BOLT_TRY(fill_u32(cnt.ptr, 0, 2, s));
  // The candidates: the first pass's guess when it holds every row of the
  // K-th row's first digit and below (a superset of the Cx rows, which the
  // steps below accept), else a collect pass for exactly the Cx rows.
  const uint32_t* cptr;
  int64_t C;
  if (first_digit >= 0 && (uint64_t)first_digit <= spec_bound && (int64_t)spec_got <= spec_cap) {
    cptr = spec.ptr;
    C = (int64_t)spec_got;
  } else {
    BOLT_TRY(cand.alloc((size_t)Cx, s));
    switch (kw) {
      case 1: topk_collect_kernel<1><<<grid, 256, 0, s>>>(tk, n, hi, prefix, cand.ptr, cnt.ptr, Cx); break;
      case 4: topk_collect_kernel<4><<<grid, 256, 0, s>>>(tk, n, hi, prefix, cand.ptr, cnt.ptr, Cx); break;
      case 8: topk_collect_kernel<8><<<grid, 256, 0, s>>>(tk, n, hi, prefix, cand.ptr, cnt.ptr, Cx); break;
      default: topk_collect_kernel<16><<<grid, 256, 0, s>>>(tk, n, hi, prefix, cand.ptr, cnt.ptr, Cx); break;
    }
    BOLT_HIP(hipGetLastError());
    unsigned long long got = 0;
    BOLT_HIP(hipMemcpyAsync(&got, cnt.ptr, 8, hipMemcpyDeviceToHost, s));
    BOLT_HIP(hipStreamSynchronize(s));
    if ((int64_t)got != Cx)
      return set_error(BOLT_ERR_OTHER, "bolt_sort_limit: collected %llu candidates, expected %lld", got, (long long)Cx);
    cptr = cand.ptr;
    C = Cx;
  }
  // One key without NULLs (its image is then exact): narrow the candidates
  // further and sort the few that remain in one workgroup.
  if (n_keys == 1 && !c0.validity && kw != 16 && c0.dtype != BOLT_BOOL && K <= kSmallSort / 2) {
    // continue the select inside the bucket of the first one: `below` rows
    // precede it, the K-th row is the need-th of its `eq`
    uint64_t p2 = prefix;
    int hi2 = hi;
    int64_t below2 = below, need2 = need, eq2 = eq;
    const unsigned cgrid = (unsigned)((C + 255) / 256);
    while (hi2 > 0 && below2 + eq2 > kSmallSort) {
      const int width = hi2 >= kSelBits ? kSelBits : hi2;
      const int shift = hi2 - width;
      BOLT_TRY(fill_u32(hist.ptr, 0, (size_t)1 << width, s));
      switch (kw) {
        case 1: topk_cand_hist_kernel<1><<<cgrid, 256, 0, s>>>(tk, cptr, C, shift, width, p2, hist.ptr); break;
        case 4: topk_cand_hist_kernel<4><<<cgrid, 256, 0, s>>>(tk, cptr, C, shift, width, p2, hist.ptr); break;
        default: topk_cand_hist_kernel<8><<<cgrid, 256, 0, s>>>(tk, cptr, C, shift, width, p2, hist.ptr); break;
      }
      BOLT_HIP(hipGetLastError());
      BOLT_HIP(hipMemcpyAsync(h.data(), hist.ptr, (size_t)(1 << width) * 4, hipMemcpyDeviceToHost, s));
      BOLT_HIP(hipStreamSynchronize(s));
      int64_t cum = 0;
      int d = 0;
      for (; d < (1 << width); d++) {
        if (cum + (int64_t)h[d] >= need2) break;
        cum += h[d];
      }
      if (d == (1 << width)) return set_error(BOLT_ERR_OTHER, "bolt_sort_limit: candidate select lost rows");
      below2 += cum;
      need2 -= cum;
      eq2 = h[d];
      p2 = (p2 << width) | (uint64_t)d;
      hi2 = shift;
    }
    int64_t C2 = below2 + eq2;
    // More rows tie on the K-th row's whole key than one workgroup sorts:
    // the order among them is by row id, so select the need2-th smallest
    // row id of the tie (14, 14 and 4 bits) and keep the tied rows up to it.
    uint32_t rmax = 0xFFFFFFFFu;
    if (C2 > kSmallSort && hi2 == 0 && below2 + need2 <= kSmallSort) {
      uint32_t rpre = 0;
      int rhi = 32;
      int64_t need3 = need2;
      while (rhi > 0) {
        const int width = rhi >= kSelBits ? kSelBits : rhi;
        const int shift = rhi - width;
        BOLT_TRY(fill_u32(hist.ptr, 0, (size_t)1 << width, s));
        switch (kw) {
          case 1: topk_tie_hist_kernel<1><<<cgrid, 256, 0, s>>>(tk, cptr, C, p2, shift, width, rpre, hist.ptr); break;
          case 4: topk_tie_hist_kernel<4><<<cgrid, 256, 0, s>>>(tk, cptr, C, p2, shift, width, rpre, hist.ptr); break;
          default: topk_tie_hist_kernel<8><<<cgrid, 256, 0, s>>>(tk, cptr, C, p2, shift, width, rpre, hist.ptr); break;
        }
        BOLT_HIP(hipGetLastError());
        BOLT_HIP(hipMemcpyAsync(h.data(), hist.ptr, (size_t)(1 << width) * 4, hipMemcpyDeviceToHost, s));
        BOLT_HIP(hipStreamSynchronize(s));
        int64_t cum = 0;
        int d = 0;
        for (; d < (1 << width); d++) {
          if (cum + (int64_t)h[d] >= need3) break;
          cum += h[d];
        }
        if (d == (1 << width)) return set_error(BOLT_ERR_OTHER, "bolt_sort_limit: tie select lost rows");
        need3 -= cum;
        rpre = (rpre << width) | (uint32_t)d;
        rhi = shift;
      }
      rmax = rpre;  // row ids are distinct: exactly need2 tied rows are <= rmax
      C2 = below2 + need2;
    }
    if (C2 <= kSmallSort) {
      DevBuf<uint32_t> cand2;
      BOLT_TRY(cand2.alloc((size_t)C2, s));
      BOLT_TRY(fill_u32(cnt.ptr, 0, 2, s));
      switch (kw) {
        case 1: topk_cand_collect_kernel<1><<<cgrid, 256, 0, s>>>(tk, cptr, C, hi2, p2, cand2.ptr, cnt.ptr, C2, rmax); break;
        case 4: topk_cand_collect_kernel<4><<<cgrid, 256, 0, s>>>(tk, cptr, C, hi2, p2, cand2.ptr, cnt.ptr, C2, rmax); break;
        default: topk_cand_collect_kernel<8><<<cgrid, 256, 0, s>>>(tk, cptr, C, hi2, p2, cand2.ptr, cnt.ptr, C2, rmax); break;
      }
      BOLT_HIP(hipGetLastError());
      int npow = 1;
      while (npow < C2) npow <<= 1;
      switch (kw) {
        case 1: topk_small_sort_kernel<1><<<1, 1024, 0, s>>>(tk, cand2.ptr, (int)C2, npow, (int)offset, (int)m, out_rows); break;
        case 4: topk_small_sort_kernel<4><<<1, 1024, 0, s>>>(tk, cand2.ptr, (int)C2, npow, (int)offset, (int)m, out_rows); break;
        default: topk_small_sort_kernel<8><<<1, 1024, 0, s>>>(tk, cand2.ptr, (int)C2, npow, (int)offset, (int)m, out_rows); break;
      }
      BOLT_HIP(hipGetLastError());
      unsigned long long got2 = 0;
      BOLT_HIP(hipMemcpyAsync(&got2, cnt.ptr, 8, hipMemcpyDeviceToHost, s));
      BOLT_HIP(hipStreamSynchronize(s));
      if ((int64_t)got2 != C2)
        return set_error(BOLT_ERR_OTHER, "bolt_sort_limit: narrowed to %llu candidates, expected %lld", got2,
                         (long long)C2);
      *n_out = m;
      return BOLT_OK;
    }
  }
  // Sort the candidates by every key, then by row id (the full sort is stable).
  std::vector<DevBuf<uint8_t>> vals(n_keys), valid(n_keys);
  std::vector<bolt_column_t> ck(n_keys + 1);
  for (int k = 0; k < n_keys; k++) {
    const int w = dtype_size(keys[k].dtype);
    BOLT_TRY(vals[k].alloc((size_t)C * w, s));
    memset(&ck[k], 0, sizeof ck[k]);
    ck[k].data = vals[k].ptr;
    if (keys[k].validity) {
      BOLT_TRY(valid[k].alloc((size_t)((C + 7) / 8 + 8), s));
      ck[k].validity = valid[k].ptr;
    }
    BOLT_TRY(bolt_take(stream, &keys[k], cptr, C, &ck[k]));
  }
  memset(&ck[n_keys], 0, sizeof ck[n_keys]);
  ck[n_keys].data = const_cast<uint32_t*>(cptr);  // row ids < 2^31: ascending as Int32
  ck[n_keys].length = C;
  ck[n_keys].dtype = BOLT_I32;
  DevBuf<uint32_t> perm;
  BOLT_TRY(perm.alloc((size_t)C, s));
  BOLT_TRY(bolt_sort_indices(stream, ck.data(), n_keys + 1, desc.data(), nf.data(), perm.ptr));
  topk_emit_kernel<<<(unsigned)((m + 255) / 256), 256, 0, s>>>(cptr, perm.ptr, offset, m, out_rows);
  BOLT_HIP(hipGetLastError());
  BOLT_HIP(hipStreamSynchronize(s));
  *n_out = m;
  return BOLT_OK;
}
