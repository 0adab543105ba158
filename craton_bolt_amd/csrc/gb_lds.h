// gb_lds.h — the LDS tier of hash GROUP BY: slot lookup in a workgroup's
// LDS table, spilling of rows that find no slot, and the flush of the
// table as partial records.
//
// Self-contained on top of device_acc.h (no includes, no standard-library
// types): the same text is compiled into the library by hipcc and prepended
// to the GROUP BY kernels generated per query and compiled with hipRTC, so
// both produce the same partial records, spill segments and sketch.
#pragma once

namespace bolt {

constexpr u64_t kEmpty = ~0ull;
constexpr int kGbBlock = 1024;
constexpr int kHllBits = 11;
constexpr int kHllRegs = 1 << kHllBits;
constexpr int kLdsProbes = 16;

// Outputs of the LDS tier.
struct PartOut {
  u64_t* keys;          // [cap]
  unsigned char* tag;   // [cap] 0 regular, 1 NULL group, 2 empty-marker key
  u64_t* words;         // [words][cap]
  i64_t cap;
  u64_t* count;
  u32_t* spill;         // [grid][per] row ids: each workgroup spills into its own segment
  u64_t* spill_count;   // total spilled rows (one atomic per workgroup)
  u32_t* spill_seg;     // [grid] rows in each workgroup's segment
  i64_t per;            // segment capacity
  u32_t* hll;           // [kHllRegs]
  // Optional, for large inputs: the spilled rows' keys beside their row ids,
  // so the partitioned tier reads them in order instead of gathering them by
  // row id (the key is in a register when the row spills).
  u64_t* spill_key;     // [grid][per] or null
};

// Float keys group -0.0 with +0.0 and every NaN together
// (src/exec/groupby_common.rs:302).
__device__ __forceinline__ u32_t canon_f32_bits(u32_t b) {
  if ((b & 0x7FFFFFFFu) > 0x7F800000u) return 0x7FC00000u;
  if ((b & 0x7FFFFFFFu) == 0) return 0;
  return b;
}
__device__ __forceinline__ u64_t canon_f64_bits(u64_t b) {
  if ((b & 0x7FFFFFFFFFFFFFFFull) > 0x7FF0000000000000ull) return 0x7FF8000000000000ull;
  if ((b & 0x7FFFFFFFFFFFFFFFull) == 0) return 0;
  return b;
}

// HyperLogLog register update from an independent hash of the key.
__device__ __forceinline__ void hll_add(u32_t* regs, u64_t key) {
  const u64_t h = fmix64(key ^ 0x9E3779B97F4A7C15ull);
  const u32_t idx = (u32_t)(h >> (64 - kHllBits));
  const u64_t rest = (h << kHllBits) | (1ull << (kHllBits - 1));
  const u32_t rho = (u32_t)__clzll(rest) + 1;
  atomicMax(&regs[idx], rho);
}

// Carry-propagating 128-bit add: the carry out of the low word is computed
// from the value the atomic returned, so concurrent adds compose exactly.
__device__ __forceinline__ void atomic_add128(u64_t* lo, u64_t* hi, u64_t vlo, u64_t vhi) {
  const u64_t old = atomicAdd(lo, vlo);
  const u64_t carry = (old + vlo) < old;
  const u64_t add_hi = vhi + carry;
  if (add_hi) atomicAdd(hi, add_hi);
}

// 192-bit add of a sign-extended 128-bit value (Decimal128 SUM): detects
// overflow of the i128 result exactly at finalize.
__device__ __forceinline__ void atomic_add192(u64_t* w, u64_t stride, u64_t vlo, u64_t vhi, u64_t vtop) {
  const u64_t old = atomicAdd(&w[0], vlo);
  const u64_t c1 = (old + vlo) < old;
  const u64_t mid_add = vhi + c1;
  const u64_t ca = (c1 && vhi == ~0ull) ? 1 : 0;  // carry out of vhi + c1
  u64_t c2 = 0;
  if (mid_add) {
    const u64_t old2 = atomicAdd(&w[stride], mid_add);
    c2 = (old2 + mid_add) < old2;
  }
  const u64_t top_add = vtop + ca + c2;
  if (top_add) atomicAdd(&w[2 * stride], top_add);
}

// Wave-aggregated append: returns this lane's position for lanes with `want`.
__device__ __forceinline__ u64_t wave_append(bool want, u64_t* counter) {
  const u64_t m = __ballot(want);
  if (m == 0) return 0;
  const int lane = threadIdx.x & 63;
  const int leader = __builtin_ctzll(m);
  u64_t base = 0;
  if (lane == leader) base = atomicAdd(counter, (u64_t)__popcll(m));
  base = __shfl(base, leader);
  const u64_t below = lane == 0 ? 0ull : (m & (~0ull >> (64 - lane)));
  return base + (u64_t)__popcll(below);
}

// Home slot of `key` in an S-slot LDS table: the top 32 bits of the key
// times a 64-bit odd constant (every key bit reaches them), scaled to S.
__device__ __forceinline__ u32_t gb_lds_home(u64_t key, int S) {
  const u32_t h = (u32_t)((key * 0x9E3779B97F4A7C15ull) >> 32);
  return (u32_t)(((u64_t)h * (u32_t)S) >> 32);
}

// Linear probe for `key` (never kEmpty) from slot `s`, claiming an empty slot
// on the way (counted in *occupied); -1 if no slot turns up within `probes`
// slots.
__device__ __forceinline__ int gb_lds_probe(u64_t* skeys, int S, u64_t key, u32_t s, int probes, u32_t* occupied) {
  for (int p = 0; p < probes; p++) {
    const u64_t cur = skeys[s];
    if (cur == key) return (int)s;
    if (cur == kEmpty) {
      const u64_t prev = atomicCAS(&skeys[s], kEmpty, key);
      if (prev == kEmpty) {
        atomicAdd(occupied, 1u);
        return (int)s;
      }
      if (prev == key) return (int)s;
    }
    s = (s + 1 == (u32_t)S) ? 0 : s + 1;
  }
  return -1;
}

// Probe bound: kLdsProbes while the table has room; 2 once 15/16 of the
// slots are taken, where a key not found near its home almost never finds a
// slot further on, and a long walk before every spill is what costs.
__device__ __forceinline__ int gb_lds_probes(int S, const u32_t* occupied) {
  return *occupied >= (u32_t)(S - (S >> 4)) ? 2 : kLdsProbes;
}

// Slot of `key` near its home, or -1.
__device__ __forceinline__ int gb_lds_slot(u64_t* skeys, int S, u64_t key, u32_t* occupied) {
  return gb_lds_probe(skeys, S, key, gb_lds_home(key, S), gb_lds_probes(S, occupied), occupied);
}

// Append `row` to this workgroup's spill segment (called by the whole wave;
// one LDS atomic per wave on the workgroup's own counter, no global one).
__device__ __forceinline__ void gb_spill(bool spill, i64_t row, u64_t key, u32_t* sused, u32_t* shll,
                                         const PartOut& out) {
  const u64_t m = __ballot(spill);
  if (m) {
    const int lane = threadIdx.x & 63;
    const int leader = __builtin_ctzll(m);
    u32_t base = 0;
    if (lane == leader) base = atomicAdd(&sused[2], (u32_t)__popcll(m));
    base = __shfl(base, leader);
    if (spill) {
      const u64_t below = lane == 0 ? 0ull : (m & (~0ull >> (64 - lane)));
      const i64_t at = (i64_t)blockIdx.x * out.per + base + (u32_t)__popcll(below);
      out.spill[at] = (u32_t)row;
      if (out.spill_key) out.spill_key[at] = key;
      hll_add(shll, key);
    }
  }
}

// After the row loop: record the spill segment, append the table's groups as
// partial records (slot S is the NULL group, S + 1 the key equal to the empty
// marker) and merge the workgroup's sketch.  `skeys` / `swords` hold S + 2
// slots; word w of slot i is swords[w * (S + 2) + i].
__device__ __forceinline__ void gb_lds_flush(const u64_t* skeys, const u64_t* swords, int S, int W, u32_t* shll,
                                             const u32_t* sused, const PartOut& out) {
  const int SS = S + 2;
  __syncthreads();
  if (threadIdx.x == 0) {
    out.spill_seg[blockIdx.x] = sused[2];
    if (sused[2]) atomicAdd(out.spill_count, (u64_t)sused[2]);
  }
  for (int base = 0; base < SS; base += (int)blockDim.x) {
    const int i = base + (int)threadIdx.x;
    bool used = false;
    unsigned char tag = 0;
    if (i < S) used = skeys[i] != kEmpty;
    else if (i < SS) { used = sused[i - S] != 0; tag = (unsigned char)(i - S + 1); }
    const u64_t pos = wave_append(used, out.count);
    if (used) {
      out.keys[pos] = tag == 2 ? kEmpty : (tag == 1 ? 0ull : skeys[i]);
      out.tag[pos] = tag;
      for (int w = 0; w < W; w++) out.words[(u64_t)w * (u64_t)out.cap + pos] = swords[(u64_t)w * SS + i];
      if (tag == 0) hll_add(shll, skeys[i]);
    }
  }
  __syncthreads();
  for (int r = threadIdx.x; r < kHllRegs; r += (int)blockDim.x)
    if (shll[r]) atomicMax(&out.hll[r], shll[r]);
}

}  // namespace bolt
