// join_dev.h — the hash-join table shared by the join operators (join.hip)
// and the served form (serve.hip): 16-byte slots
//   { i64 key (x = low word, y = high word), u32 value, u32 state }
// with state 0 = empty, linear probing from fmix64(key).
#pragma once

#include "bolt_internal.h"

struct bolt_join_table {
  uint4* tbl = nullptr;
  uint64_t mask = 0;  // capacity - 1 (power of two)
  int64_t n_build = 0;
  int32_t key_dtype = 0;
  int32_t carry_dtype = 0;  // 0 = values are build row ids
  uint32_t* flags = nullptr; // [0] nonunique, [1] lost rows
  hipStream_t stream = nullptr;  // the build's stream (its kernels may still run when build returns)
  hipEvent_t built = nullptr;    // recorded after the build's last kernel; probes on other streams wait on it
  uint4* bloom = nullptr;     // split-block Bloom filter of the build keys (nullptr = none)
  uint64_t bloom_mask = 0;    // blocks - 1 (power of two)
};

namespace bolt {

// Optional split-block Bloom filter over the build keys (BOLT_JOIN_BLOOM=1),
// 16 bits per key: one 128-bit
// block per key, chosen by bits 40..63 of fmix64(key), with one bit set in
// each of its four 32-bit words (bit positions from four 5-bit fields of bits
// 20..39); the table slot uses the low bits of the same hash.  About 0.25% of
// absent keys pass.  A probe reads one 16-byte block, which stays in the
// Infinity Cache for build sides of tens of millions of rows, and only
// possible hits go on to the table in HBM.  (bloom_pattern / bloom_hit live in
// device_acc.h, shared with the generated probe kernels.)
__device__ __forceinline__ bool bloom_maybe(const uint4* bloom, uint64_t bmask, int64_t k) {
  if (!bloom) return true;
  const uint64_t h = fmix64((uint64_t)k);
  return bloom_hit(bloom[(h >> 40) & bmask], h);
}

__device__ __forceinline__ int64_t slot_key(const uint4& e) { return (int64_t)(((uint64_t)e.y << 32) | e.x); }

// Walk the probe sequence of `k`: on_match(value) for every stored copy;
// stops at the first empty slot, or after the first match when `first_only`.
template <typename F>
__device__ __forceinline__ void probe_chain(const uint4* __restrict__ tbl, uint64_t mask, int64_t k, bool first_only,
                                            F&& on_match) {
  uint64_t s = fmix64((uint64_t)k) & mask;
  for (uint64_t step = 0; step <= mask; step++) {
    const uint4 e = tbl[s];
    if (e.w == 0u) return;
    if (slot_key(e) == k) {
      on_match(e.z);
      if (first_only) return;
    }
    s = (s + 1) & mask;
  }
}

}  // namespace bolt
