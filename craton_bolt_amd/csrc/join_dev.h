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
  hipStream_t stream = nullptr;
};

namespace bolt {

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
