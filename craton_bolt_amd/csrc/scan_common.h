// scan_common.h — single-pass decoupled look-back over tiles.
//
// Each tile publishes its aggregate, then one wave inspects 64 predecessors
// at a time until it meets an inclusive prefix (Merrill & Garland,
// "Single-pass Parallel Prefix Scan with Decoupled Look-back", 2016).  Status
// words are 64-bit: 2 flag bits + a 62-bit count, so the reference's 2^30-row
// limit of its 32-bit look-back words (src/jit/prefix_scan.rs:117,1164) does
// not apply.  Tile ids come from an atomic ticket, so every predecessor of a
// tile is already running and the wait always ends; the spin is still
// bounded, and a timeout raises an error flag instead of hanging.
#pragma once

#include "bolt_internal.h"

namespace bolt {

constexpr uint64_t kFlagAgg = 1ull << 62;
constexpr uint64_t kFlagInc = 2ull << 62;
constexpr uint64_t kValueMask = (1ull << 62) - 1;
constexpr uint32_t kSpinLimit = 1u << 24;

// Status words carry their whole payload (flag + count in one 64-bit word),
// and no other memory is read across tiles, so relaxed device-scope atomics
// suffice.  Acquire / release at agent scope would invalidate or write back
// the per-XCD L2 on every poll and publish.
__device__ __forceinline__ uint64_t ld_status(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_status(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Called by one whole wave (64 lanes) of tile `tile` whose own count is
// `total`.  Publishes the tile's aggregate and inclusive prefix and returns
// the exclusive prefix (in every lane).  `err` is set on a spin timeout.
__device__ __forceinline__ uint64_t lookback_wave(uint64_t* __restrict__ status, int64_t tile, uint64_t total,
                                                  uint32_t* __restrict__ err) {
  int lane = threadIdx.x & 63;
  if (tile == 0) {
    if (lane == 0) st_status(&status[0], kFlagInc | total);
    return 0;
  }
  if (lane == 0) st_status(&status[tile], kFlagAgg | total);
  uint64_t excl = 0;
  int64_t base = tile - 1;
  for (;;) {
    int64_t idx = base - lane;
    uint64_t v = idx >= 0 ? ld_status(&status[idx]) : (kFlagInc | 0);
    uint32_t spins = 0;
    while (__ballot((v >> 62) == 0) != 0) {
      if ((v >> 62) == 0) v = ld_status(&status[idx]);
      if (++spins > kSpinLimit) {
        if (lane == 0) atomicOr(err, 1u);
        v = kFlagInc;  // give up: the host reports the error
      }
      __builtin_amdgcn_s_sleep(1);
    }
    uint64_t inc = __ballot((v >> 62) == 2);
    uint64_t val = v & kValueMask;
    int stop = inc ? __builtin_ctzll(inc) : 64;
    uint64_t part = lane <= stop ? val : 0;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) part += __shfl_xor(part, m);
    excl += part;
    if (inc) break;
    base -= 64;
  }
  if (lane == 0) st_status(&status[tile], kFlagInc | (excl + total));
  return excl;
}

}  // namespace bolt
