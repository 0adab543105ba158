// reduce_common.h — block-level reduction of aggregate accumulators and the
// deterministic finishing kernel shared by the scalar-aggregate and the
// join-aggregate operators.
#pragma once

#include "ops_common.h"

namespace bolt {

constexpr int kReduceBlock = 256;
constexpr int kWaves = kReduceBlock / 64;

// Reduce NA accumulators across a 256-thread block; thread 0 writes them to
// out[0..NA).  `ops` is read from the uniform AggList.
template <int NA>
__device__ __forceinline__ void block_reduce_store(Acc (&acc)[NA], const AggList& aggs, Acc* out) {
  __shared__ Acc sh[kWaves][NA];
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int a = 0; a < NA; a++) wave_reduce_acc(acc[a], aggs.a[a].op, aggs.a[a].dtype);
  if (lane == 0) {
#pragma unroll
    for (int a = 0; a < NA; a++) sh[wave][a] = acc[a];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int a = 0; a < NA; a++) {
      Acc r = sh[0][a];
      for (int w = 1; w < kWaves; w++) acc_merge(r, sh[w][a], aggs.a[a].op, aggs.a[a].dtype);
      out[a] = r;
    }
  }
}

// One block folds `n_parts` per-block partials (layout [part][NA]) in a fixed
// order, so float results do not depend on scheduling.
// (layout [part][stride], this chunk's aggregates at offset `first`).
template <int NA>
__global__ __launch_bounds__(kReduceBlock) void finish_partials_kernel(const Acc* __restrict__ parts, int n_parts,
                                                                      int stride, int first, AggList aggs,
                                                                      Acc* __restrict__ out) {
  Acc acc[NA];
#pragma unroll
  for (int a = 0; a < NA; a++) acc_init(acc[a], aggs.a[a].op);
  for (int p = threadIdx.x; p < n_parts; p += kReduceBlock) {
#pragma unroll
    for (int a = 0; a < NA; a++)
      acc_merge(acc[a], parts[(int64_t)p * stride + first + a], aggs.a[a].op, aggs.a[a].dtype);
  }
  block_reduce_store<NA>(acc, aggs, out);
}

// Launch the finishing kernel for a chunk of `na` (<= 4) aggregates whose
// partials sit at offset `first` of rows of `stride` accumulators.
int launch_finish(int na, const Acc* parts, int n_parts, int stride, int first, const AggList& chunk, Acc* out,
                  hipStream_t s);

// Split an AggList into chunks of at most 4 (kernels are instantiated for
// 1..4 aggregates so that accumulators stay in registers).
inline AggList agg_chunk(const AggList& all, int first, int count) {
  AggList c;
  c.n = count;
  for (int i = 0; i < count; i++) c.a[i] = all.a[first + i];
  return c;
}

}  // namespace bolt
