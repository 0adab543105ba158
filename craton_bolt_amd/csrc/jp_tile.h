// jp_tile.h — one tile step of the partitioned join's scatter: rows that
// already carry their partition (or -1 when the filter dropped them) are
// ranked within their partition with LDS atomics, staged in LDS in partition
// order and written to the partitioned key and carried-column arrays at each
// partition's running cursor, so a tile's rows of one partition leave as one
// contiguous run.
//
// Self-contained on top of device_acc.h (no includes, no standard-library
// types): the same text is compiled into the library by hipcc and prepended
// to the scatter kernels generated per query and compiled with hipRTC, so
// both write the same runs.
#pragma once

namespace bolt {

constexpr int kJpBlock = 1024;    // scatter workgroup (one per CU)
constexpr int kJpMaxParts = 1024;
constexpr int kJpMaxPay = 2;      // probe columns carried with the key

// Destination arrays of the carried columns and their widths (4 or 8 bytes).
struct JpDst {
  void* d[kJpMaxPay];
  int w[kJpMaxPay];
};

// LDS of a scatter workgroup whose tiles hold T rows carrying NPAY columns.
template <int T, int NPAY>
struct JpTileLds {
  i64_t key[T];
  u64_t pay[NPAY > 0 ? NPAY : 1][T];
  u32_t cur[kJpMaxParts];  // next output row of each partition in this chunk's run
  u32_t cnt[kJpMaxParts];  // rows of the current tile per partition
  u32_t off[kJpMaxParts];  // exclusive prefix of cnt
  unsigned short part[T];
  u32_t wsum[kJpBlock / 64];
};

// Block-wide exclusive scan of P (<= kJpMaxParts) counters into `off`.
__device__ __forceinline__ void jp_block_scan(const u32_t* cnt, u32_t* off, int P, u32_t* wsum) {
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const u32_t v = t < P ? cnt[t] : 0u;
  u32_t incl = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const u32_t o = __shfl_up(incl, d);
    if (lane >= d) incl += o;
  }
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  if (t < 64) {
    const u32_t w = t < kJpBlock / 64 ? wsum[t] : 0u;
    u32_t wi = w;
#pragma unroll
    for (int d = 1; d < 16; d <<= 1) {
      const u32_t o = __shfl_up(wi, d);
      if (lane >= d) wi += o;
    }
    if (t < kJpBlock / 64) wsum[t] = wi - w;
  }
  __syncthreads();
  if (t < P) off[t] = wsum[wave] + incl - v;
}

// Chunk g's cursors start at its run of each partition.
template <int T, int NPAY>
__device__ __forceinline__ void jp_tile_begin(JpTileLds<T, NPAY>& L, int P, int g, const u32_t* counts,
                                              const u32_t* base) {
  for (int p = threadIdx.x; p < P; p += kJpBlock) {
    L.cur[p] = base[p] + counts[(i64_t)g * P + p];
    L.cnt[p] = 0;
  }
  __syncthreads();
}

// Rank and stage a tile: lane rows u with part[u] >= 0 are placed in LDS in
// partition order.  Ends with a barrier; the caller's registers are free.
template <int T, int NPAY, int ITEMS>
__device__ __forceinline__ void jp_tile_stage(JpTileLds<T, NPAY>& L, int P, const i64_t (&k)[ITEMS],
                                              const int (&part)[ITEMS],
                                              const u64_t (&pv)[ITEMS][NPAY > 0 ? NPAY : 1]) {
  u32_t rank[ITEMS];
#pragma unroll
  for (int u = 0; u < ITEMS; u++) rank[u] = part[u] >= 0 ? atomicAdd(&L.cnt[part[u]], 1u) : 0u;
  __syncthreads();
  jp_block_scan(L.cnt, L.off, P, L.wsum);
  __syncthreads();
#pragma unroll
  for (int u = 0; u < ITEMS; u++) {
    if (part[u] < 0) continue;
    const u32_t j = L.off[part[u]] + rank[u];
    L.key[j] = k[u];
#pragma unroll
    for (int c = 0; c < NPAY; c++) L.pay[c][j] = pv[u][c];
    L.part[j] = (unsigned short)part[u];
  }
  __syncthreads();
}

// Write the staged tile at the partitions' cursors and advance them.
template <int T, int NPAY>
__device__ __forceinline__ void jp_tile_flush(JpTileLds<T, NPAY>& L, int P, i64_t* __restrict__ okeys,
                                              const JpDst& dst) {
  const u32_t m = L.off[P - 1] + L.cnt[P - 1];
  for (u32_t j = threadIdx.x; j < m; j += kJpBlock) {
    const int p = L.part[j];
    const u64_t pos = (u64_t)L.cur[p] + (j - L.off[p]);
    okeys[pos] = L.key[j];
#pragma unroll
    for (int c = 0; c < NPAY; c++) {
      if (dst.w[c] == 8) static_cast<u64_t*>(dst.d[c])[pos] = L.pay[c][j];
      else static_cast<u32_t*>(dst.d[c])[pos] = (u32_t)L.pay[c][j];
    }
  }
  __syncthreads();
  for (int p = threadIdx.x; p < P; p += kJpBlock) {
    L.cur[p] += L.cnt[p];
    L.cnt[p] = 0;
  }
  __syncthreads();
}

// actual[g][p] = rows chunk g wrote to partition p.
template <int T, int NPAY>
__device__ __forceinline__ void jp_tile_end(JpTileLds<T, NPAY>& L, int P, int g, const u32_t* counts,
                                            const u32_t* base, u32_t* actual) {
  for (int p = threadIdx.x; p < P; p += kJpBlock)
    actual[(i64_t)g * P + p] = L.cur[p] - (base[p] + counts[(i64_t)g * P + p]);
}

}  // namespace bolt
