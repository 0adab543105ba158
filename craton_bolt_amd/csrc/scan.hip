// scan.hip — filter + stream compaction and mask prefix scan.
//
// Replaces launch_predicate_kernel (src/exec/compact.rs:183), the
// predicate / prefix-scan / gather kernels (src/jit/scan_kernel.rs:265,
// src/jit/prefix_scan.rs:175,430,834,1227) and compact_columns_on_gpu
// (src/exec/gpu_compact.rs:1005).  The reference writes a u8 mask, scans it
// per block, scans the block sums on the host (or multipass above 16.8M
// rows), then gathers each column in another launch.  Here one kernel per
// 4096-row tile evaluates the predicate, ranks survivors with wave64 ballots,
// obtains the tile's output offset by decoupled look-back and writes the
// survivors of every column: each input byte is read once and each output
// byte written once.
#include <algorithm>

#include "ops_common.h"
#include "scan_common.h"

namespace bolt {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;  // rows per thread per tile
constexpr int kTileRows = kScanBlock * kScanItems;
constexpr int kMaxCompactCols = 16;

struct GatherCol {
  const uint8_t* src;
  uint8_t* dst;
  const uint8_t* src_valid;
  uint32_t* dst_valid;  // zeroed bitmap words
  int32_t width;
  int32_t pad;
};
struct GatherList {
  GatherCol c[kMaxCompactCols];
  int32_t n;
  int64_t cap;  // rows the outputs hold: survivors beyond it are counted, not written
};

__device__ __forceinline__ void copy_elem(const GatherCol& g, int64_t from, int64_t to) {
  switch (g.width) {
    case 1: g.dst[to] = g.src[from]; break;
    case 4: reinterpret_cast<uint32_t*>(g.dst)[to] = reinterpret_cast<const uint32_t*>(g.src)[from]; break;
    case 8: reinterpret_cast<uint64_t*>(g.dst)[to] = reinterpret_cast<const uint64_t*>(g.src)[from]; break;
    case 16: reinterpret_cast<uint4*>(g.dst)[to] = reinterpret_cast<const uint4*>(g.src)[from]; break;
  }
  if (g.dst_valid && bit_valid(g.src_valid, from)) atomicOr(&g.dst_valid[to >> 5], 1u << (to & 31));
}

// MODE 0: compact the survivors of `cols`; MODE 1: write the exclusive
// prefix of the keep flags for every row into `offsets`.
template <int MODE>
__global__ __launch_bounds__(kScanBlock) void compact_kernel(DevFilter f, const uint8_t* __restrict__ mask,
                                                             int64_t n, GatherList cols,
                                                             uint32_t* __restrict__ offsets,
                                                             uint64_t* __restrict__ status,
                                                             uint32_t* __restrict__ ticket,
                                                             unsigned long long* __restrict__ total_out,
                                                             uint32_t* __restrict__ err) {
  __shared__ uint32_t sh_tile;
  __shared__ uint32_t sh_cnt[kScanItems * 4];
  __shared__ uint64_t sh_prefix;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) sh_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  const int64_t tile = sh_tile;
  const int64_t base = tile * kTileRows;

  uint32_t keep = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; j++) {
    int64_t row = base + (int64_t)j * kScanBlock + threadIdx.x;
    bool k = row < n && (mask == nullptr || mask[row] != 0) && eval_filter(f, row);
    keep |= (uint32_t)k << j;
    uint64_t b = __ballot(k);
    if (lane == 0) sh_cnt[j * 4 + wave] = (uint32_t)__popcll(b);
  }
  __syncthreads();
  // exclusive scan of the 64 (item, wave) counts in row order, by wave 0
  if (wave == 0) {
    uint32_t c = sh_cnt[lane];
    uint32_t incl = c;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      uint32_t t = __shfl_up(incl, d);
      if (lane >= d) incl += t;
    }
    uint32_t tile_total = __shfl(incl, 63);
    sh_cnt[lane] = incl - c;
    uint64_t excl = lookback_wave(status, tile, tile_total, err);
    if (lane == 0) {
      sh_prefix = excl;
      if (tile == (int64_t)gridDim.x - 1) *total_out = excl + tile_total;
    }
  }
  __syncthreads();
  const uint64_t prefix = sh_prefix;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
  for (int j = 0; j < kScanItems; j++) {
    int64_t row = base + (int64_t)j * kScanBlock + threadIdx.x;
    bool k = (keep >> j) & 1u;
    uint64_t b = __ballot(k);
    uint64_t pos = prefix + sh_cnt[j * 4 + wave] + (uint64_t)__popcll(b & below);
    if (MODE == 1) {
      if (row < n) offsets[row] = (uint32_t)pos;
    } else if (k && (int64_t)pos < cols.cap) {
      for (int c = 0; c < cols.n; c++) copy_elem(cols.c[c], row, (int64_t)pos);
    }
  }
}

struct ScanScratch {
  DevBuf<uint64_t> status;
  DevBuf<uint32_t> misc;  // [0] ticket, [1] error, [2..3] total (u64)
};

static int scan_prepare(ScanScratch& sc, int64_t n_tiles, hipStream_t s) {
  BOLT_TRY(sc.status.alloc((size_t)n_tiles, s));
  BOLT_TRY(sc.misc.alloc(4, s));
  BOLT_TRY(fill_u32(sc.status.ptr, 0, (size_t)n_tiles * 2, s));
  BOLT_TRY(fill_u32(sc.misc.ptr, 0, 4, s));
  return BOLT_OK;
}

// Gather: dst[i] = src[rows[i]].  The validity bitmap is rebuilt a wave at a
// time (64 bits -> 8 bytes, written by the wave's first lane), so it needs no
// clearing; a row id out of range sets *err and gathers zero.
__global__ __launch_bounds__(256) void take_kernel(const uint8_t* __restrict__ src, const uint8_t* __restrict__ sval,
                                                   int64_t src_len, int width, const uint32_t* __restrict__ rows,
                                                   int64_t n, uint8_t* __restrict__ dst, uint8_t* __restrict__ dval,
                                                   uint32_t* __restrict__ err) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t end = (n + 63) / 64 * 64;  // wave-uniform trip count for the ballot
  const int64_t nbytes = (n + 7) / 8;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < end; i += stride) {
    const bool in = i < n;
    int64_t r = in ? (int64_t)rows[i] : 0;
    const bool null_row = r == 0xFFFFFFFFll;  // the NULL index of an outer join
    const bool ok = r < src_len && !null_row;
    if (in && !ok) atomicOr(err, null_row ? (dval ? 0u : 2u) : 1u);
    if (!ok) r = 0;
    if (in) {
      switch (width) {
        case 1: dst[i] = ok ? src[r] : 0; break;
        case 4: reinterpret_cast<uint32_t*>(dst)[i] = ok ? reinterpret_cast<const uint32_t*>(src)[r] : 0u; break;
        case 8: reinterpret_cast<uint64_t*>(dst)[i] = ok ? reinterpret_cast<const uint64_t*>(src)[r] : 0ull; break;
        case 16:
          reinterpret_cast<uint4*>(dst)[i] = ok ? reinterpret_cast<const uint4*>(src)[r] : make_uint4(0u, 0u, 0u, 0u);
          break;
      }
    }
    if (dval) {
      const uint64_t b = __ballot(in && ok && bit_valid(sval, r));
      if ((threadIdx.x & 63) == 0) {
        const int64_t b0 = i / 8;  // i is a multiple of 64 on the wave's first lane
        for (int k = 0; k < 8 && b0 + k < nbytes; k++) dval[b0 + k] = (uint8_t)(b >> (8 * k));
      }
    }
  }
}

}  // namespace bolt

using namespace bolt;

extern "C" int bolt_filter_compact(bolt_stream_t stream, const bolt_column_t* cols, int32_t n_cols,
                                   const bolt_filter_t* filter, const uint8_t* mask, bolt_column_t* out,
                                   int64_t* n_out) {
  BOLT_CHECK_DEVICE();
  clear_error();
  hipStream_t s = (hipStream_t)stream;
  *n_out = 0;
  if (n_cols < 1 || n_cols > kMaxCompactCols || cols == nullptr || out == nullptr)
    return set_error(BOLT_ERR_PLAN, "bolt_filter_compact: 1..%d columns required, got %d", kMaxCompactCols, n_cols);
  int64_t n = cols[0].length;
  GatherList gl;
  memset(&gl, 0, sizeof gl);
  gl.n = n_cols;
  gl.cap = out[0].length;
  for (int i = 0; i < n_cols; i++) {
    if (cols[i].length != n)
      return set_error(BOLT_ERR_PLAN, "bolt_filter_compact: column %d has %lld rows, column 0 has %lld", i,
                       (long long)cols[i].length, (long long)n);
    int w = dtype_size(cols[i].dtype);
    if (w == 0) return set_error(BOLT_ERR_TYPE, "bolt_filter_compact: column %d has unknown dtype", i);
    if (out[i].dtype != cols[i].dtype)
      return set_error(BOLT_ERR_PLAN, "bolt_filter_compact: output column %d dtype differs from input", i);
    if (out[i].length < 0)
      return set_error(BOLT_ERR_PLAN, "bolt_filter_compact: output column %d has a negative capacity", i);
    gl.cap = std::min<int64_t>(gl.cap, out[i].length);
    if (cols[i].validity && !out[i].validity)
      return set_error(BOLT_ERR_PLAN, "bolt_filter_compact: input column %d has NULLs, output has no bitmap", i);
    gl.c[i].src = static_cast<const uint8_t*>(cols[i].data);
    gl.c[i].dst = static_cast<uint8_t*>(out[i].data);
    gl.c[i].src_valid = cols[i].validity;
    gl.c[i].dst_valid = cols[i].validity ? reinterpret_cast<uint32_t*>(const_cast<uint8_t*>(out[i].validity)) : nullptr;
    gl.c[i].width = w;
    if (gl.c[i].dst_valid && (reinterpret_cast<uintptr_t>(gl.c[i].dst_valid) & 3))
      return set_error(BOLT_ERR_PLAN, "bolt_filter_compact: output bitmap %d must be 4-byte aligned", i);
  }
  DevFilter f;
  BOLT_TRY(make_dev_filter(filter, cols, n_cols, n, &f));
  if (n == 0) return BOLT_OK;
  for (int i = 0; i < n_cols; i++)
    if (gl.c[i].dst_valid) BOLT_TRY(fill_u32(gl.c[i].dst_valid, 0, (size_t)(std::min(gl.cap, n) + 31) / 32, s));
  int64_t n_tiles = (n + kTileRows - 1) / kTileRows;
  ScanScratch sc;
  BOLT_TRY(scan_prepare(sc, n_tiles, s));
  unsigned long long* total = reinterpret_cast<unsigned long long*>(sc.misc.ptr + 2);
  compact_kernel<0><<<(unsigned)n_tiles, kScanBlock, 0, s>>>(f, mask, n, gl, nullptr, sc.status.ptr, sc.misc.ptr,
                                                             total, sc.misc.ptr + 1);
  BOLT_HIP(hipGetLastError());
  uint32_t h[4];
  BOLT_HIP(hipMemcpyAsync(h, sc.misc.ptr, 16, hipMemcpyDeviceToHost, s));
  BOLT_HIP(hipStreamSynchronize(s));
  if (h[1]) return set_error(BOLT_ERR_OTHER, "bolt_filter_compact: look-back timed out");
  *n_out = (int64_t)(((uint64_t)h[3] << 32) | h[2]);
  if (*n_out > gl.cap)
    return set_error(BOLT_DECLINE, "bolt_filter_compact: %lld rows survive, the outputs hold %lld",
                     (long long)*n_out, (long long)gl.cap);
  return BOLT_OK;
}

extern "C" int bolt_prefix_scan_mask(bolt_stream_t stream, const uint8_t* mask, int64_t n, uint32_t* offsets,
                                     int64_t* total_out) {
  BOLT_CHECK_DEVICE();
  clear_error();
  hipStream_t s = (hipStream_t)stream;
  *total_out = 0;
  if (n < 0 || n > 0xFFFFFFFFll) return set_error(BOLT_ERR_PLAN, "bolt_prefix_scan_mask: n out of range");
  if (n == 0) return BOLT_OK;
  if (mask == nullptr) return set_error(BOLT_ERR_PLAN, "bolt_prefix_scan_mask: mask is NULL");
  int64_t n_tiles = (n + kTileRows - 1) / kTileRows;
  ScanScratch sc;
  BOLT_TRY(scan_prepare(sc, n_tiles, s));
  DevFilter f;
  memset(&f, 0, sizeof f);
  GatherList gl;
  memset(&gl, 0, sizeof gl);
  unsigned long long* total = reinterpret_cast<unsigned long long*>(sc.misc.ptr + 2);
  compact_kernel<1><<<(unsigned)n_tiles, kScanBlock, 0, s>>>(f, mask, n, gl, offsets, sc.status.ptr, sc.misc.ptr,
                                                             total, sc.misc.ptr + 1);
  BOLT_HIP(hipGetLastError());
  uint32_t h[4];
  BOLT_HIP(hipMemcpyAsync(h, sc.misc.ptr, 16, hipMemcpyDeviceToHost, s));
  BOLT_HIP(hipStreamSynchronize(s));
  if (h[1]) return set_error(BOLT_ERR_OTHER, "bolt_prefix_scan_mask: look-back timed out");
  *total_out = (int64_t)(((uint64_t)h[3] << 32) | h[2]);
  return BOLT_OK;
}

extern "C" int bolt_take(bolt_stream_t stream, const bolt_column_t* src, const uint32_t* rows, int64_t n,
                         bolt_column_t* dst) {
  BOLT_CHECK_DEVICE();
  clear_error();
  hipStream_t s = (hipStream_t)stream;
  if (!src || !dst) return set_error(BOLT_ERR_PLAN, "bolt_take: NULL column");
  if (n < 0) return set_error(BOLT_ERR_PLAN, "bolt_take: negative length");
  const int width = dtype_size(src->dtype);
  if (width == 0) return set_error(BOLT_ERR_TYPE, "bolt_take: unknown dtype %d", src->dtype);
  if (src->validity && !dst->validity)
    return set_error(BOLT_ERR_PLAN, "bolt_take: the source has a validity bitmap, the destination none");
  dst->dtype = src->dtype;
  dst->scale = src->scale;
  dst->length = n;
  if (n == 0) return BOLT_OK;
  if (!rows || !dst->data) return set_error(BOLT_ERR_PLAN, "bolt_take: NULL rows or destination");
  DevBuf<uint32_t> err;
  BOLT_TRY(err.alloc(1, s));
  BOLT_TRY(fill_u32(err.ptr, 0, 1, s));
  take_kernel<<<grid_for(n, 256, 4), 256, 0, s>>>(static_cast<const uint8_t*>(src->data), src->validity, src->length,
                                                  width, rows, n, static_cast<uint8_t*>(const_cast<void*>(dst->data)),
                                                  const_cast<uint8_t*>(dst->validity),
                                                  err.ptr);
  BOLT_HIP(hipGetLastError());
  uint32_t e = 0;
  BOLT_HIP(hipMemcpyAsync(&e, err.ptr, 4, hipMemcpyDeviceToHost, s));
  BOLT_HIP(hipStreamSynchronize(s));
  if (e & 1u) return set_error(BOLT_ERR_PLAN, "bolt_take: a row id is out of range (source has %lld rows)",
                               (long long)src->length);
  if (e & 2u) return set_error(BOLT_ERR_PLAN, "bolt_take: a NULL row id needs a destination validity bitmap");
  return BOLT_OK;
}
