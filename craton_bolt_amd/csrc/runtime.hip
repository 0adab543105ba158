// runtime.hip — the bolt_hip_* device-runtime C ABI over HIP.
//
// Mirrors the CUDA-driver entry points bound in src/cuda/cuda_sys.rs:89-224
// one for one (see include/bolt_hip.h).  HIP has one primary context per
// device, so the reference's context calls reduce to device selection.
// Allocation goes through HIP's stream-ordered memory pool with the release
// threshold raised, so freed blocks stay cached for the next query instead of
// returning to the driver (the role of DeviceMemPool, src/cuda/mem_pool.rs:1144).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "bolt_internal.h"

namespace bolt {

static thread_local std::string g_last_error;
static thread_local hipEvent_t g_prof_start = nullptr;
static thread_local hipEvent_t g_prof_end = nullptr;

int set_error(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}
void clear_error() { g_last_error.clear(); }

void prof_begin(hipStream_t s) {
  if (g_prof_start) (void)hipEventRecord(g_prof_start, s);
}
void prof_end(hipStream_t s) {
  if (g_prof_end) (void)hipEventRecord(g_prof_end, s);
}

static std::once_flag g_dev_once;
static int g_dev_status = BOLT_ERR_NO_DEVICE;

int device_status() {
  std::call_once(g_dev_once, [] {
    int n = 0;
    if (hipGetDeviceCount(&n) == hipSuccess && n > 0) {
      g_dev_status = 0;
      // Keep freed pool memory cached across queries.
      for (int d = 0; d < n; d++) {
        hipMemPool_t pool;
        if (hipDeviceGetDefaultMemPool(&pool, d) == hipSuccess) {
          uint64_t thr = ~0ull;
          (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
        }
      }
    } else {
      (void)hipGetLastError();
    }
  });
  if (g_dev_status != 0) set_error(BOLT_ERR_NO_DEVICE, "no HIP device available");
  return g_dev_status;
}

// ---- pool accounting --------------------------------------------------------

static std::mutex g_pool_mu;
static std::unordered_map<void*, size_t> g_live;
static size_t g_in_use = 0, g_peak = 0;

static void account_alloc(void* p, size_t bytes) {
  std::lock_guard<std::mutex> lk(g_pool_mu);
  g_live[p] = bytes;
  g_in_use += bytes;
  if (g_in_use > g_peak) g_peak = g_in_use;
}
static void account_free(void* p) {
  std::lock_guard<std::mutex> lk(g_pool_mu);
  auto it = g_live.find(p);
  if (it != g_live.end()) {
    g_in_use -= it->second;
    g_live.erase(it);
  }
}

// ---- block cache ------------------------------------------------------------
// ROCm's stream-ordered allocator maps and unmaps large blocks on every
// hipMallocAsync / hipFreeAsync even with the pool's release threshold at its
// maximum (about 0.5 ms of host time per GiB on MI355X), which was a quarter
// of a 256M-row GROUP BY call, and each hipFreeAsync of a small block costs
// 8-28 us of host time (about 70 us of a 1000-group GROUP BY call, whose
// temporaries are freed before it returns).  Blocks freed outside graph
// capture are therefore kept per device and handed to later allocations:
// blocks under 32 MiB by power-of-two size class (a request is rounded up to
// its class when allocated), larger ones to a request of at most their size
// and more than half of it.  Each kept block carries an event recorded on the
// freeing stream.  A block is reused on the stream that freed it, where the
// stream's own order puts the reuse after every use before the free, or on
// another stream once that event has completed; no stream ever waits on
// another's queued work through the cache, and no wait (a barrier in the
// queue, which keeps kernels from overlapping) is added.  Blocks freed on a
// stream destroyed through bolt_hip_stream_destroy lose the same-stream
// shortcut, so a later stream with the same handle checks completion too.
// The events come from a pool, so a free costs one record.  At most 32 GiB is kept; an allocation that fails empties the cache
// and is retried.
struct CachedBlock {
  void* p;
  size_t bytes;
  hipEvent_t freed;
  int dev;
  hipStream_t stream;  // where it was freed
};
static std::mutex g_cache_mu;
static std::vector<CachedBlock> g_cache;                                    // blocks of kCacheMinBlock and more
static std::unordered_map<uint64_t, std::vector<CachedBlock>> g_small;      // (device, class) -> blocks
static std::unordered_map<void*, size_t> g_block_bytes;  // size of every block from hipMallocAsync
static std::unordered_map<int, std::vector<hipEvent_t>> g_event_pool;       // device -> idle events
static size_t g_cached_bytes = 0;
constexpr size_t kCacheMinBlock = (size_t)32 << 20;
constexpr size_t kCacheMaxBytes = (size_t)32 << 30;
constexpr size_t kSmallClassMin = 256;

static bool capturing(hipStream_t s) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return cs != hipStreamCaptureStatusNone;
}

// The size a request is allocated at: its power-of-two class below
// kCacheMinBlock, itself above.
static size_t block_size_for(size_t bytes) {
  if (bytes >= kCacheMinBlock) return bytes;
  size_t c = kSmallClassMin;
  while (c < bytes) c <<= 1;
  return c;
}

static uint64_t small_key(int dev, size_t bytes) { return ((uint64_t)(uint32_t)dev << 40) | (uint64_t)bytes; }

// An idle event of device `dev` (the caller holds no lock), or null.
static hipEvent_t event_get(int dev) {
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    auto& pool = g_event_pool[dev];
    if (!pool.empty()) {
      const hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
  }
  hipEvent_t e = nullptr;
  if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return e;
}

// Called with g_cache_mu held.
static void event_put_locked(int dev, hipEvent_t e) { g_event_pool[dev].push_back(e); }

static void* cache_take(size_t bytes, hipStream_t s) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  // A block freed on this stream is taken first (its event is behind this
  // stream's own order); one freed on another stream only once its event has
  // completed, so that concurrent streams never wait on each other's queued
  // work through the cache (a block still in flight elsewhere is left, and a
  // new one allocated).
  auto usable = [&](const CachedBlock& b, bool same_only) {
    if (b.stream == s) return true;
    if (same_only) return false;
    const hipError_t q = hipEventQuery(b.freed);
    if (q != hipSuccess) (void)hipGetLastError();
    return q == hipSuccess;
  };
  CachedBlock got{};
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    bool found = false;
    if (bytes < kCacheMinBlock) {
      auto it = g_small.find(small_key(dev, bytes));
      if (it == g_small.end() || it->second.empty()) return nullptr;
      std::vector<CachedBlock>& v = it->second;
      for (int pass = 0; pass < 2 && !found; pass++) {
        // this stream's newest block first; else one of the eight oldest
        // (the likeliest to have completed) freed elsewhere
        const size_t lim = pass == 0 ? v.size() : std::min<size_t>(v.size(), 8);
        for (size_t k = 0; k < lim && !found; k++) {
          const size_t i = pass == 0 ? v.size() - 1 - k : k;
          if (!usable(v[i], pass == 0)) continue;
          got = v[i];
          v.erase(v.begin() + (std::ptrdiff_t)i);
          found = true;
        }
      }
    } else {
      for (int pass = 0; pass < 2 && !found; pass++) {
        size_t best = g_cache.size();
        for (size_t i = 0; i < g_cache.size(); i++) {
          const CachedBlock& b = g_cache[i];
          if (b.dev == dev && b.bytes >= bytes && b.bytes / 2 < bytes &&
              (best == g_cache.size() || b.bytes < g_cache[best].bytes) && usable(b, pass == 0))
            best = i;
        }
        if (best == g_cache.size()) continue;
        got = g_cache[best];
        g_cache.erase(g_cache.begin() + (std::ptrdiff_t)best);
        found = true;
      }
    }
    if (!found) return nullptr;
    g_cached_bytes -= got.bytes;
  }
  // no stream wait: a block is taken on the stream that freed it (ordered by
  // that stream) or after its event completed.  A wait would put a barrier in
  // the stream's queue and keep its kernels from overlapping.
  std::lock_guard<std::mutex> lk(g_cache_mu);
  event_put_locked(dev, got.freed);
  return got.p;
}

static bool cache_put(void* p, hipStream_t s) {
  size_t bytes = 0;
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    const auto it = g_block_bytes.find(p);
    if (it == g_block_bytes.end() || g_cached_bytes + it->second > kCacheMaxBytes) return false;
    bytes = it->second;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  const hipEvent_t ev = event_get(dev);
  if (!ev) return false;
  if (hipEventRecord(ev, s) != hipSuccess) {
    (void)hipGetLastError();
    std::lock_guard<std::mutex> lk(g_cache_mu);
    event_put_locked(dev, ev);
    return false;
  }
  std::lock_guard<std::mutex> lk(g_cache_mu);
  const CachedBlock b{p, bytes, ev, dev, s};
  if (bytes < kCacheMinBlock) g_small[small_key(dev, bytes)].push_back(b);
  else g_cache.push_back(b);
  g_cached_bytes += bytes;
  return true;
}

// Blocks freed on `s` may only be reused once complete (the stream is going away).
void cache_forget_stream(hipStream_t s) {
  const hipStream_t gone = reinterpret_cast<hipStream_t>(~(uintptr_t)0);
  std::lock_guard<std::mutex> lk(g_cache_mu);
  for (CachedBlock& b : g_cache)
    if (b.stream == s) b.stream = gone;
  for (auto& kv : g_small)
    for (CachedBlock& b : kv.second)
      if (b.stream == s) b.stream = gone;
}

static void cache_flush() {
  std::vector<CachedBlock> all;
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    all.swap(g_cache);
    for (auto& kv : g_small)
      for (const CachedBlock& b : kv.second) all.push_back(b);
    g_small.clear();
    g_cached_bytes = 0;
    for (const CachedBlock& b : all) g_block_bytes.erase(b.p);
  }
  for (const CachedBlock& b : all) {
    (void)hipEventSynchronize(b.freed);
    (void)hipFree(b.p);
    std::lock_guard<std::mutex> lk(g_cache_mu);
    event_put_locked(b.dev, b.freed);
  }
}

int dev_alloc(void** p, size_t bytes, hipStream_t s) {
  BOLT_CHECK_DEVICE();
  *p = nullptr;
  const bool cap = capturing(s);
  // outside capture a request takes its block size (class), so the block can
  // be kept and reused; under capture it is allocated as asked and never kept
  const size_t alloc_bytes = cap ? (bytes ? bytes : 1) : block_size_for(bytes ? bytes : 1);
  if (!cap) *p = cache_take(alloc_bytes, s);
  if (!*p) {
    hipError_t e = hipMallocAsync(p, alloc_bytes, s);
    if (e != hipSuccess && !cap) {
      (void)hipGetLastError();
      cache_flush();
      e = hipMallocAsync(p, alloc_bytes, s);
    }
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return set_error(BOLT_ERR_MEMORY, "device allocation of %zu bytes failed: %s", bytes,
                       hipGetErrorString(e));
    }
    if (!cap) {
      std::lock_guard<std::mutex> lk(g_cache_mu);
      g_block_bytes[*p] = alloc_bytes;
    }
  }
  account_alloc(*p, bytes);
  return BOLT_OK;
}

int dev_free(void* p, hipStream_t s) {
  if (!p) return BOLT_OK;
  account_free(p);
  if (!capturing(s) && cache_put(p, s)) return BOLT_OK;
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    g_block_bytes.erase(p);
  }
  BOLT_HIP(hipFreeAsync(p, s));
  return BOLT_OK;
}

// ---- fill kernels (graph-capture safe) ----------------------------------------

__global__ void fill_u32_kernel(uint32_t* __restrict__ dst, uint32_t v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = v;
}
__global__ void fill_u8_kernel(uint8_t* __restrict__ dst, uint8_t v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = v;
}

int fill_u32(void* dst, uint32_t value, size_t n_words, hipStream_t s) {
  if (n_words == 0) return BOLT_OK;
  unsigned grid = grid_for((int64_t)n_words, 256, 4);
  fill_u32_kernel<<<grid, 256, 0, s>>>(static_cast<uint32_t*>(dst), value, n_words);
  BOLT_HIP(hipGetLastError());
  return BOLT_OK;
}
int fill_u8(void* dst, uint8_t value, size_t n, hipStream_t s) {
  if (n == 0) return BOLT_OK;
  // word-aligned bulk as u32, tail as bytes
  uintptr_t a = reinterpret_cast<uintptr_t>(dst);
  if ((a & 3) == 0 && n >= 4) {
    uint32_t w = value * 0x01010101u;
    BOLT_TRY(fill_u32(dst, w, n / 4, s));
    size_t done = n & ~size_t(3);
    if (done == n) return BOLT_OK;
    dst = static_cast<uint8_t*>(dst) + done;
    n -= done;
  }
  unsigned grid = grid_for((int64_t)n, 256, 4);
  fill_u8_kernel<<<grid, 256, 0, s>>>(static_cast<uint8_t*>(dst), value, n);
  BOLT_HIP(hipGetLastError());
  return BOLT_OK;
}

const char* dtype_name(int32_t dt) {
  switch (dt) {
    case BOLT_BOOL: return "Bool";
    case BOLT_I32: return "Int32";
    case BOLT_I64: return "Int64";
    case BOLT_F32: return "Float32";
    case BOLT_F64: return "Float64";
    case BOLT_DEC128: return "Decimal128";
    case BOLT_DATE32: return "Date32";
    case BOLT_TIMESTAMP: return "Timestamp";
  }
  return "Unknown";
}

// ---- stream pool (src/cuda/stream_pool.rs:136,170) ----------------------------

static std::mutex g_stream_mu;
static std::vector<hipStream_t> g_free_streams;

}  // namespace bolt

using namespace bolt;

#define RT_CHECK() BOLT_CHECK_DEVICE()
#define RT_HIP(expr)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) {                                                               \
      set_error((int)e_, "%s failed: %s", #expr, hipGetErrorString(e_));                 \
      return (int)e_;                                                                     \
    }                                                                                     \
  } while (0)

extern "C" {

const char* bolt_last_error(void) { return g_last_error.c_str(); }

const char* bolt_build_info(void) { return "craton_bolt_amd 0.1 gfx950"; }

int bolt_profile_events(bolt_event_t start, bolt_event_t end) {
  g_prof_start = reinterpret_cast<hipEvent_t>(start);
  g_prof_end = reinterpret_cast<hipEvent_t>(end);
  return 0;
}

static const char* status_name(int e) {
  if (e == BOLT_HIP_ERROR_NO_DEVICE) return "BOLT_HIP_ERROR_NO_DEVICE";
  const char* n = hipGetErrorName((hipError_t)e);
  return (n && strcmp(n, "hipErrorUnknown") != 0) || e == (int)hipErrorUnknown ? n : nullptr;
}

int bolt_hip_get_error_name(int error, const char** str) {
  *str = status_name(error);
  return *str ? 0 : (int)hipErrorInvalidValue;
}

int bolt_hip_get_error_string(int error, const char** str) {
  if (error == BOLT_HIP_ERROR_NO_DEVICE) {
    *str = "no usable HIP device (the library was loaded on a host without a GPU)";
    return 0;
  }
  if (!status_name(error)) {
    *str = nullptr;
    return (int)hipErrorInvalidValue;
  }
  *str = hipGetErrorString((hipError_t)error);
  return 0;
}

// A context handle is device + 1, so that NULL stays "no context".
static thread_local bolt_ctx_t t_current_ctx = nullptr;
static bolt_ctx_t ctx_of(int device) { return reinterpret_cast<bolt_ctx_t>((uintptr_t)device + 1); }
static int device_of(bolt_ctx_t ctx) { return (int)(reinterpret_cast<uintptr_t>(ctx) - 1); }

int bolt_hip_device_get(int* device, int ordinal) {
  RT_CHECK();
  int n = 0;
  RT_HIP(hipGetDeviceCount(&n));
  if (ordinal < 0 || ordinal >= n)
    return set_error((int)hipErrorInvalidDevice, "bolt_hip_device_get: ordinal %d of %d devices", ordinal, n);
  *device = ordinal;
  return 0;
}

int bolt_hip_ctx_create(bolt_ctx_t* ctx, unsigned int flags, int device) {
  (void)flags;
  *ctx = nullptr;
  RT_CHECK();
  hipCtx_t primary;
  RT_HIP(hipDevicePrimaryCtxRetain(&primary, device));
  RT_HIP(hipSetDevice(device));
  *ctx = t_current_ctx = ctx_of(device);
  return 0;
}

int bolt_hip_ctx_destroy(bolt_ctx_t ctx) {
  RT_CHECK();
  if (!ctx) return set_error((int)hipErrorInvalidContext, "bolt_hip_ctx_destroy: NULL context");
  RT_HIP(hipDevicePrimaryCtxRelease(device_of(ctx)));
  if (t_current_ctx == ctx) t_current_ctx = nullptr;
  return 0;
}

int bolt_hip_ctx_set_current(bolt_ctx_t ctx) {
  RT_CHECK();
  if (ctx) RT_HIP(hipSetDevice(device_of(ctx)));
  t_current_ctx = ctx;
  return 0;
}

int bolt_hip_ctx_get_current(bolt_ctx_t* ctx) {
  RT_CHECK();
  *ctx = t_current_ctx;
  return 0;
}

int bolt_hip_init(unsigned int flags) {
  if (flags != 0) return set_error(hipErrorInvalidValue, "bolt_hip_init: flags must be 0");
  RT_CHECK();
  RT_HIP(hipInit(0));
  return 0;
}

int bolt_hip_device_get_count(int* count) {
  *count = 0;
  RT_CHECK();
  RT_HIP(hipGetDeviceCount(count));
  return 0;
}

int bolt_hip_device_get_name(char* name, int len, int device) {
  RT_CHECK();
  hipDeviceProp_t p;
  RT_HIP(hipGetDeviceProperties(&p, device));
  snprintf(name, (size_t)len, "%s", p.name);
  return 0;
}

int bolt_hip_device_get_attribute(int* value, int attrib, int device) {
  RT_CHECK();
  RT_HIP(hipDeviceGetAttribute(value, (hipDeviceAttribute_t)attrib, device));
  return 0;
}

int bolt_hip_device_total_mem(size_t* bytes, int device) {
  RT_CHECK();
  RT_HIP(hipDeviceTotalMem(bytes, device));
  return 0;
}

int bolt_hip_set_device(int device) {
  RT_CHECK();
  RT_HIP(hipSetDevice(device));
  return 0;
}

int bolt_hip_get_device(int* device) {
  RT_CHECK();
  RT_HIP(hipGetDevice(device));
  return 0;
}

int bolt_hip_device_synchronize(void) {
  RT_CHECK();
  RT_HIP(hipDeviceSynchronize());
  return 0;
}

int bolt_hip_malloc(void** dptr, size_t bytes) {
  RT_CHECK();
  *dptr = nullptr;
  RT_HIP(hipMallocAsync(dptr, bytes ? bytes : 1, nullptr));
  RT_HIP(hipStreamSynchronize(nullptr));
  account_alloc(*dptr, bytes);
  return 0;
}

int bolt_hip_free(void* dptr) {
  RT_CHECK();
  if (!dptr) return 0;
  account_free(dptr);
  RT_HIP(hipFreeAsync(dptr, nullptr));
  RT_HIP(hipStreamSynchronize(nullptr));
  return 0;
}

int bolt_hip_malloc_async(void** dptr, size_t bytes, bolt_stream_t stream) {
  RT_CHECK();
  *dptr = nullptr;
  RT_HIP(hipMallocAsync(dptr, bytes ? bytes : 1, (hipStream_t)stream));
  account_alloc(*dptr, bytes);
  return 0;
}

int bolt_hip_free_async(void* dptr, bolt_stream_t stream) {
  RT_CHECK();
  if (!dptr) return 0;
  account_free(dptr);
  RT_HIP(hipFreeAsync(dptr, (hipStream_t)stream));
  return 0;
}

int bolt_hip_host_alloc(void** ptr, size_t bytes, unsigned int flags) {
  RT_CHECK();
  (void)flags;
  RT_HIP(hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault));
  return 0;
}

int bolt_hip_host_free(void* ptr) {
  RT_CHECK();
  RT_HIP(hipHostFree(ptr));
  return 0;
}

int bolt_hip_mem_get_info(size_t* free_bytes, size_t* total_bytes) {
  RT_CHECK();
  RT_HIP(hipMemGetInfo(free_bytes, total_bytes));
  return 0;
}

// Staged upload of a large pageable host buffer: kUpLanes host threads each
// copy a contiguous share through two pinned buffers of kUpChunk bytes and
// hand every filled buffer to the DMA engines on their own stream, so the
// host copies into pinned memory and the PCIe transfers overlap.  (A plain
// hipMemcpy from pageable memory moves about 26 GB/s on this path.)  The
// pinned buffers and streams are made once and kept; one staged upload runs
// at a time, and a concurrent caller takes the plain copy.
namespace {
constexpr int kUpLanes = 4;
constexpr size_t kUpChunk = (size_t)4 << 20;
constexpr size_t kUpMin = (size_t)256 << 10;  // smaller copies take hipMemcpy
constexpr size_t kUpPerLane = (size_t)2 << 20;  // a lane per 2 MB, up to kUpLanes

struct UpLane {
  char* buf[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
  hipStream_t stream = nullptr;
};
std::mutex g_up_mu;
UpLane g_up[kUpLanes];
bool g_up_ready = false;

bool is_pageable(const void* p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();  // unknown to HIP: plain pageable memory
    return true;
  }
  return a.type == hipMemoryTypeUnregistered;  // pinned / registered / device memory is copied as it is
}

// Creates what is still missing (a failed attempt resumes where it stopped;
// the staged path is skipped until everything exists).
int up_init() {
  if (g_up_ready) return 0;
  for (UpLane& l : g_up) {
    if (!l.stream && hipStreamCreateWithFlags(&l.stream, hipStreamNonBlocking) != hipSuccess) {
      l.stream = nullptr;
      (void)hipGetLastError();
      return -1;
    }
    for (int b = 0; b < 2; b++) {
      if (!l.buf[b] && hipHostMalloc(reinterpret_cast<void**>(&l.buf[b]), kUpChunk, hipHostMallocDefault) != hipSuccess) {
        l.buf[b] = nullptr;
        (void)hipGetLastError();
        return -1;
      }
      if (!l.done[b] && hipEventCreateWithFlags(&l.done[b], hipEventDisableTiming) != hipSuccess) {
        l.done[b] = nullptr;
        (void)hipGetLastError();
        return -1;
      }
    }
  }
  g_up_ready = true;
  return 0;
}

// Copies [off, off + len) of src into dst through lane l.
hipError_t up_lane(UpLane& l, char* dst, const char* src, size_t len) {
  bool used[2] = {false, false};
  int b = 0;
  for (size_t pos = 0; pos < len; pos += kUpChunk, b ^= 1) {
    const size_t n = len - pos < kUpChunk ? len - pos : kUpChunk;
    if (used[b]) {
      hipError_t e = hipEventSynchronize(l.done[b]);  // the buffer's previous transfer has finished
      if (e != hipSuccess) return e;
    }
    memcpy(l.buf[b], src + pos, n);
    hipError_t e = hipMemcpyAsync(dst + pos, l.buf[b], n, hipMemcpyHostToDevice, l.stream);
    if (e == hipSuccess) e = hipEventRecord(l.done[b], l.stream);
    if (e != hipSuccess) return e;
    used[b] = true;
  }
  return hipStreamSynchronize(l.stream);
}
}  // namespace

int bolt_hip_memcpy_htod(void* dst, const void* src, size_t bytes) {
  RT_CHECK();
  std::unique_lock<std::mutex> lk(g_up_mu, std::try_to_lock);
  if (bytes < kUpMin || !lk.owns_lock() || !is_pageable(src) || up_init() != 0) {
    if (lk.owns_lock()) lk.unlock();
    RT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return 0;
  }
  // a synchronous copy is ordered after the device's earlier work, as hipMemcpy is
  RT_HIP(hipDeviceSynchronize());
  int lanes = (int)(bytes / kUpPerLane);
  lanes = lanes < 1 ? 1 : (lanes > kUpLanes ? kUpLanes : lanes);
  if (lanes == 1) {  // one lane: in this thread
    RT_HIP(up_lane(g_up[0], static_cast<char*>(dst), static_cast<const char*>(src), bytes));
    return 0;
  }
  // ceil(bytes / lanes) rounded up to 4 KB, so that lanes * share >= bytes and
  // the last lane takes what remains (possibly nothing)
  const size_t share = ((bytes + lanes - 1) / lanes + 4095) / 4096 * 4096;
  hipError_t err[kUpLanes];
  std::thread th[kUpLanes];
  for (int i = 0; i < lanes; i++) {
    const size_t off = (size_t)i * share;
    const size_t len = off >= bytes ? 0 : (bytes - off < share ? bytes - off : share);
    err[i] = hipSuccess;
    th[i] = std::thread([&, i, off, len] {
      if (len) err[i] = up_lane(g_up[i], static_cast<char*>(dst) + off, static_cast<const char*>(src) + off, len);
    });
  }
  for (int i = 0; i < lanes; i++) th[i].join();
  for (int i = 0; i < lanes; i++) RT_HIP(err[i]);
  return 0;
}
int bolt_hip_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
  RT_CHECK();
  RT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return 0;
}
int bolt_hip_memcpy_dtod(void* dst, const void* src, size_t bytes) {
  RT_CHECK();
  RT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return 0;
}
int bolt_hip_memcpy_htod_async(void* dst, const void* src, size_t bytes, bolt_stream_t stream) {
  RT_CHECK();
  RT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
  return 0;
}
int bolt_hip_memcpy_dtoh_async(void* dst, const void* src, size_t bytes, bolt_stream_t stream) {
  RT_CHECK();
  RT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
  return 0;
}
int bolt_hip_memcpy_dtod_async(void* dst, const void* src, size_t bytes, bolt_stream_t stream) {
  RT_CHECK();
  RT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return 0;
}

int bolt_hip_memset_d8(void* dst, unsigned char value, size_t count) {
  RT_CHECK();
  int st = fill_u8(dst, value, count, nullptr);
  if (st != BOLT_OK) return hipErrorLaunchFailure;
  RT_HIP(hipStreamSynchronize(nullptr));
  return 0;
}
int bolt_hip_memset_d8_async(void* dst, unsigned char value, size_t count, bolt_stream_t stream) {
  RT_CHECK();
  int st = fill_u8(dst, value, count, (hipStream_t)stream);
  return st == BOLT_OK ? 0 : (int)hipErrorLaunchFailure;
}

int bolt_hip_module_load_data(bolt_module_t* module, const void* image) {
  RT_CHECK();
  hipModule_t m;
  RT_HIP(hipModuleLoadData(&m, image));
  *module = (bolt_module_t)m;
  return 0;
}
int bolt_hip_module_load_data_ex(bolt_module_t* module, const void* image, unsigned int num_options, int* options,
                                 void** option_values) {
  (void)num_options;
  (void)options;
  (void)option_values;
  return bolt_hip_module_load_data(module, image);
}
int bolt_hip_module_unload(bolt_module_t module) {
  RT_CHECK();
  RT_HIP(hipModuleUnload((hipModule_t)module));
  return 0;
}
int bolt_hip_module_get_function(bolt_function_t* func, bolt_module_t module, const char* name) {
  RT_CHECK();
  hipFunction_t f;
  RT_HIP(hipModuleGetFunction(&f, (hipModule_t)module, name));
  *func = (bolt_function_t)f;
  return 0;
}
int bolt_hip_launch_kernel(bolt_function_t func, unsigned int gx, unsigned int gy, unsigned int gz,
                           unsigned int bx, unsigned int by, unsigned int bz, unsigned int shmem,
                           bolt_stream_t stream, void** kernel_params, void** extra) {
  RT_CHECK();
  RT_HIP(hipModuleLaunchKernel((hipFunction_t)func, gx, gy, gz, bx, by, bz, shmem, (hipStream_t)stream,
                               kernel_params, extra));
  return 0;
}

int bolt_hip_stream_create(bolt_stream_t* stream, unsigned int flags) {
  RT_CHECK();
  hipStream_t s;
  RT_HIP(hipStreamCreateWithFlags(&s, flags ? hipStreamNonBlocking : hipStreamDefault));
  *stream = (bolt_stream_t)s;
  return 0;
}
int bolt_hip_stream_destroy(bolt_stream_t stream) {
  RT_CHECK();
  bolt::cache_forget_stream((hipStream_t)stream);
  RT_HIP(hipStreamDestroy((hipStream_t)stream));
  return 0;
}
int bolt_hip_stream_synchronize(bolt_stream_t stream) {
  RT_CHECK();
  RT_HIP(hipStreamSynchronize((hipStream_t)stream));
  return 0;
}
int bolt_hip_stream_pool_acquire(bolt_stream_t* stream) {
  RT_CHECK();
  {
    std::lock_guard<std::mutex> lk(g_stream_mu);
    if (!g_free_streams.empty()) {
      *stream = (bolt_stream_t)g_free_streams.back();
      g_free_streams.pop_back();
      return 0;
    }
  }
  hipStream_t s;
  RT_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = (bolt_stream_t)s;
  return 0;
}
int bolt_hip_stream_pool_release(bolt_stream_t stream) {
  RT_CHECK();
  std::lock_guard<std::mutex> lk(g_stream_mu);
  g_free_streams.push_back((hipStream_t)stream);
  return 0;
}

int bolt_hip_event_create(bolt_event_t* event, unsigned int flags) {
  RT_CHECK();
  hipEvent_t e;
  RT_HIP(hipEventCreateWithFlags(&e, (flags & 2) ? hipEventDisableTiming : hipEventDefault));
  *event = (bolt_event_t)e;
  return 0;
}
int bolt_hip_event_record(bolt_event_t event, bolt_stream_t stream) {
  RT_CHECK();
  RT_HIP(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
  return 0;
}
int bolt_hip_event_query(bolt_event_t event) {
  RT_CHECK();
  hipError_t e = hipEventQuery((hipEvent_t)event);
  if (e == hipErrorNotReady) {
    (void)hipGetLastError();
    return BOLT_HIP_ERROR_NOT_READY;
  }
  RT_HIP(e);
  return 0;
}
int bolt_hip_event_synchronize(bolt_event_t event) {
  RT_CHECK();
  RT_HIP(hipEventSynchronize((hipEvent_t)event));
  return 0;
}
int bolt_hip_event_elapsed_ms(float* ms, bolt_event_t start, bolt_event_t end) {
  RT_CHECK();
  RT_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)end));
  return 0;
}
int bolt_hip_event_destroy(bolt_event_t event) {
  RT_CHECK();
  RT_HIP(hipEventDestroy((hipEvent_t)event));
  return 0;
}

int bolt_hip_stream_begin_capture(bolt_stream_t stream, unsigned int mode) {
  RT_CHECK();
  RT_HIP(hipStreamBeginCapture((hipStream_t)stream, (hipStreamCaptureMode)mode));
  return 0;
}
int bolt_hip_stream_end_capture(bolt_stream_t stream, bolt_graph_t* graph) {
  RT_CHECK();
  hipGraph_t g;
  RT_HIP(hipStreamEndCapture((hipStream_t)stream, &g));
  *graph = (bolt_graph_t)g;
  return 0;
}
int bolt_hip_graph_instantiate(bolt_graph_exec_t* exec, bolt_graph_t graph, void** error_node, char* log_buffer,
                               size_t buffer_size) {
  RT_CHECK();
  hipGraphExec_t x;
  hipGraphNode_t bad = nullptr;
  if (log_buffer && buffer_size) log_buffer[0] = '\0';
  RT_HIP(hipGraphInstantiate(&x, (hipGraph_t)graph, &bad, log_buffer, log_buffer ? buffer_size : 0));
  if (error_node) *error_node = bad;
  *exec = (bolt_graph_exec_t)x;
  return 0;
}
int bolt_hip_graph_launch(bolt_graph_exec_t exec, bolt_stream_t stream) {
  RT_CHECK();
  RT_HIP(hipGraphLaunch((hipGraphExec_t)exec, (hipStream_t)stream));
  return 0;
}
int bolt_hip_graph_exec_destroy(bolt_graph_exec_t exec) {
  RT_CHECK();
  RT_HIP(hipGraphExecDestroy((hipGraphExec_t)exec));
  return 0;
}
int bolt_hip_graph_destroy(bolt_graph_t graph) {
  RT_CHECK();
  RT_HIP(hipGraphDestroy((hipGraph_t)graph));
  return 0;
}

int bolt_hip_pool_stats(size_t* in_use_bytes, size_t* peak_bytes) {
  std::lock_guard<std::mutex> lk(g_pool_mu);
  *in_use_bytes = g_in_use;
  *peak_bytes = g_peak;
  return 0;
}

}  // extern "C"
