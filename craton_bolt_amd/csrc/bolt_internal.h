// bolt_internal.h — shared host/device helpers of the MI355X execution core.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/bolt_ops.h"
#include "device_acc.h"

namespace bolt {

// ---- errors ---------------------------------------------------------------

// Records `msg` as the calling thread's last error and returns `code`.
int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

// 0 when a GPU is usable, BOLT_ERR_NO_DEVICE otherwise (cached after the first call).
int device_status();

#define BOLT_CHECK_DEVICE()                 \
  do {                                      \
    int st_ = ::bolt::device_status();      \
    if (st_ != 0) return st_;               \
  } while (0)

// Runs a HIP call; on failure records "<expr>: <hip error text>" and returns
// BOLT_ERR_DEVICE from the enclosing function.
#define BOLT_HIP(expr)                                                                    \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return ::bolt::set_error(BOLT_ERR_DEVICE, "%s failed: %s (%d) at %s:%d", #expr,    \
                               hipGetErrorString(e_), (int)e_, __FILE__, __LINE__);       \
  } while (0)

// Propagates a non-zero status.
#define BOLT_TRY(expr)          \
  do {                          \
    int s_ = (expr);            \
    if (s_ != BOLT_OK) return s_; \
  } while (0)

// ---- measurement hook -----------------------------------------------------

// Events registered with bolt_profile_events on this thread (nullptr = none).
void prof_begin(hipStream_t s);
void prof_end(hipStream_t s);

// ---- memory ---------------------------------------------------------------

// Stream-ordered allocation from the device pool (accounted in pool stats).
int dev_alloc(void** p, size_t bytes, hipStream_t s);
int dev_free(void* p, hipStream_t s);
// Before destroying stream s: blocks the device cache holds from frees on s are
// then reused only once complete.
void cache_forget_stream(hipStream_t s);

// RAII device allocation freed (stream-ordered) on destruction.
template <typename T>
struct DevBuf {
  T* ptr = nullptr;
  size_t count = 0;
  hipStream_t stream = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), count(o.count), stream(o.stream) { o.ptr = nullptr; }
  ~DevBuf() { reset(); }
  int alloc(size_t n, hipStream_t s) {
    reset();
    stream = s;
    count = n;
    void* p = nullptr;
    int st = dev_alloc(&p, (n ? n : 1) * sizeof(T), s);
    ptr = static_cast<T*>(p);
    return st;
  }
  void reset() {
    if (ptr) dev_free(ptr, stream);
    ptr = nullptr;
    count = 0;
  }
  T* release() {
    T* p = ptr;
    ptr = nullptr;
    return p;
  }
};

// Graph-capture-safe fill (a kernel, never hipMemsetAsync).
int fill_u32(void* dst, uint32_t value, size_t n_words, hipStream_t s);
int fill_u8(void* dst, uint8_t value, size_t n, hipStream_t s);

// ---- dtypes ---------------------------------------------------------------

inline int dtype_size(int32_t dt) {
  switch (dt) {
    case BOLT_BOOL: return 1;
    case BOLT_I32: case BOLT_F32: case BOLT_DATE32: return 4;
    case BOLT_I64: case BOLT_F64: case BOLT_TIMESTAMP: return 8;
    case BOLT_DEC128: return 16;
  }
  return 0;
}
inline bool dtype_is_int(int32_t dt) { return dt == BOLT_I32 || dt == BOLT_I64 || dt == BOLT_DATE32; }
inline bool dtype_is_float(int32_t dt) { return dt == BOLT_F32 || dt == BOLT_F64; }
const char* dtype_name(int32_t dt);

// Occupancy-sized grid for a grid-stride kernel: enough workgroups to fill
// all 256 CUs several times over, capped by the work available.
inline unsigned grid_for(int64_t work_items, int block, int per_thread, int max_blocks = 256 * 16) {
  int64_t need = (work_items + (int64_t)block * per_thread - 1) / ((int64_t)block * per_thread);
  if (need < 1) need = 1;
  if (need > max_blocks) need = max_blocks;
  return (unsigned)need;
}

// Grid of a streaming kernel: grid_for's, lowered to BOLT_STREAM_GRID when
// that is set and smaller (a tuning knob; the kernels stride over the rows).
inline unsigned stream_grid(int64_t work_items, int block, int per_thread, int max_blocks) {
  unsigned g = grid_for(work_items, block, per_thread, max_blocks);
  if (const char* e = getenv("BOLT_STREAM_GRID")) {
    const long v = atol(e);
    if (v >= 1 && (unsigned long)v < g) g = (unsigned)v;
  }
  return g;
}

// ---- device-side helpers --------------------------------------------------

__device__ __forceinline__ bool bit_valid(const uint8_t* validity, int64_t i) {
  return validity == nullptr || ((validity[i >> 3] >> (i & 7)) & 1);
}

}  // namespace bolt
