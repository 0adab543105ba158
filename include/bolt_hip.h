/*
 * bolt_hip.h — device-runtime C ABI of the MI355X execution core.
 *
 * One entry point per CUDA-driver call that the reference binds in its raw
 * FFI block (src/cuda/cuda_sys.rs:89-224).  A Rust `cuda_sys.rs` can be
 * re-pointed at these symbols one for one: every function returns an int
 * status (0 = success), handles are opaque pointers, sizes are size_t.
 *
 * Status codes keep the reference's conventions:
 *   - 0 is success (CUDA_SUCCESS, src/cuda/cuda_sys.rs:60);
 *   - BOLT_HIP_ERROR_NOT_READY (600) is the "still in flight" answer of
 *     bolt_hip_event_query (CUDA_ERROR_NOT_READY, src/cuda/cuda_sys.rs:66);
 *   - BOLT_HIP_ERROR_NO_DEVICE (100000) is returned by every call when no
 *     GPU is usable, matching the stub sentinel CUDA_ERROR_STUB
 *     (src/cuda/cuda_sys.rs:77) that tests/cuda_stub_contract_test.rs:40-60
 *     maps to a typed "no GPU" error;
 *   - any other non-zero value is the underlying hipError_t; the text of the
 *     last failure on the calling thread is returned by bolt_last_error().
 */
#ifndef BOLT_HIP_H
#define BOLT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BOLT_HIP_SUCCESS 0
#define BOLT_HIP_ERROR_NOT_READY 600
#define BOLT_HIP_ERROR_NO_DEVICE 100000

/* Opaque handles (hipStream_t, hipEvent_t, hipModule_t, hipFunction_t,
 * hipGraph_t, hipGraphExec_t underneath). The null stream is NULL. */
typedef struct bolt_stream_s* bolt_stream_t;
typedef struct bolt_event_s* bolt_event_t;
typedef struct bolt_module_s* bolt_module_t;
typedef struct bolt_function_s* bolt_function_t;
typedef struct bolt_graph_s* bolt_graph_t;
typedef struct bolt_graph_exec_s* bolt_graph_exec_t;

/* Context handle (CUcontext underneath in the reference).  HIP has one
 * primary context per device; a bolt_ctx_t names a device's primary context.
 * NULL means "no context". */
typedef struct bolt_ctx_s* bolt_ctx_t;

/* Text of the last error raised on the calling thread ("" if none). */
const char* bolt_last_error(void);

/* cuGetErrorName / cuGetErrorString (cuda_sys.rs:151-152): static text for a
 * status code of this ABI (hipError_t values, 600, 100000).  Unknown codes
 * set *str to NULL and return hipErrorInvalidValue (1), as the driver does. */
int bolt_hip_get_error_name(int error, const char** str);
int bolt_hip_get_error_string(int error, const char** str);

/* cuInit (cuda_sys.rs:90). flags must be 0. Safe to call repeatedly. */
int bolt_hip_init(unsigned int flags);
/* cuDeviceGetCount (cuda_sys.rs:91). */
int bolt_hip_device_get_count(int* count);
/* cuDeviceGet (cuda_sys.rs:92): the device handle of ordinal `ordinal`
 * (device handles are ordinals, as CUdevice is). */
int bolt_hip_device_get(int* device, int ordinal);
/* cuDeviceGetName (cuda_sys.rs:93). */
int bolt_hip_device_get_name(char* name, int len, int device);
/* cuDeviceGetAttribute (cuda_sys.rs:94); attrib is a hipDeviceAttribute_t. */
int bolt_hip_device_get_attribute(int* value, int attrib, int device);
/* cuDeviceTotalMem_v2 (cuda_sys.rs:99). */
int bolt_hip_device_total_mem(size_t* bytes, int device);
/* cuCtxCreate_v2 / cuCtxDestroy_v2 / cuCtxSetCurrent / cuCtxGetCurrent
 * (cuda_sys.rs:100-102,116).  create retains `device`'s primary context,
 * makes it current on the calling thread and returns its handle (flags are
 * accepted and ignored: scheduling flags have no HIP counterpart); destroy
 * releases that retain (the primary context itself lives until the process
 * ends, so destroying a context does not free memory allocated in it);
 * get_current writes the context made current on this thread by create or
 * set_current, or NULL if none was. */
int bolt_hip_ctx_create(bolt_ctx_t* ctx, unsigned int flags, int device);
int bolt_hip_ctx_destroy(bolt_ctx_t ctx);
int bolt_hip_ctx_set_current(bolt_ctx_t ctx);
int bolt_hip_ctx_get_current(bolt_ctx_t* ctx);
/* Device selection by ordinal, and cuCtxGetDevice (cuda_sys.rs:106): the
 * device of the calling thread's current context. */
int bolt_hip_set_device(int device);
int bolt_hip_get_device(int* device);
/* cuCtxSynchronize (cuda_sys.rs:170). */
int bolt_hip_device_synchronize(void);

/* cuMemAlloc_v2 / cuMemFree_v2 (cuda_sys.rs:117-118). Allocations come from
 * the stream-ordered pool of the current device and are 256-byte aligned
 * (stricter than ARROW_ALIGNMENT = 64, src/cuda/buffer.rs:30). */
int bolt_hip_malloc(void** dptr, size_t bytes);
int bolt_hip_free(void* dptr);
/* Stream-ordered variants (no reference counterpart: the reference pool
 * src/cuda/mem_pool.rs:1490,1648 is synchronous). */
int bolt_hip_malloc_async(void** dptr, size_t bytes, bolt_stream_t stream);
int bolt_hip_free_async(void* dptr, bolt_stream_t stream);
/* cuMemAllocHost_v2 / cuMemHostAlloc / cuMemFreeHost (cuda_sys.rs:119-127). */
int bolt_hip_host_alloc(void** ptr, size_t bytes, unsigned int flags);
int bolt_hip_host_free(void* ptr);
/* cuMemGetInfo_v2 (cuda_sys.rs:199). */
int bolt_hip_mem_get_info(size_t* free_bytes, size_t* total_bytes);

/* cuMemcpy{HtoD,DtoH,DtoD}_v2 and the Async forms (cuda_sys.rs:128-147).
 * bolt_hip_memcpy_htod is synchronous and ordered after the device's earlier
 * work; from pageable host memory (256 KB and up) it stages the copy through
 * pinned buffers with up to four host threads, so the host-side copies and
 * the DMA overlap (pinned sources are copied directly). */
int bolt_hip_memcpy_htod(void* dst, const void* src, size_t bytes);
int bolt_hip_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int bolt_hip_memcpy_dtod(void* dst, const void* src, size_t bytes);
int bolt_hip_memcpy_htod_async(void* dst, const void* src, size_t bytes, bolt_stream_t stream);
int bolt_hip_memcpy_dtoh_async(void* dst, const void* src, size_t bytes, bolt_stream_t stream);
int bolt_hip_memcpy_dtod_async(void* dst, const void* src, size_t bytes, bolt_stream_t stream);
/* cuMemsetD8_v2 / cuMemsetD8Async (cuda_sys.rs:148-149).  The async form is
 * a fill kernel, not hipMemsetAsync, so it is safe inside graph capture. */
int bolt_hip_memset_d8(void* dst, unsigned char value, size_t count);
int bolt_hip_memset_d8_async(void* dst, unsigned char value, size_t count, bolt_stream_t stream);

/* cuModuleLoadData / cuModuleUnload / cuModuleGetFunction (cuda_sys.rs:153,161-166).
 * `image` is a gfx950 code object (not PTX). */
int bolt_hip_module_load_data(bolt_module_t* module, const void* image);
/* cuModuleLoadDataEx (cuda_sys.rs:154-160): the same load; the JIT options
 * are accepted and ignored, since a gfx950 code object is already final
 * (options may be NULL when num_options is 0). */
int bolt_hip_module_load_data_ex(bolt_module_t* module, const void* image, unsigned int num_options,
                                 int* options, void** option_values);
int bolt_hip_module_unload(bolt_module_t module);
int bolt_hip_module_get_function(bolt_function_t* func, bolt_module_t module, const char* name);
/* cuLaunchKernel (cuda_sys.rs:180).  kernel_params follows the reference
 * KernelArgs convention (src/exec/launch.rs:162): an array of pointers to the
 * argument values in declaration order. Does not synchronise. */
int bolt_hip_launch_kernel(bolt_function_t func,
                           unsigned int grid_x, unsigned int grid_y, unsigned int grid_z,
                           unsigned int block_x, unsigned int block_y, unsigned int block_z,
                           unsigned int shared_mem_bytes, bolt_stream_t stream,
                           void** kernel_params, void** extra);

/* cuStreamCreate / cuStreamDestroy_v2 / cuStreamSynchronize (cuda_sys.rs:167-169).
 * Streams are always created at default priority. */
int bolt_hip_stream_create(bolt_stream_t* stream, unsigned int flags);
int bolt_hip_stream_destroy(bolt_stream_t stream);
int bolt_hip_stream_synchronize(bolt_stream_t stream);
/* Stream pool (src/cuda/stream_pool.rs:136,170): reusable streams that are
 * only destroyed at process exit. */
int bolt_hip_stream_pool_acquire(bolt_stream_t* stream);
int bolt_hip_stream_pool_release(bolt_stream_t stream);

/* cuEvent{Create,Record,Query,Synchronize,Destroy} (cuda_sys.rs:175-179)
 * plus elapsed time, used by the benchmark to time kernels on the stream they
 * run on. flags: 0 = timing enabled, 2 = timing disabled (CU_EVENT_DISABLE_TIMING). */
int bolt_hip_event_create(bolt_event_t* event, unsigned int flags);
int bolt_hip_event_record(bolt_event_t event, bolt_stream_t stream);
int bolt_hip_event_query(bolt_event_t event);
int bolt_hip_event_synchronize(bolt_event_t event);
int bolt_hip_event_elapsed_ms(float* ms, bolt_event_t start, bolt_event_t end);
int bolt_hip_event_destroy(bolt_event_t event);

/* cuStreamBeginCapture_v2 / cuStreamEndCapture / cuGraphInstantiate_v2 /
 * cuGraphLaunch / cuGraphExecDestroy / cuGraphDestroy (cuda_sys.rs:213-224).
 * mode: 0 global, 1 thread-local, 2 relaxed (hipStreamCaptureMode). */
int bolt_hip_stream_begin_capture(bolt_stream_t stream, unsigned int mode);
int bolt_hip_stream_end_capture(bolt_stream_t stream, bolt_graph_t* graph);
/* graph_instantiate keeps cuGraphInstantiate_v2's arguments: on failure the
 * failing node goes to *error_node and the diagnostic text (truncated to
 * buffer_size - 1 bytes, NUL-terminated) to log_buffer; both may be NULL. */
int bolt_hip_graph_instantiate(bolt_graph_exec_t* exec, bolt_graph_t graph, void** error_node, char* log_buffer,
                               size_t buffer_size);
int bolt_hip_graph_launch(bolt_graph_exec_t exec, bolt_stream_t stream);
int bolt_hip_graph_exec_destroy(bolt_graph_exec_t exec);
int bolt_hip_graph_destroy(bolt_graph_t graph);

/* Pool statistics (src/cuda/mem_pool.rs:809,840): bytes currently handed out
 * by bolt_hip_malloc*, and the high-water mark. */
int bolt_hip_pool_stats(size_t* in_use_bytes, size_t* peak_bytes);

#ifdef __cplusplus
}
#endif
#endif /* BOLT_HIP_H */
