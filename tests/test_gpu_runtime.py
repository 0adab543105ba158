"""GPU: the device-runtime C ABI (bolt_hip.h) driven the way the reference's
FFI layer drives the CUDA driver (src/cuda/cuda_sys.rs:89-224): contexts,
device handles, a code object loaded with the options form, a launch with
the KernelArgs convention (src/exec/launch.rs:162-260), and a captured graph
instantiated with cuGraphInstantiate_v2's argument list and replayed."""
import ctypes as C

import numpy as np
import pytest

from craton_bolt_amd import _native as N

pytestmark = pytest.mark.gpu

KERNEL = N.PKG.parent / "tests" / "kernels" / "abi_fill.hsaco"


def ok(rc, what):
    assert rc == 0, f"{what}: status {rc}: {N.lib().bolt_last_error()}"


def test_contexts_and_device_handles(gpu):
    L = N.lib()
    cur = C.c_void_p(1)
    ok(L.bolt_hip_ctx_get_current(C.byref(cur)), "get_current")
    assert cur.value is None  # no context made current on this thread yet
    dev = C.c_int(-1)
    ok(L.bolt_hip_device_get(C.byref(dev), 0), "device_get")
    assert dev.value == 0
    bad = C.c_int(-1)
    assert L.bolt_hip_device_get(C.byref(bad), gpu) != 0  # ordinal past the last device
    ctx = C.c_void_p()
    ok(L.bolt_hip_ctx_create(C.byref(ctx), 0, dev.value), "ctx_create")
    assert ctx.value
    ok(L.bolt_hip_ctx_get_current(C.byref(cur)), "get_current")
    assert cur.value == ctx.value
    d = C.c_int(-1)
    ok(L.bolt_hip_get_device(C.byref(d)), "get_device")
    assert d.value == 0
    ok(L.bolt_hip_ctx_set_current(None), "set_current(NULL)")
    ok(L.bolt_hip_ctx_get_current(C.byref(cur)), "get_current")
    assert cur.value is None
    ok(L.bolt_hip_ctx_set_current(ctx), "set_current")
    ok(L.bolt_hip_ctx_destroy(ctx), "ctx_destroy")
    ok(L.bolt_hip_ctx_get_current(C.byref(cur)), "get_current")
    assert cur.value is None


def test_module_launch_and_graph_replay(gpu):
    L = N.lib()
    image = KERNEL.read_bytes()
    buf = C.create_string_buffer(image, len(image))
    mod, fn = C.c_void_p(), C.c_void_p()
    opts = (C.c_int * 2)(0, 0)
    vals = (C.c_void_p * 2)(None, None)
    ok(L.bolt_hip_module_load_data_ex(C.byref(mod), buf, 2, opts, vals), "module_load_data_ex")
    ok(L.bolt_hip_module_get_function(C.byref(fn), mod, b"abi_fill"), "module_get_function")
    n = 100_003
    out = C.c_void_p()
    ok(L.bolt_hip_malloc(C.byref(out), n * 8), "malloc")
    s = C.c_void_p()
    ok(L.bolt_hip_stream_create(C.byref(s), 0), "stream_create")
    a_out, a_n, a_v = C.c_uint64(out.value), C.c_uint32(n), C.c_uint32(3)
    params = (C.c_void_p * 3)(C.addressof(a_out), C.addressof(a_n), C.addressof(a_v))
    grid = (n + 255) // 256
    ok(L.bolt_hip_launch_kernel(fn, grid, 1, 1, 256, 1, 1, 0, s, params, None), "launch_kernel")
    ok(L.bolt_hip_stream_synchronize(s), "stream_synchronize")
    host = np.zeros(n, np.uint64)
    ok(L.bolt_hip_memcpy_dtoh(host.ctypes.data, out, n * 8), "memcpy_dtoh")
    want3 = np.arange(n, dtype=np.uint64) * np.uint64(3)
    assert np.array_equal(host, want3)
    # the same launch with value 7, captured and replayed three times, with
    # eager work (an overwrite back to 3 * i) between the replays
    a_v.value = 7
    ok(L.bolt_hip_stream_begin_capture(s, 0), "begin_capture")
    ok(L.bolt_hip_launch_kernel(fn, grid, 1, 1, 256, 1, 1, 0, s, params, None), "captured launch")
    g = C.c_void_p()
    ok(L.bolt_hip_stream_end_capture(s, C.byref(g)), "end_capture")
    ex, err_node = C.c_void_p(), C.c_void_p(1)
    log = C.create_string_buffer(256)
    ok(L.bolt_hip_graph_instantiate(C.byref(ex), g, C.byref(err_node), log, 256), "graph_instantiate")
    assert err_node.value is None and log.value == b""
    a_v.value = 3
    want7 = np.arange(n, dtype=np.uint64) * np.uint64(7)
    for rep in range(3):
        ok(L.bolt_hip_launch_kernel(fn, grid, 1, 1, 256, 1, 1, 0, s, params, None), "eager launch")
        ok(L.bolt_hip_stream_synchronize(s), "sync")
        ok(L.bolt_hip_memcpy_dtoh(host.ctypes.data, out, n * 8), "memcpy_dtoh")
        assert np.array_equal(host, want3), f"eager before replay {rep}"
        ok(L.bolt_hip_graph_launch(ex, s), "graph_launch")
        ok(L.bolt_hip_stream_synchronize(s), "sync")
        ok(L.bolt_hip_memcpy_dtoh(host.ctypes.data, out, n * 8), "memcpy_dtoh")
        assert np.array_equal(host, want7), f"replay {rep}"
    ok(L.bolt_hip_graph_exec_destroy(ex), "graph_exec_destroy")
    ok(L.bolt_hip_graph_destroy(g), "graph_destroy")
    ok(L.bolt_hip_stream_destroy(s), "stream_destroy")
    ok(L.bolt_hip_free(out), "free")
    ok(L.bolt_hip_module_unload(mod), "module_unload")
