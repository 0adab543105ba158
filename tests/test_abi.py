"""The C-ABI library loads, exports every symbol its headers declare, and
honours the no-device contract (tests/cuda_stub_contract_test.rs:1-90: every
GPU entry fails with a typed no-GPU error instead of crashing)."""
import ctypes as C

import pytest

from craton_bolt_amd import _native as N
from craton_bolt_amd.device import device_count


def test_library_exports_every_header_symbol():
    L = N.lib()
    names = N.header_symbols()
    assert len(names) > 50
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_library_is_gfx950_only():
    blob = N.LIB_PATH.read_bytes()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob
    for other in (b"--gfx90a", b"--gfx942", b"--gfx1100"):
        assert other not in blob


@pytest.mark.skipif(device_count() > 0, reason="stub contract is only observable without a GPU")
def test_no_device_contract():
    L = N.lib()
    n = C.c_int(-1)
    assert L.bolt_hip_device_get_count(C.byref(n)) == N.BOLT_ERR_NO_DEVICE
    assert n.value == 0
    p = C.c_void_p()
    assert L.bolt_hip_malloc(C.byref(p), 1024) == N.BOLT_ERR_NO_DEVICE
    assert b"no HIP device" in L.bolt_last_error()
    col = N.BoltColumn(None, None, 10, N.I64, 0)
    res = (N.BoltScalar * 1)()
    aggs = (N.BoltAgg * 1)(N.BoltAgg(N.SUM, 0))
    assert L.bolt_aggregate(None, C.byref(col), 1, aggs, 1, None, res) == N.BOLT_ERR_NO_DEVICE
    t = C.c_void_p()
    assert L.bolt_join_build(None, C.byref(col), None, C.byref(t)) == N.BOLT_ERR_NO_DEVICE
    assert L.bolt_groupby(None, C.byref(col), 1, C.byref(col), 1, aggs, 1, None, 0, C.byref(t)) == \
        N.BOLT_ERR_NO_DEVICE
    with pytest.raises(N.NoDeviceError):
        N.check(L.bolt_aggregate(None, C.byref(col), 1, aggs, 1, None, res))


def test_error_names_and_strings():
    """cuGetErrorName / cuGetErrorString (src/cuda/cuda_sys.rs:151-152):
    static text for every status of this ABI, and an error for unknown codes."""
    L = N.lib()
    for code, name in [(0, b"hipSuccess"), (1, b"hipErrorInvalidValue"), (600, b"hipErrorNotReady"),
                       (100000, b"BOLT_HIP_ERROR_NO_DEVICE")]:
        p, q = C.c_char_p(), C.c_char_p()
        assert L.bolt_hip_get_error_name(code, C.byref(p)) == 0
        assert p.value == name
        assert L.bolt_hip_get_error_string(code, C.byref(q)) == 0
        assert q.value
    p = C.c_char_p(b"x")
    assert L.bolt_hip_get_error_name(123456, C.byref(p)) == 1 and p.value is None
    assert L.bolt_hip_get_error_string(123456, C.byref(p)) == 1 and p.value is None


@pytest.mark.skipif(device_count() > 0, reason="stub contract is only observable without a GPU")
def test_no_device_contract_for_contexts_and_modules():
    L = N.lib()
    d, ctx, mod = C.c_int(-1), C.c_void_p(), C.c_void_p()
    assert L.bolt_hip_device_get(C.byref(d), 0) == N.BOLT_ERR_NO_DEVICE
    assert L.bolt_hip_ctx_create(C.byref(ctx), 0, 0) == N.BOLT_ERR_NO_DEVICE
    assert ctx.value is None
    assert L.bolt_hip_ctx_get_current(C.byref(ctx)) == N.BOLT_ERR_NO_DEVICE
    assert L.bolt_hip_module_load_data_ex(C.byref(mod), b"\0" * 16, 0, None, None) == N.BOLT_ERR_NO_DEVICE
    ex = C.c_void_p()
    assert L.bolt_hip_graph_instantiate(C.byref(ex), None, None, None, 0) == N.BOLT_ERR_NO_DEVICE
