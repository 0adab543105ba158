"""GPU: bolt_hip_memcpy_htod, whose pageable copies of 256 KB and up are staged
through pinned buffers by up to four host threads.  Round trips are
byte-exact at sizes around every boundary of the staging (its threshold, the
4 MB chunk, the 2 MB per lane, odd byte counts), from pinned memory, and
after device work that the copy must follow."""
import ctypes as C

import numpy as np
import pytest

from craton_bolt_amd import _native as N
from craton_bolt_amd.device import DeviceBuffer

pytestmark = pytest.mark.gpu

KB, MB = 1 << 10, 1 << 20
SIZES = [1, 4095, 256 * KB - 1, 256 * KB, 256 * KB + 1, 2 * MB - 7, 2 * MB, 4 * MB - 1, 4 * MB, 4 * MB + 1,
         6 * MB + 3, 8 * MB + 5, 33 * MB + 17, 257 * MB + 9]


def _roundtrip(host: np.ndarray) -> np.ndarray:
    d = DeviceBuffer(host.nbytes)
    N.check_rt(N.lib().bolt_hip_memcpy_htod(d.ptr, host.ctypes.data, host.nbytes), "htod")
    return d.to_numpy(np.uint8, host.nbytes)


@pytest.mark.parametrize("nbytes", SIZES)
def test_pageable_upload_roundtrip(gpu, nbytes):
    print(f"[case] upload {nbytes} bytes", flush=True)
    rng = np.random.default_rng(nbytes)
    buf = rng.integers(0, 256, nbytes + 3, dtype=np.uint8)
    host = buf[3:]  # an unaligned start
    assert np.array_equal(_roundtrip(host), host)


def test_pinned_upload_roundtrip(gpu):
    n = 12 * MB + 1
    p = C.c_void_p()
    N.check_rt(N.lib().bolt_hip_host_alloc(C.byref(p), n, 0), "host alloc")
    try:
        host = np.ctypeslib.as_array((C.c_uint8 * n).from_address(p.value))
        host[:] = np.random.default_rng(1).integers(0, 256, n, dtype=np.uint8)
        assert np.array_equal(_roundtrip(host), host)
    finally:
        N.lib().bolt_hip_host_free(p)


def test_upload_follows_device_work(gpu):
    """A synchronous upload is ordered after earlier work on the device: a
    device-to-device copy into the same buffer, enqueued first, must not
    overwrite it afterwards."""
    n = 64 * MB
    old = np.full(n, 7, np.uint8)
    src = DeviceBuffer.from_numpy(old)
    dst = DeviceBuffer(n)
    for _ in range(3):
        N.check_rt(N.lib().bolt_hip_memcpy_dtod_async(dst.ptr, src.ptr, n, None), "dtod")
        new = np.random.default_rng(2).integers(0, 256, n, dtype=np.uint8)
        N.check_rt(N.lib().bolt_hip_memcpy_htod(dst.ptr, new.ctypes.data, n), "htod")
        assert np.array_equal(dst.to_numpy(np.uint8, n), new)
