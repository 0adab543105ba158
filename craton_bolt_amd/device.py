"""Device memory, streams and events over the bolt_hip_* runtime ABI.

`DeviceBuffer` is the owning handle (the reference's GpuBuffer / GpuVec,
src/cuda/buffer.rs:154, src/cuda/smart_ptrs.rs:36); `DeviceColumn` pairs a
values buffer with an optional validity bitmap and a dtype, and converts to
the C `bolt_column_t` descriptor.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N

_NP_DTYPE = {
    N.BOOL: np.uint8,
    N.I32: np.int32,
    N.I64: np.int64,
    N.F32: np.float32,
    N.F64: np.float64,
    N.DATE32: np.int32,
    N.TIMESTAMP: np.int64,
}


def dtype_of_numpy(a: np.ndarray) -> int:
    m = {np.dtype(np.int32): N.I32, np.dtype(np.int64): N.I64, np.dtype(np.float32): N.F32,
         np.dtype(np.float64): N.F64, np.dtype(np.uint8): N.BOOL, np.dtype(np.bool_): N.BOOL}
    if a.dtype not in m:
        raise TypeError(f"unsupported numpy dtype {a.dtype}")
    return m[a.dtype]


class Stream:
    """A HIP stream (default priority). `Stream.null()` is the null stream."""

    def __init__(self, handle: int | None = None, owned: bool = True):
        if handle is None:
            h = C.c_void_p()
            N.check_rt(N.lib().bolt_hip_stream_create(C.byref(h), 0), "stream create")
            handle = h.value
        self.handle = handle
        self._owned = owned and handle is not None

    @classmethod
    def null(cls) -> "Stream":
        s = cls.__new__(cls)
        s.handle = None
        s._owned = False
        return s

    def synchronize(self) -> None:
        N.check_rt(N.lib().bolt_hip_stream_synchronize(self.handle), "stream sync")

    def __del__(self):
        if getattr(self, "_owned", False) and self.handle:
            try:
                N.lib().bolt_hip_stream_destroy(self.handle)
            except Exception:
                pass
            self.handle = None


class Event:
    def __init__(self, timing: bool = True):
        h = C.c_void_p()
        N.check_rt(N.lib().bolt_hip_event_create(C.byref(h), 0 if timing else 2), "event create")
        self.handle = h.value

    def record(self, stream: Stream | None = None) -> None:
        N.check_rt(N.lib().bolt_hip_event_record(self.handle, stream.handle if stream else None), "event record")

    def synchronize(self) -> None:
        N.check_rt(N.lib().bolt_hip_event_synchronize(self.handle), "event sync")

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        N.check_rt(N.lib().bolt_hip_event_elapsed_ms(C.byref(ms), self.handle, end.handle), "elapsed")
        return float(ms.value)

    def __del__(self):
        if getattr(self, "handle", None):
            try:
                N.lib().bolt_hip_event_destroy(self.handle)
            except Exception:
                pass
            self.handle = None


class DeviceBuffer:
    """Owning device allocation from the library's pool."""

    def __init__(self, nbytes: int):
        p = C.c_void_p()
        N.check_rt(N.lib().bolt_hip_malloc(C.byref(p), max(int(nbytes), 1)), f"malloc {nbytes}")
        self.ptr = p.value
        self.nbytes = int(nbytes)

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            N.check_rt(N.lib().bolt_hip_memcpy_htod(b.ptr, a.ctypes.data, a.nbytes), "htod")
        return b

    def copy_from(self, a: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(a)
        assert offset + a.nbytes <= self.nbytes
        if a.nbytes:
            N.check_rt(N.lib().bolt_hip_memcpy_htod(self.ptr + offset, a.ctypes.data, a.nbytes), "htod")

    def to_numpy(self, dtype, count: int | None = None, offset: int = 0) -> np.ndarray:
        itemsize = np.dtype(dtype).itemsize
        if count is None:
            count = (self.nbytes - offset) // itemsize
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            N.check_rt(N.lib().bolt_hip_memcpy_dtoh(out.ctypes.data, self.ptr + offset, out.nbytes), "dtoh")
        return out

    def fill(self, value: int = 0) -> None:
        N.check_rt(N.lib().bolt_hip_memset_d8(self.ptr, value, self.nbytes), "memset")

    def free(self) -> None:
        if self.ptr:
            N.lib().bolt_hip_free(self.ptr)
            self.ptr = None

    def __del__(self):
        if getattr(self, "ptr", None):
            try:
                N.lib().bolt_hip_free(self.ptr)
            except Exception:
                pass
            self.ptr = None


def pack_validity(valid: np.ndarray) -> np.ndarray:
    """Arrow LSB-first validity bitmap from a boolean array."""
    return np.packbits(np.asarray(valid, dtype=bool), bitorder="little")


class DeviceColumn:
    """A device-resident Arrow-layout column."""

    def __init__(self, values: DeviceBuffer, length: int, dtype: int, validity: DeviceBuffer | None = None,
                 scale: int = 0):
        self.values = values
        self.validity = validity
        self.length = int(length)
        self.dtype = int(dtype)
        self.scale = int(scale)

    @classmethod
    def from_numpy(cls, a: np.ndarray, dtype: int | None = None, valid: np.ndarray | None = None,
                   scale: int = 0) -> "DeviceColumn":
        """Upload `a`; `valid` (bool per row) adds a validity bitmap.
        Decimal128 columns are passed as an (n, 2) int64 array [lo, hi]."""
        if dtype is None:
            dtype = dtype_of_numpy(a)
        if dtype == N.DEC128:
            a = np.ascontiguousarray(a, dtype=np.int64).reshape(-1, 2)
            length = a.shape[0]
        else:
            a = np.ascontiguousarray(a, dtype=_NP_DTYPE[dtype])
            length = a.shape[0]
        vb = DeviceBuffer.from_numpy(pack_validity(valid)) if valid is not None else None
        return cls(DeviceBuffer.from_numpy(a), length, dtype, vb, scale)

    @classmethod
    def empty(cls, length: int, dtype: int, with_validity: bool = False, scale: int = 0) -> "DeviceColumn":
        vb = None
        if with_validity:
            vb = DeviceBuffer((length + 7) // 8 + 8)
            vb.fill(0)
        return cls(DeviceBuffer(max(length, 1) * N.DTYPE_SIZE[dtype]), length, dtype, vb, scale)

    def c(self, length: int | None = None) -> N.BoltColumn:
        return N.BoltColumn(self.values.ptr, self.validity.ptr if self.validity else None,
                            self.length if length is None else length, self.dtype, self.scale)

    def to_numpy(self, count: int | None = None) -> np.ndarray:
        n = self.length if count is None else count
        if self.dtype == N.DEC128:
            return self.values.to_numpy(np.int64, 2 * n).reshape(n, 2)
        return self.values.to_numpy(_NP_DTYPE[self.dtype], n)

    def valid_numpy(self, count: int | None = None) -> np.ndarray | None:
        if self.validity is None:
            return None
        n = self.length if count is None else count
        bits = self.validity.to_numpy(np.uint8, (n + 7) // 8)
        return np.unpackbits(bits, bitorder="little")[:n].astype(bool)


def columns_array(cols: list[DeviceColumn]):
    arr = (N.BoltColumn * max(len(cols), 1))()
    for i, c in enumerate(cols):
        arr[i] = c.c()
    return arr


def device_count() -> int:
    n = C.c_int()
    st = N.lib().bolt_hip_device_get_count(C.byref(n))
    if st == N.BOLT_ERR_NO_DEVICE:
        return 0
    N.check_rt(st, "device count")
    return n.value
