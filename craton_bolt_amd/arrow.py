"""Arrow record batches (in memory, or decoded from Arrow IPC files and
streams) -> device-resident tables, and device columns -> Arrow arrays.

The C side is bolt_arrow.h (bolt_gpu_table_*): it consumes the Arrow C Data
Interface, so any Arrow producer can feed it; here pyarrow is that producer.
`read_ipc` mirrors the reference's ingestion of record batches into a
resident table (GpuTable::from_record_batch, src/exec/gpu_table.rs:898,
called per registered batch; Flight DoPut batches, src/flight/mod.rs:129):
the IPC decode stays on the host in pyarrow, each decoded batch crosses the
C boundary once and is copied straight into the table's device storage.
"""
from __future__ import annotations

import ctypes as C

from . import _native as N
from .device import DeviceBuffer, DeviceColumn, Stream


class ArrowSchema(C.Structure):
    pass


class ArrowArray(C.Structure):
    pass


_SCHEMA_RELEASE = C.CFUNCTYPE(None, C.POINTER(ArrowSchema))
_ARRAY_RELEASE = C.CFUNCTYPE(None, C.POINTER(ArrowArray))
ArrowSchema._fields_ = [("format", C.c_char_p), ("name", C.c_char_p), ("metadata", C.c_char_p),
                        ("flags", C.c_int64), ("n_children", C.c_int64),
                        ("children", C.POINTER(C.POINTER(ArrowSchema))), ("dictionary", C.POINTER(ArrowSchema)),
                        ("release", C.c_void_p), ("private_data", C.c_void_p)]
ArrowArray._fields_ = [("length", C.c_int64), ("null_count", C.c_int64), ("offset", C.c_int64),
                       ("n_buffers", C.c_int64), ("n_children", C.c_int64), ("buffers", C.POINTER(C.c_void_p)),
                       ("children", C.POINTER(C.POINTER(ArrowArray))), ("dictionary", C.POINTER(ArrowArray)),
                       ("release", C.c_void_p), ("private_data", C.c_void_p)]


def _release(struct, fn_type) -> None:
    if struct.release:
        fn_type(struct.release)(C.byref(struct))


class _Exported:
    """A pyarrow object exported through the C Data Interface; released on exit."""

    def __init__(self, obj):
        self.schema, self.array = ArrowSchema(), ArrowArray()
        obj._export_to_c(C.addressof(self.array), C.addressof(self.schema))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        _release(self.array, _ARRAY_RELEASE)
        _release(self.schema, _SCHEMA_RELEASE)


def format_dtype(fmt: str) -> tuple[int, int]:
    """Arrow format string -> (dtype, scale) under the library's mapping (bolt_arrow_format_dtype)."""
    dt, sc = C.c_int32(), C.c_int32()
    N.check(N.lib().bolt_arrow_format_dtype(fmt.encode(), C.byref(dt), C.byref(sc)), "arrow format")
    return dt.value, sc.value


class _TableBuffer(DeviceBuffer):
    """A view of device memory owned by a GpuTable (never freed here)."""

    def __init__(self, ptr: int, nbytes: int, owner: "GpuTable"):
        self.ptr, self.nbytes, self._owner = ptr, int(nbytes), owner

    def free(self) -> None:
        self.ptr = None

    def __del__(self):
        pass


class GpuTable:
    """A device-resident table built from Arrow record batches (bolt_gpu_table_t)."""

    def __init__(self, batch, stream: Stream | None = None):
        self._stream = stream
        h = C.c_void_p()
        with _Exported(batch) as e:
            N.check(N.lib().bolt_gpu_table_from_arrow(self._sh(), C.addressof(e.schema), C.addressof(e.array),
                                                      C.byref(h)), "gpu table from arrow")
        self.handle = h.value
        self._generation = 0

    def _sh(self):
        return self._stream.handle if self._stream else None

    def append(self, batch) -> None:
        with _Exported(batch) as e:
            N.check(N.lib().bolt_gpu_table_append_arrow(self.handle, self._sh(), C.addressof(e.schema),
                                                        C.addressof(e.array)), "gpu table append")
        self._generation += 1

    def reserve(self, rows: int) -> None:
        N.check(N.lib().bolt_gpu_table_reserve(self.handle, int(rows), self._sh()), "gpu table reserve")
        self._generation += 1

    @property
    def num_rows(self) -> int:
        n = C.c_int64()
        N.check(N.lib().bolt_gpu_table_num_rows(self.handle, C.byref(n)), "num rows")
        return n.value

    @property
    def num_columns(self) -> int:
        n = C.c_int32()
        N.check(N.lib().bolt_gpu_table_num_columns(self.handle, C.byref(n)), "num columns")
        return n.value

    @property
    def names(self) -> list[str]:
        return [self.field(i)[0] for i in range(self.num_columns)]

    def field(self, i: int) -> tuple[str, str]:
        """(name, time zone) of column i."""
        name, tz = C.c_char_p(), C.c_char_p()
        N.check(N.lib().bolt_gpu_table_column_name(self.handle, i, C.byref(name), C.byref(tz)), "column name")
        return name.value.decode(), tz.value.decode()

    def column(self, key: int | str) -> DeviceColumn:
        """Column `key` (index or name) as an operator input; valid until the next append / reserve."""
        i = self.names.index(key) if isinstance(key, str) else int(key)
        c = N.BoltColumn()
        N.check(N.lib().bolt_gpu_table_column(self.handle, i, C.byref(c)), "column")
        width = N.DTYPE_SIZE[c.dtype]
        vals = _TableBuffer(c.data, max(c.length, 1) * width, self)
        valid = _TableBuffer(c.validity, (c.length + 7) // 8, self) if c.validity else None
        return DeviceColumn(vals, c.length, c.dtype, valid, c.scale)

    def release(self) -> None:
        if getattr(self, "handle", None):
            N.lib().bolt_gpu_table_release(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def table_from_arrow(data, stream: Stream | None = None) -> GpuTable:
    """A pyarrow RecordBatch, Table or iterable of batches -> one GpuTable."""
    import pyarrow as pa

    if isinstance(data, pa.RecordBatch):
        return GpuTable(data, stream)
    batches = data.to_batches() if isinstance(data, pa.Table) else list(data)
    if not batches:
        raise ValueError("no record batches")
    t = GpuTable(batches[0], stream)
    total = sum(b.num_rows for b in batches)
    if total > t.num_rows:
        t.reserve(total)
    for b in batches[1:]:
        t.append(b)
    return t


def read_ipc(source, stream: Stream | None = None) -> GpuTable:
    """An Arrow IPC file or stream (path, buffer or file object) -> one GpuTable, batch by batch (a path
    is memory-mapped, so batches are not copied on the host).  An IPC file's batches are all known from
    its footer, so the table is sized once; a stream's grow the table as they arrive."""
    import pyarrow as pa
    import pyarrow.ipc as ipc

    if isinstance(source, str) or hasattr(source, "__fspath__"):
        source = pa.memory_map(str(source))
    elif isinstance(source, (bytes, bytearray, memoryview)):
        source = pa.py_buffer(source)
    try:
        reader = ipc.open_file(source)
    except pa.ArrowInvalid:
        reader = None
    if reader is not None:  # the footer lists every batch: size the table once
        batches = [reader.get_batch(i) for i in range(reader.num_record_batches)]
        total = sum(b.num_rows for b in batches)
    else:
        if hasattr(source, "seek"):
            source.seek(0)
        batches = iter(ipc.open_stream(source))
        total = None
    t = None
    for b in batches:
        if t is None:
            t = GpuTable(b, stream)
            if total is not None and total > t.num_rows:
                t.reserve(total)
        else:
            t.append(b)
    if t is None:
        raise ValueError("the IPC source holds no record batches")
    return t


def column_to_arrow(col: DeviceColumn, name: str = "", stream: Stream | None = None):
    """A device column -> a pyarrow Array (copied to host; bolt_column_to_arrow)."""
    import pyarrow as pa

    schema, array = ArrowSchema(), ArrowArray()
    c = col.c()
    N.check(N.lib().bolt_column_to_arrow(stream.handle if stream else None, C.byref(c), name.encode(),
                                         C.addressof(schema), C.addressof(array)), "column to arrow")
    # pyarrow takes ownership of both structs (it calls their release callbacks)
    return pa.Array._import_from_c(C.addressof(array), C.addressof(schema))


def table_to_arrow(table: GpuTable):
    """A GpuTable -> a pyarrow Table (host copy of every column)."""
    import pyarrow as pa

    arrays = [column_to_arrow(table.column(i)) for i in range(table.num_columns)]
    return pa.table(arrays, names=table.names)
