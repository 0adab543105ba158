"""CPU: the Arrow format-string -> dtype mapping of the ingestion boundary
(bolt_arrow_format_dtype; GpuTable::upload_column_from_batch,
src/exec/gpu_table.rs:805-892), checked against pyarrow's own exported
format strings, and the no-device contract of the table entry points."""
import ctypes as C

import pytest

from craton_bolt_amd import _native as N
from craton_bolt_amd.device import device_count

pa = pytest.importorskip("pyarrow")
from craton_bolt_amd.arrow import ArrowArray, ArrowSchema, _Exported, format_dtype  # noqa: E402


def _format_of(t):
    with _Exported(pa.array([], t)) as e:
        return e.schema.format.decode()


@pytest.mark.parametrize("t,dtype,scale", [
    (pa.int32(), N.I32, 0), (pa.int64(), N.I64, 0), (pa.float32(), N.F32, 0), (pa.float64(), N.F64, 0),
    (pa.bool_(), N.BOOL, 0), (pa.date32(), N.DATE32, 0),
    (pa.timestamp("s"), N.TIMESTAMP, 0), (pa.timestamp("ms"), N.TIMESTAMP, 3),
    (pa.timestamp("us", tz="UTC"), N.TIMESTAMP, 6), (pa.timestamp("ns", tz="Europe/Paris"), N.TIMESTAMP, 9),
    (pa.decimal128(12, 2), N.DEC128, 2), (pa.decimal128(38, 0), N.DEC128, 0), (pa.decimal128(38, 38), N.DEC128, 38),
])
def test_supported_formats(t, dtype, scale):
    assert format_dtype(_format_of(t)) == (dtype, scale)


@pytest.mark.parametrize("t", [pa.utf8(), pa.large_utf8(), pa.int8(), pa.int16(), pa.uint32(), pa.uint64(),
                               pa.float16(), pa.date64(), pa.decimal256(40, 2), pa.list_(pa.int32()),
                               pa.struct([("a", pa.int32())]), pa.time64("us"), pa.binary()])
def test_unsupported_formats_are_type_errors(t):
    with pytest.raises(N.BoltTypeError, match="unsupported Arrow dtype"):
        format_dtype(_format_of(t))


@pytest.mark.parametrize("fmt", ["", "d:", "d:12", "d:12,", "d:12,-2", "d:39,2", "d:5,6", "d:12,2,256", "d:12,2,",
                                 "ts", "tsx:", "tdm", "x"])
def test_malformed_formats_are_type_errors(fmt):
    with pytest.raises(N.BoltTypeError):
        format_dtype(fmt)


@pytest.mark.skipif(device_count() > 0, reason="stub contract is only observable without a GPU")
def test_table_entry_points_without_a_device():
    L = N.lib()
    batch = pa.record_batch({"x": pa.array([1, 2], pa.int64())})
    h = C.c_void_p()
    with _Exported(batch) as e:
        assert L.bolt_gpu_table_from_arrow(None, C.addressof(e.schema), C.addressof(e.array), C.byref(h)) == \
            N.BOLT_ERR_NO_DEVICE
    assert h.value is None
    assert L.bolt_gpu_table_release(None) == 0
    assert isinstance(ArrowSchema(), C.Structure) and isinstance(ArrowArray(), C.Structure)
