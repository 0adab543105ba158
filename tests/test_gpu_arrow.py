"""GPU: Arrow record batches -> device-resident tables (bolt_gpu_table_*,
GpuTable::from_record_batch, src/exec/gpu_table.rs:898) and device columns
-> Arrow arrays.  Byte-exact: every column's device values equal the Arrow
values buffer over the batch's rows (Bool unpacked to bytes), validity equals
Arrow's, across sliced batches (non-zero offsets), appends that land at
unaligned rows, NULLs that first appear in a later batch, and IPC files and
streams; operators over an ingested table equal the oracle on the same data."""
import io

import numpy as np
import pytest

pa = pytest.importorskip("pyarrow")
import pyarrow.ipc as ipc  # noqa: E402

import oracle as O  # noqa: E402
from craton_bolt_amd import _native as N  # noqa: E402
from craton_bolt_amd.arrow import GpuTable, column_to_arrow, read_ipc, table_from_arrow, table_to_arrow  # noqa: E402
from craton_bolt_amd.ops import execute_groupby  # noqa: E402

pytestmark = pytest.mark.gpu

WANT_DTYPE = {"i32": N.I32, "i64": N.I64, "f32": N.F32, "f64": N.F64, "b": N.BOOL, "d": N.DATE32,
              "ts_us": N.TIMESTAMP, "ts_ns": N.TIMESTAMP, "price": N.DEC128}


def _table(n, seed, nulls=True):
    """Every supported Arrow dtype; about 10% NULLs in five columns when `nulls`."""
    rng = np.random.default_rng(seed)

    def mask():
        return rng.random(n) < 0.1 if nulls else None

    # Decimal128(12, 2) cents in [100.00, 1_000_000.00): 16-byte little-endian two's complement
    cents = rng.integers(10_000, 100_000_000, n).astype(np.int64)
    dec_buf = pa.py_buffer(np.stack([cents, np.zeros(n, np.int64)], axis=1).tobytes())
    if nulls:
        valid = rng.random(n) >= 0.1
        price = pa.Array.from_buffers(pa.decimal128(12, 2), n,
                                      [pa.py_buffer(np.packbits(valid, bitorder="little").tobytes()), dec_buf],
                                      null_count=int((~valid).sum()))
    else:
        price = pa.Array.from_buffers(pa.decimal128(12, 2), n, [None, dec_buf], null_count=0)
    return pa.table({
        "i32": pa.array(rng.integers(-1000, 1000, n).astype(np.int32), mask=mask()),
        "i64": pa.array(rng.integers(-(1 << 62), 1 << 62, n)),
        "f32": pa.array(rng.normal(size=n).astype(np.float32), mask=mask()),
        "f64": pa.array(rng.normal(size=n)),
        "b": pa.array(rng.random(n) < 0.5, mask=mask()),
        "d": pa.array(rng.integers(0, 20_000, n).astype(np.int32), pa.date32()),
        "ts_us": pa.array(rng.integers(0, 1 << 50, n), pa.timestamp("us", tz="UTC"), mask=mask()),
        "ts_ns": pa.array(rng.integers(-(1 << 62), 1 << 62, n), pa.timestamp("ns")),
        "price": price,
    })


def _arrow_values(arr):
    """The Arrow values of a (possibly sliced) array as numpy, in the operators' layout."""
    n, off = len(arr), arr.offset
    bufs = arr.buffers()
    t = arr.type
    if pa.types.is_boolean(t):
        bits = np.frombuffer(bufs[1], np.uint8)
        return np.unpackbits(bits, bitorder="little")[off: off + n].astype(np.uint8)
    if pa.types.is_decimal(t):
        return np.frombuffer(bufs[1], np.int64).reshape(-1, 2)[off: off + n]
    width = {pa.int32(): 4, pa.date32(): 4, pa.float32(): 4}.get(t, 8)
    npdt = {4: np.int32, 8: np.int64}[width]
    if pa.types.is_floating(t):
        npdt = np.float32 if width == 4 else np.float64
    return np.frombuffer(bufs[1], npdt)[off: off + n]


def _arrow_valid(arr):
    return np.asarray(arr.is_valid().to_numpy(zero_copy_only=False), bool)


def _same(gt: GpuTable, table):
    assert gt.num_rows == table.num_rows
    assert gt.names == table.column_names
    for i, name in enumerate(table.column_names):
        arr = table.column(i).combine_chunks() if table.column(i).num_chunks != 1 else table.column(i).chunk(0)
        col = gt.column(i)
        assert col.dtype == WANT_DTYPE[name], name
        got = col.to_numpy()
        want = _arrow_values(arr)
        valid = _arrow_valid(arr)
        if pa.types.is_floating(arr.type):
            assert np.array_equal(got.view(np.uint8), np.ascontiguousarray(want).view(np.uint8)), name
        else:
            assert np.array_equal(got[valid], want[valid]), name
        if arr.null_count:
            assert np.array_equal(col.valid_numpy(), valid), name
        else:
            assert col.validity is None or col.valid_numpy().all(), name
    assert gt.column("ts_us").scale == 6 and gt.field(gt.names.index("ts_us"))[1] == "UTC"
    assert gt.column("ts_ns").scale == 9 and gt.column("price").scale == 2


@pytest.mark.parametrize("n", [1, 8, 1000, 300_007])
def test_record_batch_upload_is_byte_exact(gpu, n):
    print(f"[case] arrow upload n={n}", flush=True)
    t = _table(n, n)
    gt = table_from_arrow(t.to_batches()[0])
    _same(gt, t)
    gt.release()


def test_sliced_batches_and_unaligned_appends(gpu):
    print("[case] arrow slices", flush=True)
    t = _table(100_000, 5)
    parts = [t.slice(0, 13), t.slice(13, 7), t.slice(20, 49_991), t.slice(50_011, 1), t.slice(50_012, 49_988)]
    gt = GpuTable(parts[0].to_batches()[0])
    for p in parts[1:]:
        for b in p.to_batches():
            assert b.num_rows == 0 or b.column(0).offset == p.column(0).chunk(0).offset
            gt.append(b)
    _same(gt, t)


def test_nulls_first_appearing_in_a_later_batch(gpu):
    print("[case] arrow late nulls", flush=True)
    a = _table(1003, 1, nulls=False)
    b = _table(5001, 2, nulls=True)
    gt = GpuTable(a.to_batches()[0])
    assert gt.column("i32").validity is None
    gt.append(b.to_batches()[0])
    _same(gt, pa.concat_tables([a, b]))


@pytest.mark.parametrize("kind", ["file", "stream"])
def test_ipc_round_trip(gpu, kind, tmp_path):
    print(f"[case] arrow ipc {kind}", flush=True)
    t = pa.concat_tables([_table(70_001, 7), _table(30_000, 8), _table(3, 9)])
    path = tmp_path / f"t.{kind}"
    with pa.OSFile(str(path), "wb") as f:
        w = ipc.new_file(f, t.schema) if kind == "file" else ipc.new_stream(f, t.schema)
        for b in t.to_batches(max_chunksize=25_000):
            w.write_batch(b)
        w.close()
    gt = read_ipc(str(path))
    _same(gt, t)
    # and back out: every column equals the original Arrow array (decimals re-typed to precision 38)
    back = table_to_arrow(gt)
    for name in t.column_names:
        want = t.column(name).combine_chunks()
        got = back.column(name).combine_chunks()
        if pa.types.is_decimal(want.type):
            want = want.cast(pa.decimal128(38, 2))
        if pa.types.is_timestamp(want.type):
            want = want.cast(pa.timestamp(want.type.unit))
        assert got.equals(want), name


def test_column_export_round_trip(gpu):
    """bolt_column_to_arrow: values, NULLs (bits past the end cleared) and Bool packing."""
    from craton_bolt_amd.device import DeviceColumn

    for n in (1, 9, 4097):
        rng = np.random.default_rng(n)
        valid = rng.random(n) > 0.3
        v = rng.integers(-50, 50, n).astype(np.int64)
        got = column_to_arrow(DeviceColumn.from_numpy(v, N.I64, valid), "x")
        assert got.to_pylist() == [int(x) if ok else None for x, ok in zip(v, valid)]
        b = (rng.random(n) > 0.5).astype(np.uint8)
        got = column_to_arrow(DeviceColumn.from_numpy(b, N.BOOL))
        assert got.type == pa.bool_() and got.to_pylist() == [bool(x) for x in b]


def test_groupby_over_an_ingested_table_matches_oracle(gpu):
    """GROUP BY the Date32 column of a table read from an IPC stream: SUM(Int32 with NULLs),
    SUM(Decimal128), COUNT(Int32), MIN(Date32), all exact against the oracle on the Arrow buffers."""
    print("[case] arrow groupby", flush=True)
    t = _table(400_000, 11)
    gt = read_ipc(_ipc_bytes(t))
    names = ("i32", "price", "d")
    aggs = [(N.SUM, 0), (N.SUM, 1), (N.COUNT, 0), (N.MIN, 2)]
    res = execute_groupby([gt.column("d")], [gt.column(c) for c in names], aggs)
    odt = {"i32": O.I32, "price": O.DEC128, "d": O.DATE32}
    host = [O.HostCol(_arrow_values(t.column(c).chunk(0)), odt[c], _arrow_valid(t.column(c).chunk(0)),
                      2 if c == "price" else 0) for c in names]
    want = O.groupby([host[2]], host, aggs, max_groups=25_000)
    assert want["status"] == O.OK
    assert res["n_groups"] == want["n_groups"]
    assert np.array_equal(res["keys"][0][0].view(np.uint32).astype(np.int64), want["keys"])
    assert np.array_equal(res["aggs"][0][0], want["i64"][0])
    dec = res["aggs"][1][0]
    assert np.array_equal(dec[:, 0], want["i64"][1]) and np.array_equal(dec[:, 1], want["hi"][1])
    assert np.array_equal(res["aggs"][2][0], want["i64"][2])
    assert np.array_equal(res["aggs"][3][0].astype(np.int64), want["i64"][3])


def _ipc_bytes(t):
    sink = io.BytesIO()
    with ipc.new_stream(sink, t.schema) as w:
        w.write_table(t)
    return pa.py_buffer(sink.getvalue())


def test_unsupported_and_mismatched_batches_are_type_errors(gpu):
    with pytest.raises(N.BoltTypeError, match="unsupported Arrow dtype"):
        table_from_arrow(pa.table({"s": pa.array(["a", "b"])}))
    with pytest.raises(N.BoltTypeError, match="unsupported Arrow dtype"):
        table_from_arrow(pa.table({"d": pa.array([1, 2], pa.int32()).dictionary_encode()}))
    gt = table_from_arrow(pa.table({"x": pa.array([1, 2], pa.int64())}))
    with pytest.raises(N.BoltTypeError, match="appended column 0"):
        gt.append(pa.record_batch({"x": pa.array([1.0])}))
    assert gt.num_rows == 2
