"""GPU parity: filter + stream compaction and the mask prefix scan, against
the oracle's filter (src/exec/compact.rs:146 semantics).  Row sets and values
are bit-exact, order preserved."""
import numpy as np
import pytest

import oracle as O
from craton_bolt_amd import _native as N
from craton_bolt_amd.device import DeviceBuffer, DeviceColumn
from craton_bolt_amd.ops import Term, compact_columns, prefix_scan_mask, take

from gpu_util import both, dec_array

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 100, 4096, 4097, 1_000_003, 9_000_000])
def test_compact_matches_oracle(gpu, n):
    print(f"[case] compact n={n}", flush=True)
    rng = np.random.default_rng(n)
    a = rng.uniform(0, 100, n)
    b = rng.integers(-1000, 1000, n).astype(np.int32)
    c = rng.integers(-2 ** 62, 2 ** 62, n).astype(np.int64)
    vb = rng.random(n) > 0.1
    da, ha = both(a, N.F64)
    db, hb = both(b, N.I32, vb)
    dc, hc = both(c, N.I64)
    outs, k = compact_columns([da, db, dc], [Term(0, N.GT, 10.0), Term(1, N.LT, 500)])
    rows = O.filter_rows([ha, hb, hc], [(0, O.GT, 10.0), (1, O.LT, 500)], n)
    assert k == len(rows)
    assert np.array_equal(outs[0].to_numpy(), a[rows])
    assert np.array_equal(outs[1].to_numpy(), b[rows])
    assert np.array_equal(outs[1].valid_numpy(), vb[rows])
    assert np.array_equal(outs[2].to_numpy(), c[rows])


def test_compact_with_mask_and_nulls(gpu):
    rng = np.random.default_rng(3)
    n = 300_001
    x = rng.normal(size=n).astype(np.float32)
    vx = rng.random(n) > 0.3
    mask = (rng.random(n) > 0.5).astype(np.uint8)
    dx, hx = both(x, N.F32, vx)
    outs, k = compact_columns([dx], [Term(0, N.IS_NOT_NULL)], mask=DeviceBuffer.from_numpy(mask))
    want = np.nonzero(vx & (mask != 0))[0]
    assert k == len(want)
    assert np.array_equal(outs[0].to_numpy(), x[want])
    assert outs[0].valid_numpy().all()
    outs, k = compact_columns([dx], [Term(0, N.IS_NULL)])
    assert k == int((~vx).sum())


def test_compact_decimal_and_none_pass(gpu):
    vals = [10 ** 30, -5, 7, 10 ** 20 + 3]
    dd, _ = both(dec_array(vals), N.DEC128)
    outs, k = compact_columns([dd], [Term(0, N.GT, 6)])
    got = [(int(h) << 64) | (int(l) & (2 ** 64 - 1)) for l, h in outs[0].to_numpy(k)]
    assert got == [10 ** 30, 7, 10 ** 20 + 3]
    outs, k = compact_columns([dd], [Term(0, N.GT, 10 ** 31)])
    assert k == 0


def test_compact_output_capacity(gpu):
    """Outputs sized to the survivors suffice; a smaller capacity declines
    with the full count and writes nothing past it."""
    import ctypes as C

    from craton_bolt_amd.ops import _filter, columns_array

    n = 300_000
    rng = np.random.default_rng(3)
    a = rng.integers(0, 1000, n).astype(np.int64)
    da = DeviceColumn.from_numpy(a)
    keep = a < 100
    want = a[keep]
    k = len(want)
    terms = [Term(0, N.LT, 100)]
    f, _ = _filter(terms, [da])
    for cap, declines in [(k, False), (k - 1000, True)]:
        out = DeviceColumn.empty(cap + 4096, N.I64)
        out.values.copy_from(np.full(cap + 4096, -7, np.int64))
        oc = (N.BoltColumn * 1)(out.c(length=cap))
        n_out = C.c_int64()
        st = N.lib().bolt_filter_compact(None, columns_array([da]), 1, C.byref(f), None, oc, C.byref(n_out))
        assert st == (N.BOLT_DECLINE if declines else N.BOLT_OK)
        assert n_out.value == k
        got = out.values.to_numpy(np.int64, cap + 4096)
        assert np.array_equal(got[:cap], want[:cap])
        assert np.all(got[cap:] == -7)


@pytest.mark.parametrize("n", [1, 4096, 5000, 10_000_000])
def test_prefix_scan_mask(gpu, n):
    rng = np.random.default_rng(n + 1)
    mask = (rng.random(n) > 0.3).astype(np.uint8)
    off, total = prefix_scan_mask(DeviceBuffer.from_numpy(mask), n)
    want = np.concatenate([[0], np.cumsum(mask, dtype=np.int64)[:-1]]).astype(np.uint32)
    assert total == int(mask.sum())
    assert np.array_equal(off.to_numpy(np.uint32, n), want)


@pytest.mark.parametrize("n", [1, 63, 64, 65, 1000, 100_003])
@pytest.mark.parametrize("dt", [N.BOOL, N.I32, N.I64, N.F64, N.DEC128])
def test_take_gathers_values_and_validity(gpu, n, dt):
    """bolt_take: out[i] = src[rows[i]] (arrow take), validity bits likewise,
    for every width and row counts around the 64-row wave that writes the
    bitmap."""
    print(f"[case] take n={n} dt={dt}", flush=True)
    rng = np.random.default_rng(n * 10 + dt)
    m = 5000
    if dt == N.DEC128:
        src = rng.integers(-2 ** 62, 2 ** 62, (m, 2)).astype(np.int64)
    elif dt == N.BOOL:
        src = rng.integers(0, 2, m).astype(np.uint8)
    else:
        src = rng.integers(-2 ** 30, 2 ** 30, m).astype({N.I32: np.int32, N.I64: np.int64, N.F64: np.float64}[dt])
    valid = rng.random(m) > 0.2
    rows = rng.integers(0, m, n).astype(np.uint32)
    col = DeviceColumn.from_numpy(src, dt, valid=valid)
    out = take(col, DeviceBuffer.from_numpy(rows), n)
    assert out.length == n and out.dtype == dt
    assert np.array_equal(out.to_numpy(n), src[rows])
    assert np.array_equal(out.valid_numpy(n), valid[rows])
    plain = take(DeviceColumn.from_numpy(src, dt), DeviceBuffer.from_numpy(rows), n)
    assert plain.validity is None and np.array_equal(plain.to_numpy(n), src[rows])


def test_take_rejects_out_of_range_row_ids(gpu):
    col = DeviceColumn.from_numpy(np.arange(10, dtype=np.int64))
    rows = np.array([0, 5, 10, 3], np.uint32)
    with pytest.raises(N.PlanError, match="out of range"):
        take(col, DeviceBuffer.from_numpy(rows), 4)
