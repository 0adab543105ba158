"""GPU parity: EXTRACT / DATE_TRUNC on Date32 and Timestamp columns and the
Decimal128 casts of expression programs (bolt_project), bit-exact against
the oracle's interpreter of the same program, NULLs included; and a program
mixing them (DATE_TRUNC feeding a comparison, a decimal rescale feeding
arithmetic)."""
import numpy as np
import pytest

import oracle as O
from craton_bolt_amd import _native as N
from craton_bolt_amd.ops import Program, execute_projection

from gpu_util import both, dec_array

pytestmark = pytest.mark.gpu

FIELDS = ("year", "month", "day", "hour", "minute", "second")


def _check(cols, prog, out_dtypes, out_validity=None, out_scales=None):
    dev = [d for d, _ in cols]
    host = [h for _, h in cols]
    outs = execute_projection(dev, prog, out_dtypes, out_validity, out_scales=out_scales)
    n = dev[0].length
    st, want, want_valid = O.project(host, prog.ops, out_dtypes, n, out_validity, out_scales)
    assert st == O.OK
    for k, (o, w) in enumerate(zip(outs, want)):
        got = o.to_numpy()
        if o.dtype in (N.F32, N.F64):
            bits = np.uint32 if o.dtype == N.F32 else np.uint64
            nan = np.isnan(w)
            assert np.array_equal(np.isnan(got), nan), k
            assert np.array_equal(got.view(bits)[~nan], w.view(bits)[~nan]), k
        else:
            assert np.array_equal(got.reshape(w.shape), w), k
        if out_validity and out_validity[k]:
            assert np.array_equal(o.valid_numpy(), want_valid[k]), k


@pytest.mark.parametrize("n", [1, 5000, 1_000_003])
def test_date32_extract_and_trunc(gpu, n):
    print(f"[case] date32 extract/trunc n={n}", flush=True)
    rng = np.random.default_rng(n)
    d = rng.integers(-719_162, 2_932_896, n).astype(np.int32)  # 0001-01-01 .. 9999-12-31
    d[: min(n, 4)] = [0, -1, 11_016, -2_147_483_648][: min(n, 4)]  # epoch, 1969-12-31, 2000-02-29, INT32_MIN
    valid = rng.random(n) > 0.1
    cols = [both(d, N.DATE32, valid)]
    p = Program()
    a = p.load(0, N.DATE32)
    outs = []
    for f in FIELDS[:3]:
        p.store(p.extract(f, a), len(outs))
        outs.append(N.I64)
    for f in FIELDS[:3]:
        p.store(p.date_trunc(f, a), len(outs))
        outs.append(N.DATE32)
    _check(cols, p, outs, [True] * len(outs))


@pytest.mark.parametrize("unit", [N.TS_S, N.TS_MS, N.TS_US, N.TS_NS])
def test_timestamp_extract_and_trunc(gpu, unit):
    n = 1_000_003
    print(f"[case] timestamp extract/trunc unit={unit} n={n}", flush=True)
    rng = np.random.default_rng(unit)
    lim = min(253_402_300_799 * 10 ** unit, (1 << 63) - 1)
    t = rng.integers(-lim if unit < 9 else -(1 << 63), lim, n, dtype=np.int64)
    t[:6] = [0, -1, 1, -(1 << 63), (1 << 63) - 1, 951_782_400 * 10 ** unit]  # extremes, 2000-02-29
    valid = rng.random(n) > 0.1
    cols = [both(t, N.TIMESTAMP, valid, scale=unit)]
    p = Program()
    a = p.load(0, N.TIMESTAMP)
    outs, scales = [], []
    for f in FIELDS:
        p.store(p.extract(f, a), len(outs))
        outs.append(N.I64)
        scales.append(0)
    for f in FIELDS:
        p.store(p.date_trunc(f, a), len(outs))
        outs.append(N.TIMESTAMP)
        scales.append(unit)
    _check(cols, p, outs, [True] * len(outs), scales)


def test_timestamp_filter_on_truncated_month(gpu):
    """WHERE DATE_TRUNC('month', ts) = TIMESTAMP '2024-02-01' AND EXTRACT(HOUR FROM ts) >= 12, as a Bool."""
    n = 2_000_000
    rng = np.random.default_rng(3)
    base = 1_704_067_200_000  # 2024-01-01 in milliseconds
    t = base + rng.integers(0, 120 * 86_400_000, n, dtype=np.int64)
    cols = [both(t, N.TIMESTAMP, None, scale=N.TS_MS)]
    p = Program()
    a = p.load(0, N.TIMESTAMP)
    month = p.date_trunc("month", a)
    eq = p.binary("==", month, p.const(1_706_745_600_000, N.TIMESTAMP, unit=N.TS_MS), N.TIMESTAMP)
    pm = p.binary(">=", p.extract("hour", a), p.const(12, N.I64), N.I64)
    p.store(p.binary("and", eq, pm, N.BOOL), 0)
    _check(cols, p, [N.BOOL])


def _dec_vals(rng, n):
    v = [int(x) * 10 ** 12 + int(y) for x, y in zip(rng.integers(-10 ** 12, 10 ** 12, n), rng.integers(0, 10 ** 12, n))]
    v[:6] = [0, 1, -1, (1 << 127) - 1, -(1 << 127), 10 ** 38 - 1][: min(n, 6)]
    return v


@pytest.mark.parametrize("n", [7, 300_001])
def test_decimal_casts(gpu, n):
    print(f"[case] decimal casts n={n}", flush=True)
    rng = np.random.default_rng(n)
    dv = _dec_vals(rng, n)
    i64 = rng.integers(-(1 << 62), 1 << 62, n).astype(np.int64)
    i32 = rng.integers(-(1 << 31), 1 << 31, n).astype(np.int32)
    f64 = rng.normal(size=n) * 10.0 ** rng.integers(-6, 20, n)
    f64[: min(n, 5)] = [2.5, -2.5, np.nan, np.inf, -1e39][: min(n, 5)]
    f32 = (rng.normal(size=n) * 1e3).astype(np.float32)
    d32 = rng.integers(-100_000, 100_000, n).astype(np.int32)
    valid = rng.random(n) > 0.05
    cols = [both(dec_array(dv), N.DEC128, valid, scale=6), both(i64, N.I64), both(i32, N.I32, valid),
            both(f64, N.F64), both(f32, N.F32), both(d32, N.DATE32)]
    p = Program()
    dec, a64, a32, af, af32, ad = (p.load(0, N.DEC128), p.load(1, N.I64), p.load(2, N.I32), p.load(3, N.F64),
                                   p.load(4, N.F32), p.load(5, N.DATE32))
    plan = [
        (p.cast(dec, N.DEC128, scale=9, from_scale=6), N.DEC128, 9),   # rescale up
        (p.cast(dec, N.DEC128, scale=2, from_scale=6), N.DEC128, 2),   # rescale down, truncating
        (p.cast(dec, N.I64, from_scale=6), N.I64, 0),
        (p.cast(dec, N.I32, from_scale=6), N.I32, 0),
        (p.cast(dec, N.F64, from_scale=6), N.F64, 0),
        (p.cast(dec, N.F32, from_scale=6), N.F32, 0),
        (p.cast(a64, N.DEC128, scale=4), N.DEC128, 4),
        (p.cast(a32, N.DEC128, scale=18), N.DEC128, 18),
        (p.cast(af, N.DEC128, scale=2), N.DEC128, 2),
        (p.cast(af32, N.DEC128, scale=3), N.DEC128, 3),
        (p.cast(ad, N.DEC128, scale=0), N.DEC128, 0),
        # a rescale feeding decimal arithmetic: price(6) -> (2) + 1.00
        (p.binary("+", p.cast(dec, N.DEC128, scale=2, from_scale=6), p.const(100, N.DEC128), N.DEC128), N.DEC128, 2),
    ]
    for k, (r, _, _) in enumerate(plan):
        p.store(r, k)
    _check(cols, p, [dt for _, dt, _ in plan], [True] * len(plan), [s for _, _, s in plan])


def test_temporal_and_cast_type_errors(gpu):
    d = both(np.zeros(4, np.int32), N.DATE32)
    p = Program()
    p.store(p.extract("hour", p.load(0, N.DATE32)), 0)
    with pytest.raises(N.BoltTypeError, match="time-of-day field from Date32"):
        execute_projection([d[0]], p, [N.I64])
    t = both(np.zeros(4, np.int64), N.TIMESTAMP, scale=N.TS_MS)
    p = Program()
    p.store(p.binary("+", p.load(0, N.TIMESTAMP), p.load(0, N.TIMESTAMP), N.TIMESTAMP), 0)
    with pytest.raises(N.BoltTypeError, match="not defined on Timestamp"):
        execute_projection([t[0]], p, [N.TIMESTAMP], out_scales=[N.TS_MS])
    p = Program()
    p.store(p.load(0, N.TIMESTAMP), 0)
    with pytest.raises(N.BoltTypeError, match="Timestamp unit"):
        execute_projection([t[0]], p, [N.TIMESTAMP], out_scales=[N.TS_US])
    p = Program()
    p.store(p.cast(p.load(0, N.TIMESTAMP), N.DEC128, scale=2), 0)
    with pytest.raises(N.BoltTypeError, match="no cast from Timestamp to Decimal128"):
        execute_projection([t[0]], p, [N.DEC128], out_scales=[2])


def test_timestamp_columns_enter_operators_through_programs(gpu):
    """Operators that take columns directly reject a Timestamp column (its unit would be lost); COUNT
    only reads validity and is allowed."""
    from craton_bolt_amd.ops import Term, execute_aggregate, execute_groupby

    t = both(np.arange(10, dtype=np.int64), N.TIMESTAMP, scale=N.TS_US)[0]
    k = both(np.zeros(10, np.int32), N.I32)[0]
    with pytest.raises(N.BoltTypeError, match="Timestamp enters through an expression program"):
        execute_aggregate([t], [(N.MIN, 0)])
    with pytest.raises(N.BoltTypeError, match="Timestamp enters through an expression program"):
        execute_groupby([k], [t], [(N.MAX, 0)])
    with pytest.raises(N.BoltTypeError, match="Timestamp is compared inside an expression program"):
        execute_aggregate([t], [(N.COUNT_STAR, 0)], [Term(0, N.GT, 3)])
    assert execute_aggregate([t], [(N.COUNT, 0)]) == [10]
