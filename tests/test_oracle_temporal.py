"""Oracle pins for the temporal and decimal-cast ops of expression programs.

EXTRACT / DATE_TRUNC (src/jit/date_scalar.rs:459,539 over civil_from_days,
date_scalar.rs:37-60) are checked against Python's `datetime`, an independent
implementation of the proleptic Gregorian calendar, over two whole 400-year
eras and a sample of 0001-01-01 .. 9999-12-31 for Date32 and over instants
in each Timestamp unit (seconds .. nanoseconds, before and after the epoch).  The Decimal128 casts
(src/plan/physical_plan.rs:2046-2060) are checked against Python integer
arithmetic (wrapping to 128 / 64 / 32 bits, truncating division) and the
float -> decimal rounding against the known answers of the reference's
f64_to_i128_saturating tests (src/jit/ptx_gen.rs:4578-4640) and exact
rational arithmetic (`fractions`) on values where trunc(|x| + 0.5) and exact
round-half-away agree."""
import datetime as dtm
import fractions

import numpy as np
import pytest

import oracle as O

EPOCH = dtm.date(1970, 1, 1)
EPOCH_DT = dtm.datetime(1970, 1, 1)
FIELDS = {"year": 1, "month": 2, "day": 3, "hour": 4, "minute": 5, "second": 6}
X_LOAD, X_CONST, X_CAST, X_STORE, X_EXTRACT, X_DATE_TRUNC = 1, 2, 3, 8, 9, 10


def _op(kind, **kw):
    d = dict(kind=kind, dst=0, a=0, b=0, c=0, col=0, op=0, dtype=0, lit_i64=0, lit_f64=0.0, lit_hi=0)
    d.update(kw)
    return d


def _temporal(values, dtype, unit, kind, field):
    """One program: LOAD column 0, EXTRACT / DATE_TRUNC, STORE."""
    out_dt = O.I64 if kind == X_EXTRACT else dtype
    ops = [_op(X_LOAD, dst=0, col=0, dtype=dtype), _op(kind, dst=1, a=0, op=FIELDS[field]),
           _op(X_STORE, a=1, col=0)]
    st, vals, _ = O.project([O.HostCol(values, dtype, scale=unit)], ops, [out_dt], len(values),
                            out_scales=[unit])
    assert st == O.OK
    return vals[0]


def _day_sample():
    """Every day of 1600-01-01 .. 2400-12-31 (two 400-year eras, 292K days) and 200K days drawn from all
    of 0001-01-01 .. 9999-12-31."""
    lo, hi = dtm.date(1, 1, 1).toordinal(), dtm.date(9999, 12, 31).toordinal()
    every = np.arange(dtm.date(1600, 1, 1).toordinal(), dtm.date(2400, 12, 31).toordinal() + 1, dtype=np.int64)
    drawn = np.random.default_rng(7).integers(lo, hi + 1, 200_000)
    return np.concatenate([every, drawn, [lo, hi]]) - EPOCH.toordinal()


def test_date32_extract_matches_datetime():
    d = _day_sample()
    dates = [EPOCH + dtm.timedelta(days=int(x)) for x in d]
    d32 = d.astype(np.int32)
    for field in ("year", "month", "day"):
        want = np.array([getattr(x, field) for x in dates], np.int64)
        assert np.array_equal(_temporal(d32, O.DATE32, 0, X_EXTRACT, field), want), field


def test_date32_trunc_matches_datetime():
    d = _day_sample()
    dates = [EPOCH + dtm.timedelta(days=int(x)) for x in d]
    d32 = d.astype(np.int32)
    want = {
        "year": [(dtm.date(x.year, 1, 1) - EPOCH).days for x in dates],
        "month": [(dtm.date(x.year, x.month, 1) - EPOCH).days for x in dates],
        "day": list(d),
    }
    for field, w in want.items():
        assert np.array_equal(_temporal(d32, O.DATE32, 0, X_DATE_TRUNC, field), np.array(w, np.int32)), field


def _instants(unit, n, seed):
    """Ticks in `unit` digits of a second covering the unit's int64 range within years 1..9999, plus
    the instants around midnight, leap days and the epoch."""
    tps = 10 ** unit
    lo = max((dtm.datetime(1, 1, 1) - EPOCH_DT) // dtm.timedelta(seconds=1) * tps, -(1 << 63))
    hi = min((dtm.datetime(9999, 12, 31, 23, 59, 59) - EPOCH_DT) // dtm.timedelta(seconds=1) * tps, (1 << 63) - 1)
    rng = np.random.default_rng(seed)
    t = [int(x) for x in rng.integers(lo, hi, n, dtype=np.int64)]
    for day in (dtm.datetime(1970, 1, 1), dtm.datetime(2000, 2, 29), dtm.datetime(1900, 3, 1),
                dtm.datetime(1969, 12, 31, 23, 59, 59), dtm.datetime(1600, 2, 29, 12)):
        base = (day - EPOCH_DT) // dtm.timedelta(seconds=1) * tps
        t += [base + k for k in (-tps - 1, -1, 0, 1, tps - 1, tps) if lo <= base + k <= hi]
    return np.array(t, np.int64)


def _as_datetime(t, unit):
    s = t // 10 ** unit  # Python's floor division: instants before the epoch round down
    return EPOCH_DT + dtm.timedelta(seconds=s)


def _ticks(x, unit):
    """Ticks of x, wrapped to int64 as the 64-bit multiply does (years before 1677 in nanoseconds)."""
    v = (x - EPOCH_DT) // dtm.timedelta(seconds=1) * 10 ** unit
    return (v + (1 << 63)) % (1 << 64) - (1 << 63)


@pytest.mark.parametrize("unit", [0, 3, 6, 9])
def test_timestamp_extract_and_trunc_match_datetime(unit):
    t = _instants(unit, 50_000, seed=unit)
    dts = [_as_datetime(int(x), unit) for x in t]
    for field in FIELDS:
        want = np.array([getattr(x, field) for x in dts], np.int64)
        assert np.array_equal(_temporal(t, O.TIMESTAMP, unit, X_EXTRACT, field), want), field
    starts = {
        "year": lambda x: dtm.datetime(x.year, 1, 1),
        "month": lambda x: dtm.datetime(x.year, x.month, 1),
        "day": lambda x: dtm.datetime(x.year, x.month, x.day),
        "hour": lambda x: dtm.datetime(x.year, x.month, x.day, x.hour),
        "minute": lambda x: dtm.datetime(x.year, x.month, x.day, x.hour, x.minute),
        "second": lambda x: x,
    }
    for field, f in starts.items():
        want = np.array([_ticks(f(x), unit) for x in dts], np.int64)
        assert np.array_equal(_temporal(t, O.TIMESTAMP, unit, X_DATE_TRUNC, field), want), field


def test_temporal_type_errors():
    d = np.zeros(4, np.int32)
    ops = [_op(X_LOAD, dst=0, col=0, dtype=O.DATE32), _op(X_EXTRACT, dst=1, a=0, op=FIELDS["hour"]),
           _op(X_STORE, a=1, col=0)]
    assert O.project([O.HostCol(d, O.DATE32)], ops, [O.I64], 4)[0] == O.ERR_TYPE  # no HOUR on a date
    ops[1] = _op(X_EXTRACT, dst=1, a=0, op=FIELDS["year"])
    ops[0] = _op(X_LOAD, dst=0, col=0, dtype=O.I32)
    assert O.project([O.HostCol(d, O.I32)], ops, [O.I64], 4)[0] == O.ERR_TYPE  # EXTRACT of an integer
    # a Timestamp STORE into a column of another unit
    t = np.zeros(4, np.int64)
    ops = [_op(X_LOAD, dst=0, col=0, dtype=O.TIMESTAMP), _op(X_STORE, a=0, col=0)]
    assert O.project([O.HostCol(t, O.TIMESTAMP, scale=3)], ops, [O.TIMESTAMP], 4, out_scales=[6])[0] == O.ERR_TYPE
    assert O.project([O.HostCol(t, O.TIMESTAMP, scale=3)], ops, [O.TIMESTAMP], 4, out_scales=[3])[0] == O.OK


# ---- Decimal128 casts -------------------------------------------------------------

M128 = (1 << 128) - 1


def _s128(v):
    v &= M128
    return v - (1 << 128) if v >> 127 else v


def _dec_rows(vals):
    return np.array([[v & ((1 << 64) - 1), (v >> 64) & ((1 << 64) - 1)] for v in vals],
                    dtype=np.uint64).view(np.int64)


def _from_rows(a):
    a = a.view(np.uint64)
    return [_s128(int(lo) | (int(hi) << 64)) for lo, hi in a]


def _tdiv(a, b):
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


def _cast(col, src_dt, dst_dt, scale=0, from_scale=0):
    ops = [_op(X_LOAD, dst=0, col=0, dtype=src_dt), _op(X_CAST, dst=1, a=0, dtype=dst_dt, op=scale, c=from_scale),
           _op(X_STORE, a=1, col=0)]
    n = len(col.values)
    st, vals, _ = O.project([col], ops, [dst_dt], n, out_scales=[scale if dst_dt == O.DEC128 else 0])
    return st, vals[0]


def _dec_sample(seed):
    rng = np.random.default_rng(seed)
    v = [int(x) * (1 << 64) + int(y) for x, y in zip(rng.integers(-(1 << 40), 1 << 40, 300),
                                                      rng.integers(0, 1 << 63, 300))]
    v += [int(x) for x in rng.integers(-10 ** 15, 10 ** 15, 300)]
    v += [0, 1, -1, 5, -5, 15, -15, 149, -149, 150, -150, (1 << 127) - 1, -(1 << 127), 10 ** 38 - 1, -(10 ** 38 - 1)]
    return v


@pytest.mark.parametrize("sf,st", [(0, 0), (2, 4), (4, 2), (0, 9), (9, 0), (3, 3), (0, 38), (38, 0), (10, 30)])
def test_decimal_rescale_matches_integer_arithmetic(sf, st):
    v = _dec_sample(sf * 100 + st)
    status, out = _cast(O.HostCol(_dec_rows(v), O.DEC128, scale=sf), O.DEC128, O.DEC128, st, sf)
    assert status == O.OK
    want = [_s128(x * 10 ** (st - sf)) if st >= sf else _tdiv(x, 10 ** (sf - st)) for x in v]
    assert _from_rows(out) == want


@pytest.mark.parametrize("src,np_dt", [(O.I32, np.int32), (O.I64, np.int64), (O.DATE32, np.int32)])
@pytest.mark.parametrize("st", [0, 2, 18, 30])
def test_integer_to_decimal_matches_integer_arithmetic(src, np_dt, st):
    info = np.iinfo(np_dt)
    v = np.array([0, 1, -1, 7, -7, info.max, info.min, 123456789 % info.max, -987654 % info.max], np_dt)
    status, out = _cast(O.HostCol(v, src), src, O.DEC128, st)
    assert status == O.OK
    assert _from_rows(out) == [_s128(int(x) * 10 ** st) for x in v]


@pytest.mark.parametrize("dst,bits", [(O.I64, 64), (O.I32, 32)])
@pytest.mark.parametrize("sf", [0, 2, 12])
def test_decimal_to_integer_truncates_and_keeps_low_bits(dst, bits, sf):
    v = _dec_sample(sf + bits)
    status, out = _cast(O.HostCol(_dec_rows(v), O.DEC128, scale=sf), O.DEC128, dst, 0, sf)
    assert status == O.OK
    m = (1 << bits) - 1
    want = [(q & m) - ((q & m) >> (bits - 1) << bits) for q in (_tdiv(x, 10 ** sf) for x in v)]
    assert [int(x) for x in out] == want


def test_float_to_decimal_known_answers():
    """The answers of the reference's f64_to_i128_saturating tests (ptx_gen.rs:4578-4640), at scale 0."""
    cases = [(0.0, 0), (-0.0, 0), (1.0, 1), (-1.0, -1), (2.5, 3), (-2.5, -3), (0.5, 1), (-0.5, -1), (1.5, 2),
             (-1.5, -2), (2.4, 2), (-2.4, -2), (1e18, 10 ** 18), (-1e18, -10 ** 18),
             (1.2e38, 120_000_000_000_000_008_632_251_347_034_389_348_352),
             (-1.2e38, -120_000_000_000_000_008_632_251_347_034_389_348_352),
             (2.0e38, (1 << 127) - 1), (-2.0e38, -(1 << 127)), (3.3e38, (1 << 127) - 1), (-3.3e38, -(1 << 127)),
             (float("inf"), (1 << 127) - 1), (float("-inf"), -(1 << 127)), (float("nan"), 0)]
    x = np.array([c[0] for c in cases], np.float64)
    status, out = _cast(O.HostCol(x, O.F64), O.F64, O.DEC128, 0)
    assert status == O.OK
    assert _from_rows(out) == [c[1] for c in cases]


def _round_half_away(q: fractions.Fraction) -> int:
    a = abs(q)
    r = int(a + fractions.Fraction(1, 2))  # floor of a non-negative value
    return r if q >= 0 else -r


@pytest.mark.parametrize("st", [0, 2, 6, 12])
def test_float_to_decimal_rounds_half_away_from_zero(st):
    """Against exact rational rounding of the f64 product x * 10^st (the reference scales in f64 first),
    on products below 2^52, where adding 0.5 in f64 is exact and both roundings agree."""
    rng = np.random.default_rng(st)
    x = np.concatenate([rng.normal(size=2000) * 10.0 ** rng.integers(-3, 6, 2000),
                        np.array([0.125, -0.125, 2.675, -2.675, 1e-9, -1e-9, 0.005, -0.005])])
    x32 = x.astype(np.float32)
    for src, vals in ((O.F64, x), (O.F32, x32)):
        status, out = _cast(O.HostCol(vals, src), src, O.DEC128, st)
        assert status == O.OK
        prod = [float(v) * float(10.0 ** st) for v in vals]  # the f64 multiply
        keep = [abs(p) < 2.0 ** 52 for p in prod]
        got = _from_rows(out)
        want = [_round_half_away(fractions.Fraction(p)) for p in prod]
        assert all(g == w for g, w, k in zip(got, want, keep) if k)


@pytest.mark.parametrize("sf", [0, 2, 9, 22])
def test_decimal_to_float_matches_exact_division(sf):
    """Non-negative values below 2^53 convert exactly, so the result is the correctly rounded quotient (one
    division by an exact power of ten); a negative value is hi * 2^64 + lo with the UNSIGNED low word
    rounded to f64 first (ptx_gen.rs:1341-1356 converts lo with cvt.rn.f64.u64, then one fma), so
    -4147153880435083 comes out as -4147153880434688: computed here with exact rationals."""
    rng = np.random.default_rng(sf)
    pos = [int(x) for x in rng.integers(0, 1 << 53, 300)] + [0, 1, (1 << 53) - 1]
    neg = [int(x) for x in rng.integers(-(1 << 53), 0, 300)] + [-1, -4147153880435083, -(1 << 100) - 12345]
    v = pos + neg
    status, out = _cast(O.HostCol(_dec_rows(v), O.DEC128, scale=sf), O.DEC128, O.F64, 0, sf)
    assert status == O.OK
    assert [float(a) for a in out[: len(pos)]] == [x / 10 ** sf for x in pos]  # int / int rounds correctly

    def ref(x):
        hi, lo = x >> 64, x & ((1 << 64) - 1)
        raw = float(fractions.Fraction(float(hi)) * 2 ** 64 + fractions.Fraction(float(lo)))  # one rounding
        return raw / float(10 ** sf)
    assert [float(a) for a in out[len(pos):]] == [ref(x) for x in neg]
    assert ref(-4147153880435083) * 10 ** sf == -4147153880434688.0 or sf != 0
    status, out32 = _cast(O.HostCol(_dec_rows(pos), O.DEC128, scale=sf), O.DEC128, O.F32, 0, sf)
    assert status == O.OK
    assert np.array_equal(out32, np.array([x / 10 ** sf for x in pos], np.float64).astype(np.float32))


def test_decimal_cast_type_errors():
    d = O.HostCol(_dec_rows([1, 2]), O.DEC128, scale=2)
    assert _cast(d, O.DEC128, O.DATE32)[0] == O.ERR_TYPE
    assert _cast(d, O.DEC128, O.BOOL)[0] == O.ERR_TYPE
    assert _cast(d, O.DEC128, O.DEC128, 39, 2)[0] == O.ERR_TYPE
    assert _cast(d, O.DEC128, O.F64, 0, 23)[0] == O.ERR_TYPE
    assert _cast(O.HostCol(np.zeros(2, np.int64), O.TIMESTAMP), O.TIMESTAMP, O.DEC128, 2)[0] == O.ERR_TYPE
