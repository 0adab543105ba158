"""GPU parity: scalar aggregates (bolt_aggregate) against the oracle.

Integer, decimal, COUNT and MIN/MAX results must be bit-exact; float SUM /
AVG / VAR / STDDEV within REL_TOL = 1e-9 of the oracle's Neumaier / Welford
result (tests/common/mod.rs:32 uses the same tolerance)."""
import math

import numpy as np
import pytest

import oracle as O
from craton_bolt_amd import _native as N
from craton_bolt_amd.ops import Term, execute_aggregate

from gpu_util import both, dec_array, rel_close

REL_TOL = 1e-9
pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, params=["jit", "generic"])
def kernel_path(request, monkeypatch):
    """Every case runs through the per-query hipRTC kernel and through the
    precompiled generic kernel (BOLT_JIT=0)."""
    monkeypatch.setenv("BOLT_JIT", "1" if request.param == "jit" else "0")
    return request.param

ALL_OPS = [N.SUM, N.COUNT, N.COUNT_STAR, N.MIN, N.MAX, N.AVG, N.VAR_POP, N.VAR_SAMP, N.STDDEV_POP, N.STDDEV_SAMP]
FLOAT_RESULT = {N.AVG, N.VAR_POP, N.VAR_SAMP, N.STDDEV_POP, N.STDDEV_SAMP}


def _check(dev, host, ops, dtype):
    got = execute_aggregate([dev], [(op, 0) for op in ops])
    for op, g in zip(ops, got):
        want = O.scalar_value(O.scalar_agg(host, op))
        if dtype in (N.F32, N.F64) and op == N.SUM or op in FLOAT_RESULT:
            assert rel_close(g, want, REL_TOL), (op, g, want)
        elif dtype in (N.F32, N.F64) and op in (N.MIN, N.MAX):
            assert (g is None and want is None) or (math.isnan(g) and math.isnan(want)) or \
                (g == want and math.copysign(1, g) == math.copysign(1, want)), (op, g, want)
        else:
            assert g == want, (op, g, want)


@pytest.mark.parametrize("n", [0, 1, 7, 1000, 65_537, 1_000_003])
@pytest.mark.parametrize("dtype", [N.I32, N.I64, N.F32, N.F64])
def test_scalar_aggregates_match_oracle(gpu, n, dtype):
    print(f"[case] scalar n={n} dtype={dtype}", flush=True)
    rng = np.random.default_rng(n * 10 + dtype)
    if dtype == N.I32:
        x = rng.integers(-2 ** 31, 2 ** 31, size=n).astype(np.int32)
    elif dtype == N.I64:
        x = rng.integers(-2 ** 40, 2 ** 40, size=n).astype(np.int64)
    elif dtype == N.F32:
        x = rng.normal(size=n).astype(np.float32)
    else:
        x = rng.normal(size=n) * 1e3
    valid = rng.random(n) > 0.2
    dev, host = both(x, dtype, valid)
    _check(dev, host, ALL_OPS, dtype)
    dev, host = both(x, dtype)
    _check(dev, host, ALL_OPS, dtype)


def test_reference_null_fixtures(gpu):
    # tests/aggregate_nulls_e2e.rs:118-366
    cases = [
        (np.array([10, 0, -5, 30, 0, 7], np.int32), N.I32, [1, 0, 1, 1, 0, 1], [4, 6, 42, -5, 30, 10.5]),
        (np.array([100, 0, 0, 50, 25], np.int64), N.I64, [1, 0, 0, 1, 1], [3, 5, 175, 25, 100, 175 / 3]),
        (np.array([1.5, 0, 2.5, 0, -3.0, 4.0], np.float32), N.F32, [1, 0, 1, 0, 1, 1], [4, 6, 5.0, -3.0, 4.0, 1.25]),
        (np.array([10.0, 0, 20.0, 30.0, 0]), N.F64, [1, 0, 1, 1, 0], [3, 5, 60.0, 10.0, 30.0, 20.0]),
    ]
    for x, dt, valid, want in cases:
        dev, _ = both(x, dt, valid)
        got = execute_aggregate([dev], [(N.COUNT, 0), (N.COUNT_STAR, 0), (N.SUM, 0), (N.MIN, 0), (N.MAX, 0),
                                        (N.AVG, 0)])
        assert got[:5] == want[:5]
        assert abs(got[5] - want[5]) < 1e-12
    dev, _ = both(np.zeros(3, np.int64), N.I64, [0, 0, 0])
    got = execute_aggregate([dev], [(N.COUNT, 0), (N.SUM, 0), (N.MIN, 0), (N.MAX, 0), (N.AVG, 0)])
    assert got == [0, None, None, None, None]


def test_integer_sum_overflow_is_an_error(gpu):
    dev, _ = both(np.array([2 ** 63 - 1, 1], np.int64), N.I64)
    with pytest.raises(N.BoltTypeError, match="overflow"):
        execute_aggregate([dev], [(N.SUM, 0)])
    # the exact total decides, across many blocks
    x = np.full(3_000_000, 2 ** 42, np.int64)  # 3e6 * 2^42 > 2^63
    dev, _ = both(x, N.I64)
    with pytest.raises(N.BoltTypeError):
        execute_aggregate([dev], [(N.SUM, 0)])
    x[1::2] = -(2 ** 42)
    dev, host = both(x, N.I64)
    assert execute_aggregate([dev], [(N.SUM, 0)]) == [O.scalar_value(O.scalar_agg(host, O.SUM))]


def _row_order_overflows(vals):
    s = 0
    for v in vals:
        s += int(v)
        if not -2 ** 63 <= s < 2 ** 63:
            return True
    return False


@pytest.mark.parametrize("case", ["max_then_one", "climb_then_fall", "fall_then_climb", "late", "filtered", "nulls"])
def test_integer_sum_overflows_in_row_order(gpu, case):
    """SUM(Int64) fails as soon as a running total in row order leaves i64
    (checked_add, src/exec/aggregate.rs:2120-2140), also when the total fits."""
    big = 2 ** 63 - 1
    rng = np.random.default_rng(5)
    terms, valid = None, None
    if case == "max_then_one":
        x = np.array([big, 1, -1], np.int64)
    elif case == "climb_then_fall":
        x = np.array([2 ** 62, 2 ** 62, -2 ** 62, -2 ** 62], np.int64)
    elif case == "fall_then_climb":
        x = np.array([-2 ** 62, -2 ** 62, -2 ** 62, 2 ** 62, 2 ** 62, 2 ** 62], np.int64)
    elif case == "late":
        # 2M rows of small values, then the excursion near the end (many workgroups)
        x = rng.integers(-1000, 1000, 2_000_000).astype(np.int64)
        x[-3:] = [big, big, -big]
    elif case == "filtered":
        # the row that would overflow is filtered out: no error
        x = np.array([big, 1, -1, 5], np.int64)
        terms = [Term(0, N.LT, big)]
    else:
        # the row that would overflow is NULL: no error
        x = np.array([big, 1, -1, -5], np.int64)
        valid = [1, 0, 1, 1]
    dev, host = both(x, N.I64, valid)
    kept = x if terms is None else x[x < big]
    if valid is not None:
        kept = x[np.array(valid, bool)]
    want = O.scalar_agg(host, O.SUM) if terms is None else None
    if _row_order_overflows(kept):
        assert want is None or want.status != O.OK
        with pytest.raises(N.BoltTypeError, match="SUM\\(integer\\) overflow"):
            execute_aggregate([dev], [(N.SUM, 0)], terms)
    else:
        assert execute_aggregate([dev], [(N.SUM, 0), (N.COUNT, 0)], terms) == [int(sum(int(v) for v in kept)),
                                                                            len(kept)]


def test_decimal_sum_fixtures(gpu):
    # tests/decimal_type_test.rs:324-431
    dev, _ = both(dec_array([123, -456, 789, 1, 10_000]), N.DEC128, scale=2)
    assert execute_aggregate([dev], [(N.SUM, 0)]) == [123 - 456 + 789 + 1 + 10_000]
    dev, _ = both(dec_array([100, 0, -25, 0, 75]), N.DEC128, [1, 0, 1, 0, 1], scale=2)
    assert execute_aggregate([dev], [(N.SUM, 0)]) == [150]
    big = 10 ** 38 - 1
    dev, _ = both(dec_array([big, big]), N.DEC128)
    with pytest.raises(N.BoltTypeError, match="SUM\\(Decimal128\\) precision overflow"):
        execute_aggregate([dev], [(N.SUM, 0)])
    rng = np.random.default_rng(4)
    vals = [int(v) * 10 ** 20 + int(w) for v, w in zip(rng.integers(-10 ** 15, 10 ** 15, 100_000),
                                                       rng.integers(0, 10 ** 18, 100_000))]
    dev, host = both(dec_array(vals), N.DEC128)
    assert execute_aggregate([dev], [(N.SUM, 0)]) == [sum(vals)]


def test_float_min_max_nan_order(gpu):
    x = np.array([np.nan, 1.0, -2.0, np.inf, -0.0, 0.0] * 1000)
    dev, host = both(x, N.F64)
    _check(dev, host, [N.MIN, N.MAX, N.SUM], N.F64)
    dev, host = both(np.array([np.nan, np.nan]), N.F64)
    _check(dev, host, [N.MIN, N.MAX], N.F64)


def test_filtered_aggregate(gpu):
    rng = np.random.default_rng(9)
    n = 500_000
    a = rng.uniform(0, 100, n)
    b = rng.integers(-1000, 1000, n).astype(np.int32)
    vb = rng.random(n) > 0.1
    da, ha = both(a, N.F64)
    db, hb = both(b, N.I32, vb)
    got = execute_aggregate([da, db], [(N.SUM, 1), (N.COUNT_STAR, 0), (N.SUM, 0), (N.MAX, 1)],
                            [Term(0, N.GT, 10.0), Term(1, N.GE, -500)])
    rows = O.filter_rows([ha, hb], [(0, O.GT, 10.0), (1, O.GE, -500)], n)
    assert got[0] == int(b[rows].astype(np.int64).sum())
    assert got[1] == len(rows)
    assert rel_close(got[2], math.fsum(a[rows]), REL_TOL)
    assert got[3] == int(b[rows].max())
