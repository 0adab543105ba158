"""GPU parity: per-row expression programs (bolt_project) against the oracle's
interpreter of the same program — bit-exact for every dtype, floats included
(one IEEE rounding per op on both sides; NaN payloads aside), and validity
bitmaps equal."""
import numpy as np
import pytest

import oracle as O
from craton_bolt_amd import _native as N
from craton_bolt_amd.ops import Program, execute_projection

from gpu_util import both, dec_array

pytestmark = pytest.mark.gpu


def _check(cols, prog, out_dtypes, out_validity=None):
    dev = [d for d, _ in cols]
    host = [h for _, h in cols]
    outs = execute_projection(dev, prog, out_dtypes, out_validity)
    n = dev[0].length
    st, want, want_valid = O.project(host, prog.ops, out_dtypes, n, out_validity)
    assert st == O.OK
    for k, (o, w) in enumerate(zip(outs, want)):
        got = o.to_numpy()
        if o.dtype in (N.F32, N.F64):
            # bit-exact, except that a NaN's payload is not compared: x86 (the
            # oracle) makes inf - inf the negative default NaN, the GPU the
            # positive one, and SQL has one NaN
            bits = np.uint32 if o.dtype == N.F32 else np.uint64
            nan = np.isnan(w)
            assert np.array_equal(np.isnan(got), nan), k
            assert np.array_equal(got.view(bits)[~nan], w.view(bits)[~nan]), k
        else:
            assert np.array_equal(got, w), k
        if out_validity and out_validity[k]:
            assert np.array_equal(o.valid_numpy(), want_valid[k]), k


def _ints(rng, n, dtype, nulls):
    v = rng.integers(-(1 << 31), 1 << 31, n).astype(np.int64)
    v[rng.integers(0, n, max(n // 50, 1))] = 0
    v[rng.integers(0, n, 5)] = -1
    v[rng.integers(0, n, 5)] = -(1 << 31) if dtype == N.I32 else -(1 << 63)
    valid = rng.random(n) > 0.1 if nulls else None
    return both(v.astype(np.int32 if dtype == N.I32 else np.int64), dtype, valid)


@pytest.mark.parametrize("n", [1, 1000, 1_000_003])
@pytest.mark.parametrize("dtype", [N.I32, N.I64])
def test_project_integer_ops(gpu, n, dtype):
    print(f"[case] project ints dtype={dtype} n={n}", flush=True)
    rng = np.random.default_rng(n + dtype)
    cols = [_ints(rng, n, dtype, True), _ints(rng, n, dtype, False)]
    p = Program()
    a, b = p.load(0, dtype), p.load(1, dtype)
    outs = []
    for op in ("+", "-", "*", "/", "%", "&", "|"):
        p.store(p.binary(op, a, b, dtype), len(outs))
        outs.append(dtype)
    for op in ("==", "<", ">="):
        p.store(p.binary(op, a, b, dtype), len(outs))
        outs.append(N.BOOL)
    _check(cols, p, outs, [True] * len(outs))


@pytest.mark.parametrize("dtype", [N.F32, N.F64])
def test_project_float_ops_and_casts(gpu, dtype):
    n = 500_000
    rng = np.random.default_rng(dtype)
    npdt = np.float32 if dtype == N.F32 else np.float64
    x = (rng.standard_normal(n) * 1e3).astype(npdt)
    y = (rng.standard_normal(n) * 1e3).astype(npdt)
    for v in (x, y):
        v[rng.integers(0, n, 100)] = np.nan
        v[rng.integers(0, n, 100)] = 0.0
        v[rng.integers(0, n, 100)] = -0.0
        v[rng.integers(0, n, 50)] = np.inf
        v[rng.integers(0, n, 50)] = npdt(3e10)
        v[rng.integers(0, n, 50)] = npdt(-3e19)
    cols = [both(x, dtype, rng.random(n) > 0.05), both(y, dtype)]
    p = Program()
    a, b = p.load(0, dtype), p.load(1, dtype)
    specs = []
    for op in ("+", "-", "*", "/"):
        p.store(p.binary(op, a, b, dtype), len(specs))
        specs.append(dtype)
    p.store(p.binary("+", p.binary("*", a, b, dtype), a, dtype), len(specs))  # a * b + a, no contraction
    specs.append(dtype)
    for op in ("==", "!=", "<", "<=", ">", ">="):
        p.store(p.binary(op, a, b, dtype), len(specs))
        specs.append(N.BOOL)
    for to in (N.I32, N.I64, N.BOOL, N.F32 if dtype == N.F64 else N.F64):
        p.store(p.cast(a, to), len(specs))
        specs.append(to)
    _check(cols, p, specs, [True] * len(specs))


def test_project_int_to_float_and_bool_casts(gpu):
    n = 300_000
    rng = np.random.default_rng(3)
    big = rng.integers(-(1 << 62), 1 << 62, n).astype(np.int64)
    flags = rng.integers(0, 2, n).astype(np.uint8)
    small = rng.integers(-1000, 1000, n).astype(np.int32)
    cols = [both(big, N.I64), both(flags, N.BOOL), both(small, N.I32)]
    p = Program()
    g, f, s = p.load(0, N.I64), p.load(1, N.BOOL), p.load(2, N.I32)
    specs = []
    for src, to in ((g, N.F32), (g, N.F64), (g, N.I32), (s, N.I64), (s, N.F32), (f, N.I32), (f, N.F64), (g, N.BOOL),
                    (s, N.BOOL)):
        p.store(p.cast(src, to), len(specs))
        specs.append(to)
    p.store(p.binary("and", f, p.cast(s, N.BOOL), N.BOOL), len(specs))
    specs.append(N.BOOL)
    p.store(p.binary("or", p.not_(f), p.binary(">", s, p.const(0, N.I32), N.I32), N.BOOL), len(specs))
    specs.append(N.BOOL)
    _check(cols, p, specs)


def test_project_case_when_and_is_null(gpu):
    # CASE WHEN qty IS NULL THEN 0 WHEN qty > 10 THEN price * 2 ELSE price END
    n = 700_000
    rng = np.random.default_rng(11)
    qty = rng.integers(0, 20, n).astype(np.int32)
    price = rng.standard_normal(n) * 100
    cols = [both(qty, N.I32, rng.random(n) > 0.2), both(price, N.F64, rng.random(n) > 0.1)]
    p = Program()
    q, pr = p.load(0, N.I32), p.load(1, N.F64)
    inner = p.select(p.binary(">", q, p.const(10, N.I32), N.I32), p.binary("*", pr, p.const(2.0, N.F64), N.F64), pr)
    p.store(p.select(p.is_null(0), p.const(0.0, N.F64), inner), 0)
    p.store(p.is_null(1, want_null=False), 1)
    _check(cols, p, [N.F64, N.BOOL], [True, True])


def test_project_decimal_and_date(gpu):
    n = 200_000
    rng = np.random.default_rng(13)
    xs = [int(a) * (1 << 64) + int(b) for a, b in zip(rng.integers(-(1 << 40), 1 << 40, n),
                                                       rng.integers(0, 1 << 63, n))]
    ys = [int(a) for a in rng.integers(-(1 << 20), 1 << 20, n)]
    d1 = rng.integers(0, 30_000, n).astype(np.int32)
    d2 = rng.integers(0, 30_000, n).astype(np.int32)
    cols = [both(dec_array(xs), N.DEC128, rng.random(n) > 0.05, scale=2), both(dec_array(ys), N.DEC128, None, 2),
            both(d1, N.DATE32), both(d2, N.DATE32)]
    p = Program()
    x, y = p.load(0, N.DEC128), p.load(1, N.DEC128)
    specs = []
    for op in ("+", "-", "*"):
        p.store(p.binary(op, x, y, N.DEC128), len(specs))
        specs.append(N.DEC128)
    for op in ("<", "==", ">="):
        p.store(p.binary(op, x, y, N.DEC128), len(specs))
        specs.append(N.BOOL)
    p.store(p.binary("+", x, p.const(-(10 ** 30), N.DEC128), N.DEC128), len(specs))
    specs.append(N.DEC128)
    p.store(p.binary("-", p.load(2, N.DATE32), p.load(3, N.DATE32), N.DATE32), len(specs))
    specs.append(N.I32)
    _check(cols, p, specs, [True] * len(specs))


def test_project_errors(gpu):
    d, _ = both(np.arange(10, dtype=np.int32), N.I32)
    p = Program()
    a = p.load(0, N.I32)
    p.store(p.binary("+", a, p.const(1, N.I64), N.I32), 0)
    with pytest.raises(N.BoltError) as e:
        execute_projection([d], p, [N.I32])
    assert e.value.code == N.BOLT_ERR_TYPE
    q = Program()
    q.load(0, N.I32)
    with pytest.raises(N.BoltError) as e:  # output 0 never stored
        execute_projection([d], q, [N.I32])
    assert e.value.code == N.BOLT_ERR_PLAN
    r = Program()
    r.store(r.binary("%", r.load(0, N.I32), r.load(0, N.I32), N.I32), 0)
    r.ops.append(dict(r.ops[-1]))  # the same output stored twice
    with pytest.raises(N.BoltError) as e:
        execute_projection([d], r, [N.I32])
    assert e.value.code == N.BOLT_ERR_PLAN


# ---- fused WHERE + aggregates over expression programs (bolt_aggregate_expr) ----------------------

from craton_bolt_amd.ops import execute_aggregate_expr  # noqa: E402

SUM, COUNT, COUNT_STAR, MIN, MAX, AVG, VAR_POP, STDDEV_SAMP = 1, 2, 3, 4, 5, 6, 7, 10
REL_TOL = 1e-9


def _agg_check(cols, prog, value_dtypes, filter_value, aggs):
    dev = [d for d, _ in cols]
    host = [h for _, h in cols]
    got = execute_aggregate_expr(dev, prog, len(value_dtypes), aggs, filter_value)
    want = O.aggregate_expr(host, prog.ops, value_dtypes, filter_value, aggs, dev[0].length)
    for (op, col), g, w in zip(aggs, got, want):
        if op == COUNT_STAR:
            assert g == w
            continue
        assert w.status == O.OK
        if w.is_null:
            assert g is None
        elif w.dtype in (O.F32, O.F64):
            assert abs(g - w.f64) <= REL_TOL * max(abs(w.f64), 1e-300), (op, g, w.f64)
        elif w.dtype == O.DEC128:
            assert g == (int(w.hi) << 64) | (int(w.i64) & ((1 << 64) - 1))
        else:
            assert g == w.i64, (op, g, w.i64)


@pytest.mark.parametrize("n", [0, 1, 10_000, 4_000_037])
def test_aggregate_expr_q6_shape(gpu, n):
    """SUM(price * disc), COUNT(*), MIN/MAX/AVG(price * disc) WHERE qty < 24 AND disc BETWEEN 0.05 AND 0.07
    AND price + qty > 100 (a TPC-H Q6 shape with an expression predicate)."""
    print(f"[case] aggregate_expr q6 n={n}", flush=True)
    rng = np.random.default_rng(n + 1)
    price = rng.uniform(1, 1000, n)
    disc = np.round(rng.uniform(0, 0.1, n), 2)
    qty = rng.integers(1, 50, n).astype(np.int32)
    cols = [both(price, N.F64, rng.random(n) > 0.02), both(disc, N.F64), both(qty, N.I32)]
    p = Program()
    pr, d, q = p.load(0, N.F64), p.load(1, N.F64), p.load(2, N.I32)
    p.store(p.binary("*", pr, d, N.F64), 0)
    c1 = p.binary("<", q, p.const(24, N.I32), N.I32)
    c2 = p.binary(">=", d, p.const(0.05, N.F64), N.F64)
    c3 = p.binary("<=", d, p.const(0.07, N.F64), N.F64)
    c4 = p.binary(">", p.binary("+", pr, p.cast(q, N.F64), N.F64), p.const(100.0, N.F64), N.F64)
    p.store(p.binary("and", p.binary("and", c1, c2, N.BOOL), p.binary("and", c3, c4, N.BOOL), N.BOOL), 1)
    _agg_check(cols, p, [N.F64, N.BOOL], 1, [(SUM, 0), (COUNT_STAR, 0), (MIN, 0), (MAX, 0), (AVG, 0), (COUNT, 0)])


def test_aggregate_expr_integer_and_decimal(gpu):
    n = 1_500_000
    rng = np.random.default_rng(21)
    a = rng.integers(-(1 << 30), 1 << 30, n).astype(np.int64)
    b = rng.integers(-1000, 1000, n).astype(np.int32)
    prices = [int(x) for x in rng.integers(100, 10 ** 6, n)]  # Decimal128(12,2)
    cols = [both(a, N.I64, rng.random(n) > 0.1), both(b, N.I32), both(dec_array(prices), N.DEC128, None, 2)]
    p = Program()
    ra, rb, rp = p.load(0, N.I64), p.load(1, N.I32), p.load(2, N.DEC128)
    p.store(p.binary("*", ra, p.cast(rb, N.I64), N.I64), 0)                      # a * b
    p.store(p.binary("%", p.binary("-", ra, p.cast(rb, N.I64), N.I64), p.const(97, N.I64), N.I64), 1)
    p.store(p.binary("*", rp, p.const(3, N.DEC128), N.DEC128), 2)                 # price * 3
    p.store(p.binary("!=", rb, p.const(0, N.I32), N.I32), 3)                      # WHERE b <> 0
    _agg_check(cols, p, [N.I64, N.I64, N.DEC128, N.BOOL], 3,
               [(SUM, 0), (MIN, 0), (MAX, 0), (SUM, 1), (VAR_POP, 1), (STDDEV_SAMP, 0), (SUM, 2), (COUNT, 0)])


def test_aggregate_expr_no_filter_and_errors(gpu):
    n = 100_000
    rng = np.random.default_rng(5)
    x = rng.standard_normal(n).astype(np.float32)
    cols = [both(x, N.F32, rng.random(n) > 0.5)]
    p = Program()
    rx = p.load(0, N.F32)
    p.store(p.binary("*", rx, rx, N.F32), 0)
    _agg_check(cols, p, [N.F32], -1, [(SUM, 0), (COUNT, 0), (COUNT_STAR, 0), (AVG, 0)])
    with pytest.raises(N.BoltError) as e:  # the filter value must be Bool
        execute_aggregate_expr([cols[0][0]], p, 1, [(SUM, 0)], 0)
    assert e.value.code == N.BOLT_ERR_PLAN
    q = Program()
    q.store(q.binary(">", q.load(0, N.F32), q.const(0.0, N.F32), N.F32), 0)
    with pytest.raises(N.BoltError) as e:  # SUM over a Bool value
        execute_aggregate_expr([cols[0][0]], q, 1, [(SUM, 0)])
    assert e.value.code == N.BOLT_ERR_TYPE
