"""GPU: bolt_aggregate_noalloc, the allocation-free form of bolt_aggregate,
called through the C ABI and checked against bolt_aggregate on the same
inputs: two calls in a row on each case, the second with a smaller input
(the shared scratch array keeps the first call's contents), results equal
(integer, Decimal128 and count results bit-exact; float results within the
relative tolerance of the compensated sums, since the two entry points may
run different kernels), the same errors, and no device memory left held."""
import ctypes as C

import numpy as np
import pytest

from craton_bolt_amd import _native as N
from craton_bolt_amd.device import DeviceColumn
from craton_bolt_amd.ops import Term, _aggs, _filter, columns_array, scalar_value

from gpu_util import dec_array

pytestmark = pytest.mark.gpu

REL_TOL = 1e-12


def _call(fn, cols, aggs, terms):
    ca = columns_array(cols)
    f, keep = _filter(terms, cols)
    res = (N.BoltScalar * max(len(aggs), 1))()
    st = fn(None, ca, len(cols), _aggs(aggs), len(aggs), C.byref(f) if f else None, res)
    return st, [scalar_value(res[i]) for i in range(len(aggs))]


def _same(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        if isinstance(y, float) and isinstance(x, float):
            assert x == y or abs(x - y) <= REL_TOL * max(abs(y), 1e-300), (a, b)
        else:
            assert x == y, (a, b)


def _pool_in_use():
    used, peak = C.c_size_t(), C.c_size_t()
    assert N.lib().bolt_hip_pool_stats(C.byref(used), C.byref(peak)) == 0
    return used.value


def _columns(n, seed):
    rng = np.random.default_rng(seed)
    i64 = rng.integers(-10 ** 12, 10 ** 12, n).astype(np.int64)
    i32 = rng.integers(-1000, 1000, n).astype(np.int32)
    f64 = rng.normal(size=n) * 1e3
    f32 = rng.normal(size=n).astype(np.float32)
    vv = rng.random(n) > 0.1
    dec = dec_array([int(x) * 10 ** 12 + int(y) for x, y in zip(rng.integers(-10 ** 9, 10 ** 9, n),
                                                                 rng.integers(0, 10 ** 12, n))])
    return [DeviceColumn.from_numpy(i64), DeviceColumn.from_numpy(i32, valid=vv), DeviceColumn.from_numpy(f64),
            DeviceColumn.from_numpy(f32), DeviceColumn.from_numpy(dec, N.DEC128)]


CASES = [
    ("sum_count", [(N.SUM, 0), (N.COUNT_STAR, 0)], None),
    ("many", [(N.SUM, 0), (N.SUM, 1), (N.SUM, 2), (N.SUM, 3), (N.SUM, 4), (N.COUNT, 1), (N.MIN, 2), (N.MAX, 0),
              (N.AVG, 2), (N.VAR_SAMP, 2), (N.STDDEV_POP, 3), (N.MIN, 1), (N.MAX, 3), (N.COUNT_STAR, 0)],
     [Term(2, N.GT, -500.0), Term(1, N.NE, 7)]),
    ("sixteen", [(N.SUM, 1)] * 8 + [(N.MAX, 2)] * 8, [Term(0, N.LT, 0)]),
]


@pytest.mark.parametrize("name,aggs,terms", CASES)
def test_noalloc_equals_bolt_aggregate_twice(gpu, name, aggs, terms):
    L = N.lib()
    for n, seed in ((3_000_001, 1), (70_000, 2)):  # the second call is smaller
        print(f"[case] noalloc {name} n={n}", flush=True)
        cols = _columns(n, seed)
        st_a, want = _call(L.bolt_aggregate, cols, aggs, terms)
        assert st_a == 0, N.lib().bolt_last_error()
        before = _pool_in_use()
        st_b, got = _call(L.bolt_aggregate_noalloc, cols, aggs, terms)
        assert st_b == 0, N.lib().bolt_last_error()
        assert _pool_in_use() == before
        _same(got, want)


def test_noalloc_errors_like_bolt_aggregate(gpu):
    L = N.lib()
    big = DeviceColumn.from_numpy(np.array([2 ** 63 - 1, 1, -1], np.int64))
    st_a, _ = _call(L.bolt_aggregate, [big], [(N.SUM, 0)], None)
    st_b, _ = _call(L.bolt_aggregate_noalloc, [big], [(N.SUM, 0)], None)
    assert st_a == st_b == N.BOLT_ERR_TYPE
    assert b"SUM(integer) overflow" in L.bolt_last_error()
    # then a normal call still works (the scratch keeps nothing that matters)
    ok = DeviceColumn.from_numpy(np.arange(1000, dtype=np.int64))
    st, got = _call(L.bolt_aggregate_noalloc, [ok], [(N.SUM, 0), (N.COUNT_STAR, 0)], None)
    assert st == 0 and got == [499500, 1000]
    # empty input and too many aggregates
    empty = DeviceColumn.from_numpy(np.zeros(0, np.int64))
    st, got = _call(L.bolt_aggregate_noalloc, [empty], [(N.SUM, 0), (N.COUNT_STAR, 0)], None)
    assert st == 0 and got == [None, 0]
    st, _ = _call(L.bolt_aggregate_noalloc, [ok], [(N.SUM, 0)] * 17, None)
    assert st == N.BOLT_ERR_PLAN
