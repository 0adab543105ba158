"""GPU: every operator accepts zero-length inputs and returns an empty (or
SQL-NULL) result, as the reference's executors do for empty batches."""
import ctypes as C

import numpy as np
import pytest

from craton_bolt_amd import _native as N
from craton_bolt_amd.device import DeviceBuffer, DeviceColumn
from craton_bolt_amd.ops import (JoinTable, Term, compact_columns, execute_aggregate, execute_groupby,
                                 execute_inner_join, execute_sort, take)

pytestmark = pytest.mark.gpu


def _empty(dt, valid=False):
    a = np.zeros(0, dtype={N.I32: np.int32, N.I64: np.int64, N.F64: np.float64}[dt])
    return DeviceColumn.from_numpy(a, dt, valid=np.zeros(0, bool) if valid else None)


def test_empty_aggregate_is_null_or_zero(gpu):
    c = _empty(N.F64)
    got = execute_aggregate([c], [(N.SUM, 0), (N.COUNT_STAR, 0), (N.MIN, 0)], [Term(0, N.GT, 1.0)])
    assert got == [None, 0, None]


def test_empty_compaction_take_and_sort(gpu):
    c = _empty(N.I64, valid=True)
    outs, k = compact_columns([c], [Term(0, N.GT, 0)])
    assert k == 0
    t = take(c, DeviceBuffer(4), 0)
    assert t.length == 0
    assert len(execute_sort([c], [True], [False])) == 0


@pytest.mark.parametrize("flags", [0, N.GROUPBY_FORCE_GLOBAL, N.GROUPBY_FORCE_PARTITIONED])
def test_empty_groupby(gpu, flags):
    r = execute_groupby([_empty(N.I32)], [_empty(N.I64)], [(N.SUM, 0), (N.COUNT_STAR, 0)], flags=flags)
    assert r["n_groups"] == 0
    assert all(len(v) == 0 for v, _ in r["keys"] + r["aggs"])


def test_empty_join_sides(gpu):
    e, x = _empty(N.I64), DeviceColumn.from_numpy(np.arange(10, dtype=np.int64))
    for a, b in [(e, x), (x, e), (e, e)]:
        l, r = execute_inner_join(a, b)
        assert len(l) == 0 and len(r) == 0
    t = JoinTable(e)
    pr, br = t.probe([x], 0)
    assert len(pr) == 0
    assert t.probe_aggregate([x], 0, [(N.COUNT_STAR, 0), (N.SUM, 0)]) == [0, None]
