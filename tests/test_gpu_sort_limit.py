"""GPU parity: ORDER BY + LIMIT / OFFSET (bolt_sort_limit).  The row ids are
bit-exact against the slice of the oracle's stable sort (execute_sort +
execute_limit, src/exec/sort.rs:161, src/exec/limit.rs:18) and, at sizes the
oracle is slow on, of numpy's stable argsort."""
import numpy as np
import pytest

import oracle as O
from craton_bolt_amd import _native as N
from craton_bolt_amd.device import DeviceColumn
from craton_bolt_amd.ops import execute_sort, execute_sort_limit

from gpu_util import both, dec_array

pytestmark = pytest.mark.gpu


def _check(cols, desc, nf, limit, offset):
    dev = [d for d, _ in cols]
    host = [h for _, h in cols]
    got = execute_sort_limit(dev, limit, offset, desc, nf)
    want = O.sort_limit(host, desc, nf, limit, offset)
    assert np.array_equal(got, want), (limit, offset, got[:8], want[:8])


@pytest.mark.parametrize("n", [70_000, 1_500_000])
@pytest.mark.parametrize("limit,offset", [(1, 0), (10, 0), (100, 37), (5000, 1000), (0, 0), (10, 10 ** 9)])
@pytest.mark.parametrize("desc", [False, True])
def test_sort_limit_int64_matches_oracle(gpu, n, limit, offset, desc):
    print(f"[case] sort_limit i64 n={n} limit={limit} offset={offset} desc={desc}", flush=True)
    rng = np.random.default_rng(n + limit)
    cols = [both(rng.integers(-(1 << 62), 1 << 62, n).astype(np.int64), N.I64)]
    _check(cols, [desc], [True], limit, offset)


@pytest.mark.parametrize("dtype", ["i32", "f64", "f32", "date32", "dec128"])
@pytest.mark.parametrize("nulls_first", [True, False])
@pytest.mark.parametrize("desc", [False, True])
def test_sort_limit_dtypes_with_nulls(gpu, dtype, nulls_first, desc):
    print(f"[case] sort_limit {dtype} nulls_first={nulls_first} desc={desc}", flush=True)
    n = 1_200_000
    rng = np.random.default_rng(7)
    valid = rng.random(n) > 0.00025  # about 300 NULLs: a limit of 1000 reaches past them when they come first
    if dtype == "i32":
        cols = [both(rng.integers(-10 ** 9, 10 ** 9, n).astype(np.int32), N.I32, valid)]
    elif dtype == "date32":
        cols = [both(rng.integers(0, 40_000, n).astype(np.int32), N.DATE32, valid)]
    elif dtype == "f64":
        v = rng.standard_normal(n) * 1e6
        v[rng.integers(0, n, 50)] = np.nan
        v[rng.integers(0, n, 50)] = -0.0
        v[rng.integers(0, n, 50)] = 0.0
        v[rng.integers(0, n, 20)] = np.inf
        v[rng.integers(0, n, 20)] = -np.inf
        cols = [both(v, N.F64, valid)]
    elif dtype == "f32":
        v = (rng.standard_normal(n) * 1e3).astype(np.float32)
        v[rng.integers(0, n, 50)] = np.nan
        v[rng.integers(0, n, 50)] = -0.0
        cols = [both(v, N.F32, valid)]
    else:
        vals = [int(x) * (1 << 70) + int(y) for x, y in zip(rng.integers(-50, 50, n), rng.integers(0, 1 << 62, n))]
        cols = [both(dec_array(vals), N.DEC128, valid, scale=2)]
    for limit, offset in ((1, 0), (1000, 0), (200, 250), (50, n - 20)):
        _check(cols, [desc], [nulls_first], limit, offset)


def test_sort_limit_multi_key_ties(gpu):
    """A leading key with few distinct values leaves large tie buckets; the
    second key and then the row id order them (the full sort is stable)."""
    n = 2_000_000
    rng = np.random.default_rng(11)
    k0 = rng.integers(0, 1000, n).astype(np.int32)
    k1 = rng.integers(0, 5, n).astype(np.int64)
    v1 = rng.random(n) > 0.1
    cols = [both(k0, N.I32), both(k1, N.I64, v1)]
    for limit, offset in ((10, 0), (3000, 0), (777, 4000)):
        for desc in ([False, True], [True, False]):
            print(f"[case] multi-key ties limit={limit} offset={offset} desc={desc}", flush=True)
            _check(cols, desc, [True, False], limit, offset)


def test_sort_limit_leading_key_constant_falls_back(gpu):
    """Every row ties on the leading key: the select cannot narrow the rows
    and the full sort runs; the result is the same slice."""
    n = 1_200_000
    rng = np.random.default_rng(3)
    cols = [both(np.ones(n, np.uint8), N.BOOL), both(rng.integers(0, 100, n).astype(np.int32), N.I32)]
    _check(cols, [False, True], [True, True], 100, 5)


def test_sort_limit_large_against_numpy(gpu):
    n = 32 << 20
    rng = np.random.default_rng(5)
    key = rng.integers(-(1 << 40), 1 << 40, n).astype(np.int64)
    dev = DeviceColumn.from_numpy(key, N.I64)
    for limit, offset, desc in ((100, 0, False), (10_000, 123, True)):
        print(f"[case] sort_limit 32M limit={limit} offset={offset} desc={desc}", flush=True)
        got = execute_sort_limit([dev], limit, offset, [desc])
        want = np.argsort(-key if desc else key, kind="stable")[offset:offset + limit]
        assert np.array_equal(got, want.astype(np.uint32))


def test_sort_limit_equals_full_sort_slice(gpu):
    """bolt_sort_limit equals the slice of bolt_sort_indices on the same keys."""
    n = 3_000_000
    rng = np.random.default_rng(9)
    dev = DeviceColumn.from_numpy(rng.integers(0, 50_000, n).astype(np.int64), N.I64)
    full = execute_sort([dev])
    assert np.array_equal(execute_sort_limit([dev], 2500, 100), full[100:2600])


@pytest.mark.parametrize("dtype", ["i64", "i32", "f64", "f32", "date32"])
@pytest.mark.parametrize("limit,offset,desc", [(1, 0, False), (100, 0, True), (2000, 48, False), (30, 7, True)])
def test_sort_limit_small_final_set(gpu, dtype, limit, offset, desc):
    """One key without NULLs: the candidates are narrowed and the last few
    sorted in one workgroup; heavy duplicates put ties across the window's
    edge, which the row id must order."""
    print(f"[case] small final set {dtype} limit={limit} offset={offset} desc={desc}", flush=True)
    n = 3_000_000
    rng = np.random.default_rng(len(dtype) + limit)
    if dtype == "i64":
        cols = [both(rng.integers(-(1 << 62), 1 << 62, n).astype(np.int64), N.I64)]
    elif dtype == "i32":
        cols = [both(rng.integers(0, 5000, n).astype(np.int32), N.I32)]  # about 600 copies of each value
    elif dtype == "date32":
        cols = [both(rng.integers(0, 40_000, n).astype(np.int32), N.DATE32)]
    elif dtype == "f64":
        v = np.round(rng.standard_normal(n), 3)
        v[rng.integers(0, n, 40)] = np.nan
        v[rng.integers(0, n, 40)] = -0.0
        cols = [both(v, N.F64)]
    else:
        cols = [both(rng.standard_normal(n).astype(np.float32), N.F32)]
    _check(cols, [desc], [True], limit, offset)


@pytest.mark.parametrize("limit,offset,desc", [(100, 0, False), (100, 0, True), (1, 0, False), (300, 200, True),
                                               (2000, 0, False)])
def test_sort_limit_first_pass_candidates(gpu, limit, offset, desc):
    """Int64 keys over the whole range: the K-th row's top digit is within the
    first pass's guess, so the candidates come from that pass (no collect pass)."""
    print(f"[case] first-pass candidates limit={limit} offset={offset} desc={desc}", flush=True)
    n = 4_000_000
    rng = np.random.default_rng(limit + offset)
    keys = rng.integers(np.iinfo(np.int64).min, np.iinfo(np.int64).max, n, dtype=np.int64)
    keys[rng.integers(0, n, 20)] = keys[0]  # a few ties
    _check([both(keys, N.I64)], [desc], [True], limit, offset)


def test_sort_limit_first_pass_candidates_two_keys(gpu):
    """The same guess with a second key: the general candidate sort over the guessed superset."""
    n = 3_000_000
    rng = np.random.default_rng(4)
    k0 = rng.integers(np.iinfo(np.int64).min, np.iinfo(np.int64).max, n, dtype=np.int64)
    k0[rng.integers(0, n, 50)] = np.iinfo(np.int64).min  # ties at the very front
    k1 = rng.integers(0, 3, n).astype(np.int32)
    _check([both(k0, N.I64), both(k1, N.I32, rng.random(n) > 0.1)], [False, True], [True, False], 120, 10)


@pytest.mark.parametrize("dtype", ["i64", "i32", "date32"])
@pytest.mark.parametrize("limit,offset,desc", [(100, 0, False), (100, 0, True), (1, 0, False), (1500, 300, True),
                                               (2047, 0, False)])
def test_sort_limit_large_tie_on_the_kth_key(gpu, dtype, limit, offset, desc):
    """More rows tie on the K-th row's key than one workgroup sorts (about
    20,000 copies of each value): the tied rows are narrowed by row id, and the
    window must still be the stable order's."""
    print(f"[case] large tie {dtype} limit={limit} offset={offset} desc={desc}", flush=True)
    n = 2_000_003
    rng = np.random.default_rng(7 * limit + offset + len(dtype))
    v = rng.integers(-50, 50, n)
    if dtype == "i64":
        cols = [both((v * 1_000_000_007).astype(np.int64), N.I64)]
    elif dtype == "i32":
        cols = [both(v.astype(np.int32), N.I32)]
    else:
        cols = [both((v + 19_000).astype(np.int32), N.DATE32)]
    _check(cols, [desc], [True], limit, offset)
