"""GPU parity: LSD radix sort.  Sorted keys and permutations are bit-exact
against a stable reference order (numpy's stable argsort, and the oracle's
stable restatement of the ORDER BY key transforms, src/exec/gpu_sort.rs:
1070-1180); the reference's arrow lexsort is unstable, so ties are compared in
the canonical (key, row id) order."""
import numpy as np
import pytest

import oracle as O
from craton_bolt_amd import _native as N
from craton_bolt_amd import datagen as D
from craton_bolt_amd.device import DeviceBuffer
from craton_bolt_amd.ops import execute_sort, sort_pairs_i64

from gpu_util import both, dec_array

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, params=["hybrid", "lsd"])
def sort_path(request, monkeypatch):
    """Large sorts run the hybrid (two global passes + LDS bucket sorts) and,
    with BOLT_SORT_LSD set, the plain LSD sort."""
    if request.param == "lsd":
        monkeypatch.setenv("BOLT_SORT_LSD", "1")
    else:
        monkeypatch.delenv("BOLT_SORT_LSD", raising=False)
    return request.param


def _check_pairs(key, rid, desc=False):
    n = len(key)
    ko, vo = sort_pairs_i64(DeviceBuffer.from_numpy(key), DeviceBuffer.from_numpy(rid), n, desc)
    order = np.argsort(~key if desc else key, kind="stable")
    assert np.array_equal(ko.to_numpy(np.int64, n), key[order])
    assert np.array_equal(vo.to_numpy(np.uint32, n), rid[order])


def test_sort_hybrid_shapes(gpu):
    """Key sets that exercise the hybrid's choices: 40-bit keys (top varying
    digits are bytes 3 and 4), a skewed set whose buckets overflow LDS (LSD
    fallback), heavy duplicates, and two varying bytes only."""
    rng = np.random.default_rng(21)
    n = 3_000_000
    rid = np.arange(n, dtype=np.uint32)
    print("[case] 40-bit keys", flush=True)
    _check_pairs(rng.integers(0, 2 ** 40, n).astype(np.int64), rid)
    print("[case] skewed: two top-16 values", flush=True)
    k = rng.integers(0, 2 ** 48, n).astype(np.int64) | (rng.integers(0, 2, n).astype(np.int64) << 56)
    _check_pairs(k, rid)
    print("[case] duplicates", flush=True)
    _check_pairs((rng.integers(0, 5000, n).astype(np.int64) << 20) - (1 << 40), rid, desc=True)
    print("[case] two varying bytes", flush=True)
    _check_pairs(rng.integers(0, 65536, n).astype(np.int64), rid)


@pytest.mark.parametrize("n", [1, 2, 1000, 4096, 4097, 1_000_003, 20_000_000])
@pytest.mark.parametrize("desc", [False, True])
def test_sort_pairs_i64(gpu, n, desc):
    print(f"[case] sort_pairs n={n} desc={desc}", flush=True)
    key, rid = D.sort_pairs(n)
    ko, vo = sort_pairs_i64(DeviceBuffer.from_numpy(key), DeviceBuffer.from_numpy(rid), n, desc)
    order = np.argsort(-key if desc else key, kind="stable") if not desc else np.argsort(~key, kind="stable")
    assert np.array_equal(ko.to_numpy(np.int64, n), key[order])
    assert np.array_equal(vo.to_numpy(np.uint32, n), rid[order])


def test_sort_pairs_stability_with_duplicates(gpu):
    n = 3_000_000
    rng = np.random.default_rng(1)
    key = rng.integers(-5, 5, n).astype(np.int64) * (2 ** 40)
    rid = np.arange(n, dtype=np.uint32)
    ko, vo = sort_pairs_i64(DeviceBuffer.from_numpy(key), DeviceBuffer.from_numpy(rid), n)
    order = np.argsort(key, kind="stable")
    assert np.array_equal(vo.to_numpy(np.uint32, n), order.astype(np.uint32))


@pytest.mark.parametrize("n", [1, 5000, 700_001])
def test_order_by_matches_oracle(gpu, n):
    rng = np.random.default_rng(n)
    a = rng.integers(-3, 3, n).astype(np.int32)
    b = rng.normal(size=n)
    b[::13] = np.nan
    b[::17] = -0.0
    c = rng.integers(-2 ** 62, 2 ** 62, n).astype(np.int64)
    d = (rng.integers(0, 400, n)).astype(np.int32)
    va = rng.random(n) > 0.05
    vb = rng.random(n) > 0.05
    cols = [both(a, N.I32, va), both(b, N.F64, vb), both(c, N.I64, None), both(d, N.DATE32, None)]
    for desc, nf in [([False, True, False, True], [True, False, True, True]),
                     ([True, False, True, False], [False, True, False, False])]:
        got = execute_sort([x for x, _ in cols], desc, nf)
        want = O.sort_indices([h for _, h in cols], desc, nf)
        assert np.array_equal(got, want)


def test_order_by_float32_single_key(gpu):
    x = np.array([np.nan, 1.0, -0.0, 0.0, -np.inf, np.inf, -np.nan, -1.0], np.float32)
    d, h = both(x, N.F32, None)
    assert execute_sort([d], [False]).tolist() == [4, 7, 2, 3, 1, 5, 0, 6]
    assert execute_sort([d], [True]).tolist() == [0, 6, 5, 1, 2, 3, 7, 4]


@pytest.mark.parametrize("n", [1000, 70_000])
def test_order_by_decimal128_keys(gpu, n):
    """Decimal128 sort keys (signed 128-bit order) with NULLs, DESC and ASC,
    ahead of an int32 tie-breaker, against the oracle's stable order."""
    print(f"[case] order by decimal n={n}", flush=True)
    rng = np.random.default_rng(n + 5)
    base = [int(x) for x in rng.integers(-2 ** 40, 2 ** 40, n)]
    vals = [v * (1 << 60) + int(rng.integers(0, 4)) if i % 3 else v for i, v in enumerate(base)]
    vd = rng.random(n) > 0.05
    t = rng.integers(0, 50, n).astype(np.int32)
    cols = [both(dec_array(vals), N.DEC128, vd), both(t, N.I32, None)]
    for desc, nf in [([True, False], [False, True]), ([False, False], [True, True])]:
        got = execute_sort([x for x, _ in cols], desc, nf)
        want = O.sort_indices([h for _, h in cols], desc, nf)
        assert np.array_equal(got, want)
