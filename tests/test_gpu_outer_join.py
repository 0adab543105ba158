"""GPU parity: equi-joins of every kind (bolt_join) and the NULL-index take.

Row sets are bit-exact against the oracle's restatement of the reference's
host join (src/exec/join.rs:132), compared after canonical ordering; the GPU's
own order is checked separately: probe rows ascending with each NULL-padded
pair in its place, then (FULL) the unmatched build rows ascending."""
import numpy as np
import pytest

import oracle as O
from craton_bolt_amd import _native as N
from craton_bolt_amd.device import DeviceBuffer, DeviceColumn
from craton_bolt_amd.ops import NULL_ROW, execute_join, take

from gpu_util import both

pytestmark = pytest.mark.gpu
HOWS = ["inner", "left", "right", "full"]


def _canon(a, b):
    order = np.lexsort((b, a))
    return np.stack([a[order], b[order]])


def _check_order(l, r, how, nl, nr):
    """The GPU's emission order (bolt_join in include/bolt_ops.h)."""
    if how == "inner":
        build_is_left = nl <= nr
    else:
        build_is_left = how == "right"
    probe, build = (r, l) if build_is_left else (l, r)
    tail = int((probe == NULL_ROW).sum()) if how == "full" else 0
    head = len(probe) - tail
    assert np.all(np.diff(probe[:head].astype(np.int64)) >= 0)
    if tail:
        assert np.all(probe[head:] == NULL_ROW)
        assert np.all(np.diff(build[head:].astype(np.int64)) > 0)
    if how != "inner":  # every probe row appears
        assert np.array_equal(np.unique(probe[:head]), np.arange(nr if build_is_left else nl))


def _run(lcols, rcols, how, hl, hr):
    l, r = execute_join(lcols, rcols, how)
    ol, orr = O.join(hl, hr, how)
    assert len(l) == len(ol), (how, len(l), len(ol))
    assert np.array_equal(_canon(l, r), _canon(ol, orr))
    _check_order(l, r, how, lcols[0].length, rcols[0].length)
    return l, r


def _fixture_sides():
    # tests/gpu_join_e2e.rs:378-487: left keys 0..4096, right keys 2048..6144
    dl, hl = both(np.arange(4096, dtype=np.int32), N.I32)
    dr, hr = both(np.arange(2048, 2048 + 4096, dtype=np.int32), N.I32)
    return dl, hl, dr, hr


@pytest.mark.parametrize("how,rows,lnull,rnull", [("left", 4096, 0, 2048), ("right", 4096, 2048, 0),
                                                  ("full", 6144, 2048, 2048), ("inner", 2048, 0, 0)])
def test_outer_join_fixtures(gpu, how, rows, lnull, rnull):
    # tests/gpu_join_e2e.rs:378-487 row and NULL counts
    print(f"[case] outer fixture {how}", flush=True)
    dl, hl, dr, hr = _fixture_sides()
    l, r = _run([dl], [dr], how, [hl], [hr])
    assert len(l) == rows and (l == NULL_ROW).sum() == lnull and (r == NULL_ROW).sum() == rnull


def test_join_duplicate_keys_fixture(gpu):
    # tests/gpu_join_e2e.rs:491-524: 4 x 8 matches per key, 1024 keys
    db, hb = both((np.arange(4096) % 1024).astype(np.int32), N.I32)
    dp, hp = both((np.arange(8192) % 1024).astype(np.int32), N.I32)
    l, _ = _run([db], [dp], "inner", [hb], [hp])
    assert len(l) == 32_768


def test_join_float_nan_fixture(gpu):
    # tests/gpu_join_e2e.rs:528-597: non-canonical NaN probes meet the NaN build row
    bk = np.arange(1, 4097, dtype=np.float64)
    bk[-1] = np.nan
    pk = np.arange(1, 8193, dtype=np.float64)
    pk.view(np.uint64)[4095:] = np.uint64(0x7FF8000000000001)
    db, hb = both(bk, N.F64)
    dp, hp = both(pk, N.F64)
    l, _ = _run([db], [dp], "inner", [hb], [hp])
    assert len(l) == 8192


def test_join_two_key_fixture(gpu):
    # tests/gpu_join_e2e.rs:320-374
    i = np.arange(4096, dtype=np.int32)
    j = np.arange(8192, dtype=np.int32)
    pa = np.where(j < 4096, j // 64, 1000 + j // 64).astype(np.int32)
    (da, ha), (db_, hb_) = both(i // 64, N.I32), both(i % 64, N.I32)
    (pa_d, pa_h), (pb_d, pb_h) = both(pa, N.I32), both(j % 64, N.I32)
    l, r = _run([da, db_], [pa_d, pb_d], "inner", [ha, hb_], [pa_h, pb_h])
    assert len(l) == 4096


def _key_cols(rng, kind, n, null_frac):
    valid = rng.random(n) >= null_frac if null_frac else None
    if kind == "i64":
        return [both(rng.integers(0, max(n // 3, 1), n).astype(np.int64) * 7919 - 11, N.I64, valid)]
    if kind == "i32":
        return [both(rng.integers(-(n // 4) - 1, n // 4 + 1, n).astype(np.int32), N.I32, valid)]
    if kind == "date32":
        return [both(rng.integers(18000, 18000 + max(n // 2, 1), n).astype(np.int32), N.DATE32, valid)]
    # a few rows of each special float value (every zero meets every zero and
    # every NaN every NaN, so a fixed fraction of them would square the output)
    def few():
        return rng.integers(0, n, min(n, 12))
    if kind == "f64":
        v = rng.integers(0, max(n // 3, 1), n).astype(np.float64) - 0.5
        v[few()] = -0.0
        v[few()] = 0.0
        v[few()] = np.nan
        v.view(np.uint64)[few()] = np.uint64(0xFFF8000000000007)
        return [both(v, N.F64, valid)]
    if kind == "f32":
        v = rng.integers(0, max(n // 3, 1), n).astype(np.float32) * np.float32(0.25)
        v[few()] = np.float32(-0.0)
        v[few()] = np.float32(0.0)
        v.view(np.uint32)[few()] = np.uint32(0x7FC00005)
        return [both(v, N.F32, valid)]
    if kind == "bool":
        return [both(rng.integers(0, 2, n).astype(np.uint8), N.BOOL, valid)]
    v2 = rng.random(n) >= null_frac if null_frac else None
    if kind == "two_i64":  # hashed and checked on the device (TwoI64)
        return [both(rng.integers(-3, 4, n).astype(np.int64) << 36, N.I64, valid),
                both(rng.integers(0, max(n // 10, 1), n).astype(np.int64) - (1 << 40), N.I64, v2)]
    if kind == "three_mixed":  # MultiI32-like: three columns, hashed
        return [both(rng.integers(0, 3, n).astype(np.int32), N.I32, valid),
                both(rng.integers(0, max(n // 20, 1), n).astype(np.int64), N.I64, v2),
                both(rng.integers(18000, 18004, n).astype(np.int32), N.DATE32)]
    # two Int32 columns (the second with its own NULLs)
    return [both(rng.integers(-3, 4, n).astype(np.int32), N.I32, valid),
            both(rng.integers(0, max(n // 10, 1), n).astype(np.int32), N.I32, v2)]


@pytest.mark.parametrize("how", HOWS)
@pytest.mark.parametrize("kind", ["i64", "i32", "date32", "f64", "f32", "two_i32", "two_i64", "three_mixed"])
@pytest.mark.parametrize("nl,nr,null_frac", [(1, 1, 0.0), (700, 300, 0.1), (40_000, 90_000, 0.05),
                                             (300_000, 120_000, 0.0)])
def test_join_kinds_match_oracle(gpu, how, kind, nl, nr, null_frac):
    print(f"[case] bolt_join {how} {kind} {nl}x{nr} nulls={null_frac}", flush=True)
    rng = np.random.default_rng(nl * 31 + nr + len(kind))
    lc = _key_cols(rng, kind, nl, null_frac)
    rc = _key_cols(rng, kind, nr, null_frac)
    _run([d for d, _ in lc], [d for d, _ in rc], how, [h for _, h in lc], [h for _, h in rc])


@pytest.mark.parametrize("how", HOWS)
def test_join_bool_keys(gpu, how):
    rng = np.random.default_rng(5)
    lc = _key_cols(rng, "bool", 50, 0.2)
    rc = _key_cols(rng, "bool", 60, 0.2)
    _run([d for d, _ in lc], [d for d, _ in rc], how, [h for _, h in lc], [h for _, h in rc])


@pytest.mark.parametrize("how", HOWS)
def test_join_empty_sides(gpu, how):
    # src/exec/join.rs:320-360: one or both sides empty
    de, he = both(np.zeros(0, np.int64), N.I64)
    dx, hx = both(np.array([5, 6, 5, 9], np.int64), N.I64)
    for a, ha, b, hb in ((dx, hx, de, he), (de, he, dx, hx), (de, he, de, he)):
        _run([a], [b], how, [ha], [hb])


@pytest.mark.parametrize("how", HOWS)
@pytest.mark.parametrize("kind", ["two_i64", "three_mixed"])
def test_join_hashed_keys_checked_under_collisions(gpu, monkeypatch, how, kind):
    """With the key hash cut to 3 bits nearly every hash match is a different
    tuple: the result stays exact only if every match is checked against the
    key columns."""
    monkeypatch.setenv("BOLT_JOIN_HASH_BITS", "3")
    print(f"[case] hashed keys, 3-bit hash, {how} {kind}", flush=True)
    rng = np.random.default_rng(77)
    lc = _key_cols(rng, kind, 1500, 0.05)
    rc = _key_cols(rng, kind, 2500, 0.05)
    _run([d for d, _ in lc], [d for d, _ in rc], how, [h for _, h in lc], [h for _, h in rc])


def test_join_lossy_two_i64_fixtures(gpu):
    # tests/gpu_join_e2e.rs:762-832 (INNER) and :1063-1162 (LEFT) over two Int64 keys
    i = np.arange(4096, dtype=np.int64)
    j = np.arange(8192, dtype=np.int64)
    pa = np.where(j < 4096, j // 2, 99_999 + j)
    pb = np.where(j < 4096, j * 3 + 1, 88_888 + j)
    b = [both(i // 2, N.I64), both(i * 3 + 1, N.I64)]
    p = [both(pa, N.I64), both(pb, N.I64)]
    l, _ = _run([d for d, _ in b], [d for d, _ in p], "inner", [h for _, h in b], [h for _, h in p])
    assert len(l) == 4096
    n = np.arange(4096, dtype=np.int64)
    m = np.arange(2048, dtype=np.int64)
    lt = [both(n, N.I64), both(1_000_000 + n, N.I64)]
    rt = [both(m, N.I64), both(1_000_000 + m, N.I64)]
    l, r = _run([d for d, _ in lt], [d for d, _ in rt], "left", [h for _, h in lt], [h for _, h in rt])
    assert len(l) == 4096 and (r != NULL_ROW).sum() == 2048


def test_join_capacity_declines(gpu):
    db, _ = both((np.arange(1000) % 10).astype(np.int64), N.I64)
    dp, _ = both((np.arange(3000) % 10).astype(np.int64), N.I64)
    with pytest.raises(N.BoltError) as e:
        execute_join([db], [dp], "full", capacity=1000)
    assert e.value.code == N.BOLT_DECLINE
    l, r = execute_join([db], [dp], "full")  # sized from the declined count
    assert len(l) == 300_000


def test_join_type_errors(gpu):
    a, _ = both(np.arange(10, dtype=np.int32), N.I32)
    b, _ = both(np.arange(10, dtype=np.int64), N.I64)
    with pytest.raises(N.BoltError) as e:
        execute_join([a], [b], "inner")
    assert e.value.code == N.BOLT_ERR_TYPE
    c, _ = both(np.arange(10, dtype=np.float64), N.F64)
    with pytest.raises(N.BoltError) as e:  # multi-column keys are integer columns
        execute_join([c, c], [c, c], "inner")
    assert e.value.code == N.BOLT_ERR_UNSUPPORTED
    d, _ = both(np.arange(10, dtype=np.int64), N.I64)
    with pytest.raises(N.BoltError) as e:
        execute_join([d] * 5, [d] * 5, "inner")
    assert e.value.code == N.BOLT_ERR_UNSUPPORTED


def test_outer_join_materialise_with_null_take(gpu):
    # execute_outer_join's materialise (src/exec/join.rs:754): take with NULL indices
    dl, hl, dr, hr = _fixture_sides()
    lv = np.arange(4096, dtype=np.int32) + 100
    rv = np.arange(2048, 2048 + 4096, dtype=np.int64) + 500
    l, r = execute_join([dl], [dr], "full")
    lrows = DeviceBuffer.from_numpy(l)
    rrows = DeviceBuffer.from_numpy(r)
    lcol = take(DeviceColumn.from_numpy(lv, N.I32), lrows, len(l), nullable=True)
    rcol = take(DeviceColumn.from_numpy(rv, N.I64), rrows, len(r), nullable=True)
    lval, lok = lcol.to_numpy(), lcol.valid_numpy()
    rval, rok = rcol.to_numpy(), rcol.valid_numpy()
    assert np.array_equal(lok, l != NULL_ROW) and np.array_equal(rok, r != NULL_ROW)
    assert np.array_equal(lval[lok], lv[l[lok]]) and np.array_equal(rval[rok], rv[r[rok]])
    assert lok.sum() == 4096 and rok.sum() == 4096
    with pytest.raises(N.BoltError) as e:  # a NULL index needs a destination bitmap
        take(DeviceColumn.from_numpy(lv, N.I32), lrows, len(l))
    assert e.value.code == N.BOLT_ERR_PLAN
