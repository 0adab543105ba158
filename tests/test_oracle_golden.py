"""Pins the CPU oracle before it is trusted as the checker.

Each oracle function is checked against something other than itself:
known-answer fixtures held in the reference's own tests (cited file:line),
closed forms, an independent library routine (numpy / pandas), or brute force
on tiny inputs — chosen so that a dropped term, a wrong sign or index, or a
swapped operand fails at least one check.
"""
import math

import numpy as np
import pandas as pd
import pytest

import oracle as O

M64 = (1 << 64) - 1


# ---- Xorshift64Star (tests/common/mod.rs:46-90) -------------------------------

def _xs_reference_python(seed, n):
    """Independent restatement on Python ints with explicit 64-bit masking."""
    s = (seed + 0x9E3779B97F4A7C15) & M64
    if s == 0:
        s = 0xDEADBEEFCAFEBABE
    out = []
    for _ in range(n):
        s ^= s >> 12
        s ^= (s << 25) & M64
        s ^= s >> 27
        out.append((s * 0x2545F4914F6CDD1D) & M64)
    return out


def test_rng_matches_masked_python():
    for seed in (0, 1, 0xB0170001, M64, (1 << 63) + 12345):
        got = [int(x) for x in O.rng_draws(seed, 64)]
        assert got == _xs_reference_python(seed, 64)


def test_rng_zero_guard():
    # seed that mixes to 0 takes the 0xDEADBEEFCAFEBABE state (mod.rs:59-61)
    seed = (-0x9E3779B97F4A7C15) & M64
    assert O.lib().bo_rng_seed(seed) == 0xDEADBEEFCAFEBABE


def test_rng_unit_is_53_bit_fraction():
    u = O.rng_unit(7, 1000)
    d = O.rng_draws(7, 1000)
    assert np.array_equal(u, (d >> np.uint64(11)).astype(np.float64) / 2.0 ** 53)
    assert u.min() >= 0.0 and u.max() < 1.0


def test_shuffle_is_knuth_lcg_fisher_yates():
    xs = np.arange(50, dtype=np.uint32)
    got = O.shuffle_u32(xs, 42)
    # restatement with the LCG constants of mod.rs:106-107
    ref = list(range(50))
    s = 42
    for i in range(49, 0, -1):
        s = (s * 6364136223846793005 + 1442695040888963407) & M64
        j = s % (i + 1)
        ref[i], ref[j] = ref[j], ref[i]
    assert got.tolist() == ref
    assert sorted(got.tolist()) == list(range(50))


# ---- Neumaier (src/exec/aggregate.rs:2001) -------------------------------------

def test_neumaier_recovers_cancelled_terms():
    # Neumaier's own example: naive summation returns 0.0
    assert O.neumaier_sum(np.array([1.0, 1e100, 1.0, -1e100])) == 2.0
    assert O.neumaier_sum(np.array([1e100, 1.0, -1e100])) == 1.0


def test_neumaier_matches_exact_rational_sum():
    rng = np.random.default_rng(3)
    x = rng.normal(size=5000) * 10.0 ** rng.integers(-8, 8, size=5000)
    assert O.neumaier_sum(x) == math.fsum(x)


def test_neumaier_nonfinite_falls_back_to_naive():
    assert O.neumaier_sum(np.array([math.inf, 1.0])) == math.inf
    assert math.isnan(O.neumaier_sum(np.array([math.inf, -math.inf])))
    assert math.isnan(O.neumaier_sum(np.array([1.0, math.nan])))


# ---- scalar aggregates: tests/aggregate_nulls_e2e.rs fixtures ---------------------

def _agg(values, dtype, valid=None):
    col = O.HostCol(np.asarray(values), dtype, valid)
    return {op: O.scalar_agg(col, op) for op in (O.COUNT, O.COUNT_STAR, O.SUM, O.MIN, O.MAX, O.AVG)}


def test_scalar_i32_nulls_fixture():
    # aggregate_nulls_e2e.rs:118-183: [10, NULL, -5, 30, NULL, 7]
    r = _agg(np.array([10, 0, -5, 30, 0, 7], np.int32), O.I32, [1, 0, 1, 1, 0, 1])
    assert O.scalar_value(r[O.COUNT]) == 4
    assert O.scalar_value(r[O.COUNT_STAR]) == 6  # :379-382
    assert O.scalar_value(r[O.SUM]) == 42 and r[O.SUM].dtype == O.I64
    assert O.scalar_value(r[O.MIN]) == -5
    assert O.scalar_value(r[O.MAX]) == 30
    assert abs(O.scalar_value(r[O.AVG]) - 10.5) < 1e-12


def test_scalar_i64_nulls_fixture():
    # aggregate_nulls_e2e.rs:193-245: [100, NULL, NULL, 50, 25]
    r = _agg(np.array([100, 0, 0, 50, 25], np.int64), O.I64, [1, 0, 0, 1, 1])
    assert O.scalar_value(r[O.COUNT]) == 3
    assert O.scalar_value(r[O.SUM]) == 175
    assert O.scalar_value(r[O.MIN]) == 25
    assert O.scalar_value(r[O.MAX]) == 100
    assert abs(O.scalar_value(r[O.AVG]) - 175 / 3) < 1e-12


def test_scalar_f32_nulls_fixture():
    # aggregate_nulls_e2e.rs:250-320: [1.5, NULL, 2.5, NULL, -3.0, 4.0]
    r = _agg(np.array([1.5, 0, 2.5, 0, -3.0, 4.0], np.float32), O.F32, [1, 0, 1, 0, 1, 1])
    assert O.scalar_value(r[O.COUNT]) == 4
    assert O.scalar_value(r[O.SUM]) == 5.0
    assert O.scalar_value(r[O.MIN]) == -3.0
    assert O.scalar_value(r[O.MAX]) == 4.0
    assert O.scalar_value(r[O.AVG]) == 1.25


def test_scalar_f64_nulls_fixture():
    # aggregate_nulls_e2e.rs:326-366: [10, NULL, 20, 30, NULL]
    r = _agg(np.array([10.0, 0, 20.0, 30.0, 0], np.float64), O.F64, [1, 0, 1, 1, 0])
    assert O.scalar_value(r[O.COUNT]) == 3
    assert O.scalar_value(r[O.SUM]) == 60.0
    assert O.scalar_value(r[O.MIN]) == 10.0
    assert O.scalar_value(r[O.MAX]) == 30.0
    assert O.scalar_value(r[O.AVG]) == 20.0


def test_scalar_all_null_is_sql_null():
    # aggregate_nulls_e2e.rs:424-470
    for dt, npdt in ((O.I64, np.int64), (O.F64, np.float64)):
        r = _agg(np.zeros(3, npdt), dt, [0, 0, 0])
        assert O.scalar_value(r[O.COUNT]) == 0
        for op in (O.SUM, O.MIN, O.MAX, O.AVG):
            assert r[op].is_null == 1


def test_scalar_no_null_fast_path():
    # aggregate_nulls_e2e.rs:392-421: [1..5]
    r = _agg(np.arange(1, 6, dtype=np.int64), O.I64)
    assert [O.scalar_value(r[o]) for o in (O.COUNT, O.SUM, O.MIN, O.MAX)] == [5, 15, 1, 5]
    assert O.scalar_value(r[O.AVG]) == 3.0


def test_scalar_sum_overflow_errors():
    # aggregate.rs:2120-2140 checked_add -> "SUM(integer) overflow"
    col = O.HostCol(np.array([2 ** 63 - 1, 1], np.int64), O.I64)
    assert O.scalar_agg(col, O.SUM).status == O.ERR_TYPE
    col = O.HostCol(np.array([-(2 ** 63), -1], np.int64), O.I64)
    assert O.scalar_agg(col, O.SUM).status == O.ERR_TYPE
    col = O.HostCol(np.array([2 ** 62, 2 ** 62 - 1], np.int64), O.I64)
    assert O.scalar_value(O.scalar_agg(col, O.SUM)) == 2 ** 63 - 1


def test_scalar_float_min_max_nan_total_order():
    # aggregate.rs:2165 float_total_cmp: NaN is the largest value
    col = O.HostCol(np.array([np.nan, 1.0, -2.0]), O.F64)
    assert O.scalar_value(O.scalar_agg(col, O.MIN)) == -2.0
    assert math.isnan(O.scalar_value(O.scalar_agg(col, O.MAX)))
    col = O.HostCol(np.array([np.nan, np.nan]), O.F64)
    assert math.isnan(O.scalar_value(O.scalar_agg(col, O.MIN)))
    col = O.HostCol(np.array([0.0, -0.0]), O.F64)  # total order: -0.0 < +0.0
    assert math.copysign(1, O.scalar_value(O.scalar_agg(col, O.MIN))) == -1
    assert math.copysign(1, O.scalar_value(O.scalar_agg(col, O.MAX))) == 1


def test_scalar_stddev_hand_values():
    # tests/stddev_e2e.rs:200-260: [1,2,3,4,5] -> pop sqrt(2), samp sqrt(2.5)
    col = O.HostCol(np.arange(1, 6, dtype=np.int32), O.I32)
    assert abs(O.scalar_value(O.scalar_agg(col, O.STDDEV_POP)) - math.sqrt(2)) < 1e-12
    assert abs(O.scalar_value(O.scalar_agg(col, O.STDDEV_SAMP)) - math.sqrt(2.5)) < 1e-12
    assert abs(O.scalar_value(O.scalar_agg(col, O.VAR_SAMP)) - 2.5) < 1e-12
    one = O.HostCol(np.array([7], np.int32), O.I32)
    assert O.scalar_agg(one, O.STDDEV_SAMP).is_null == 1  # NULL when N <= 1
    assert O.scalar_value(O.scalar_agg(one, O.STDDEV_POP)) == 0.0


def test_scalar_stddev_matches_numpy():
    x = np.random.default_rng(5).normal(3.0, 2.0, size=10001)
    col = O.HostCol(x, O.F64)
    assert abs(O.scalar_value(O.scalar_agg(col, O.VAR_POP)) - np.var(x)) < 1e-9
    assert abs(O.scalar_value(O.scalar_agg(col, O.STDDEV_SAMP)) - np.std(x, ddof=1)) < 1e-9


def test_scalar_decimal_sum_fixtures():
    # tests/decimal_type_test.rs:324-431
    def dec(vals, valid=None):
        a = np.array([[v & M64, (v >> 64) & M64] for v in vals], dtype=np.uint64).view(np.int64)
        return O.HostCol(a, O.DEC128, valid, scale=2)

    r = O.scalar_agg(dec([123, -456, 789, 1, 10_000]), O.SUM)
    assert O.scalar_value(r) == 123 - 456 + 789 + 1 + 10_000 and r.dtype == O.DEC128
    r = O.scalar_agg(dec([100, 0, -25, 0, 75], [1, 0, 1, 0, 1]), O.SUM)
    assert O.scalar_value(r) == 150
    big = 10 ** 38 - 1
    assert O.scalar_agg(dec([big, big]), O.SUM).status == O.ERR_TYPE


# ---- filter (src/exec/compact.rs:146) --------------------------------------------

def test_filter_matches_numpy_and_null_semantics():
    rng = np.random.default_rng(11)
    a = rng.integers(-50, 50, size=2000).astype(np.int32)
    b = rng.uniform(0, 100, size=2000)
    b[::37] = np.nan
    va = rng.random(2000) > 0.1
    cols = [O.HostCol(a, O.I32, va), O.HostCol(b, O.F64)]
    got = O.filter_rows(cols, [(0, O.GE, -10), (1, O.LT, 60.0)], 2000)
    want = np.nonzero(va & (a >= -10) & (b < 60.0))[0]
    assert np.array_equal(got, want)
    got = O.filter_rows(cols, [(0, O.IS_NULL, 0)], 2000)
    assert np.array_equal(got, np.nonzero(~va)[0])


# ---- GROUP BY -----------------------------------------------------------------------

def test_groupby_known_sums_per_group_n2():
    # tests/tier2_multi_sum_e2e.rs:392-440
    keys = np.array([10, 20, 10, 20, 10, 30, -7, -7], np.int32)
    v0 = np.array([1.0, 4.0, 2.0, 5.0, 3.0, 6.0, 7.0, 8.0])
    v1 = v0 * 10
    r = O.groupby([O.HostCol(keys, O.I32)], [O.HostCol(v0, O.F64), O.HostCol(v1, O.F64)], [(O.SUM, 0), (O.SUM, 1)])
    got = {int(np.int64(k).astype(np.uint32).view(np.int32)): (r["f64"][0][i], r["f64"][1][i])
           for i, k in enumerate(r["keys"])}
    assert got == {-7: (15.0, 150.0), 10: (6.0, 60.0), 20: (9.0, 90.0), 30: (6.0, 60.0)}
    # int32 keys are packed as zero-extended u32, so -7 sorts last (groupby.rs:800)
    assert r["keys"].tolist() == [10, 20, 30, (2 ** 32) - 7]


def test_groupby_matches_pandas_bruteforce():
    rng = np.random.default_rng(17)
    n = 20000
    k = rng.integers(-300, 300, size=n).astype(np.int32)
    v = rng.integers(-10 ** 6, 10 ** 6, size=n).astype(np.int64)
    f = rng.normal(size=n)
    vv = rng.random(n) > 0.05
    aggs = [(O.SUM, 0), (O.COUNT_STAR, 0), (O.MIN, 0), (O.MAX, 0), (O.COUNT, 0), (O.SUM, 1), (O.AVG, 1),
            (O.MIN, 1), (O.MAX, 1), (O.VAR_SAMP, 1)]
    r = O.groupby([O.HostCol(k, O.I32)], [O.HostCol(v, O.I64, vv), O.HostCol(f, O.F64)], aggs)
    df = pd.DataFrame({"k": k, "v": np.where(vv, v, 0), "vv": vv, "f": f})
    dv = df[df.vv]
    g = df.groupby("k")
    gv = dv.groupby("k")
    keys = np.sort(np.unique(k).astype(np.uint32).astype(np.int64))
    assert np.array_equal(r["keys"], keys)
    ks = keys.astype(np.uint32).view(np.int32)
    assert np.array_equal(r["i64"][0], gv.v.sum().reindex(ks).to_numpy())
    assert np.array_equal(r["i64"][1], g.size().reindex(ks).to_numpy())
    assert np.array_equal(r["i64"][2], gv.v.min().reindex(ks).to_numpy())
    assert np.array_equal(r["i64"][3], gv.v.max().reindex(ks).to_numpy())
    assert np.array_equal(r["i64"][4], gv.v.count().reindex(ks).to_numpy())
    assert np.allclose(r["f64"][5], g.f.sum().reindex(ks).to_numpy(), rtol=1e-12, atol=1e-12)
    assert np.allclose(r["f64"][6], g.f.mean().reindex(ks).to_numpy(), rtol=1e-12, atol=1e-12)
    assert np.array_equal(r["f64"][7], g.f.min().reindex(ks).to_numpy())
    assert np.array_equal(r["f64"][8], g.f.max().reindex(ks).to_numpy())
    assert np.allclose(r["f64"][9], g.f.var(ddof=1).reindex(ks).to_numpy(), rtol=1e-9)


def test_groupby_growth_and_output_capacity():
    """Many groups (the per-group tables grow many times) against pandas, and
    the same result whether the outputs are sized for every row, for too few
    groups (one more fold at the exact size) or exactly."""
    rng = np.random.default_rng(23)
    n = 300_000
    k = rng.integers(-2 ** 40, 2 ** 40, size=n).astype(np.int64) // 7
    k[::3] = k[1::3][: len(k[::3])]  # some repeats
    v = rng.integers(-1000, 1000, size=n).astype(np.int64)
    aggs = [(O.SUM, 0), (O.COUNT_STAR, 0), (O.MAX, 0)]
    full = O.groupby([O.HostCol(k, O.I64)], [O.HostCol(v, O.I64)], aggs)
    df = pd.DataFrame({"k": k, "v": v}).groupby("k").v
    ks = np.sort(np.unique(k))
    assert np.array_equal(full["keys"], ks)
    assert np.array_equal(full["i64"][0], df.sum().reindex(ks).to_numpy())
    assert np.array_equal(full["i64"][1], df.size().reindex(ks).to_numpy())
    assert np.array_equal(full["i64"][2], df.max().reindex(ks).to_numpy())
    for cap in (10, len(ks)):
        r = O.groupby([O.HostCol(k, O.I64)], [O.HostCol(v, O.I64)], aggs, max_groups=cap)
        assert r["n_groups"] == len(ks)
        assert np.array_equal(r["keys"], full["keys"])
        assert all(np.array_equal(a, b) for a, b in zip(r["i64"], full["i64"]))


def _dec(vals):
    """Decimal128 values as (n, 2) little-endian two's complement words."""
    m = (1 << 64) - 1
    return np.array([[v & m, (v >> 64) & m] for v in vals], dtype=np.uint64).view(np.int64)


def test_groupby_decimal128_keys_bruteforce():
    """Decimal128 GROUP BY keys (the reference rejects them, groupby_common.rs:117; here a
    single key column whose values fit i64 groups by its exact value, in value order, with
    its NULL group last) against a Python fold; a value beyond 64 bits is refused."""
    rng = np.random.default_rng(31)
    n = 5000
    vals = [int(x) for x in rng.integers(-40, 40, n) * 10 ** 15 + rng.integers(-3, 3, n)]
    valid = rng.random(n) > 0.1
    v = rng.integers(-1000, 1000, n).astype(np.int64)
    r = O.groupby([O.HostCol(_dec(vals), O.DEC128, valid)], [O.HostCol(v, O.I64)], [(O.SUM, 0), (O.COUNT_STAR, 0)])
    assert r["status"] == O.OK
    want = {}
    for x, ok, y in zip(vals, valid, v):
        k = x if ok else None
        s, c = want.get(k, (0, 0))
        want[k] = (s + int(y), c + 1)
    keys = sorted(k for k in want if k is not None) + [None]
    assert r["n_groups"] == len(keys)
    assert r["key_null"].tolist() == [k is None for k in keys]
    assert [int(x) for x in r["keys"][:-1]] == keys[:-1]
    assert r["i64"][0].tolist() == [want[k][0] for k in keys]
    assert r["i64"][1].tolist() == [want[k][1] for k in keys]
    wide = _dec([1, 2 ** 63, 3])
    assert O.groupby([O.HostCol(wide, O.DEC128)], [O.HostCol(v[:3], O.I64)], [(O.SUM, 0)])["status"] != O.OK
    # a wide value in a NULL row does not count
    assert O.groupby([O.HostCol(wide, O.DEC128, [1, 0, 1])], [O.HostCol(v[:3], O.I64)], [(O.SUM, 0)])["status"] == O.OK


def test_groupby_null_group_single_column():
    # groupby.rs:528-606: all NULL-key rows form one group, key NULL, sorted last
    k = np.array([1, 2, 0, 1, 0], np.int64)
    v = np.array([10, 20, 30, 40, 50], np.int64)
    r = O.groupby([O.HostCol(k, O.I64, [1, 1, 0, 1, 0])], [O.HostCol(v, O.I64)], [(O.SUM, 0), (O.COUNT_STAR, 0)])
    assert r["n_groups"] == 3
    assert r["key_null"].tolist() == [False, False, True]
    assert r["i64"][0].tolist() == [50, 20, 80]
    assert r["i64"][1].tolist() == [2, 1, 2]


def test_groupby_two_key_packing_edges():
    # tests/tier2_twokey_e2e.rs:96-160: i32::MIN / MAX survive the (a<<32)|b pack
    a = np.array([-2 ** 31, 2 ** 31 - 1, -2 ** 31, 0], np.int32)
    b = np.array([2 ** 31 - 1, -2 ** 31, 2 ** 31 - 1, -1], np.int32)
    r = O.groupby([O.HostCol(a, O.I32), O.HostCol(b, O.I32)], [O.HostCol(np.ones(4, np.int64), O.I64)],
                  [(O.SUM, 0)])
    packed = sorted(((int(x) & 0xFFFFFFFF) << 32 | (int(y) & 0xFFFFFFFF)) for x, y in set(zip(a, b)))
    packed = [p - (1 << 64) if p >= 1 << 63 else p for p in packed]
    assert sorted(r["keys"].tolist()) == sorted(packed)
    assert dict(zip(r["keys"].tolist(), r["i64"][0].tolist()))[packed[[p for p in packed].index(
        ((0x80000000 << 32) | 0x7FFFFFFF) - (1 << 64))]] == 2


def test_groupby_float_keys_canonicalised():
    # groupby_common.rs:302: -0.0 == +0.0, all NaNs one group
    k = np.array([0.0, -0.0, np.nan, -np.nan, 1.5], np.float64)
    r = O.groupby([O.HostCol(k, O.F64)], [O.HostCol(np.ones(5, np.int64), O.I64)], [(O.COUNT_STAR, 0)])
    assert r["n_groups"] == 3
    assert sorted(r["i64"][0].tolist()) == [1, 2, 2]


def test_groupby_date32_keys_pack_like_int32():
    k = np.array([-3, 5, -3, 7], np.int32)
    r32 = O.groupby([O.HostCol(k, O.I32)], [O.HostCol(np.ones(4, np.int64), O.I64)], [(O.SUM, 0)])
    rd = O.groupby([O.HostCol(k, O.DATE32)], [O.HostCol(np.ones(4, np.int64), O.I64)], [(O.SUM, 0)])
    assert rd["keys"].tolist() == r32["keys"].tolist() == [5, 7, 2 ** 32 - 3]
    assert rd["i64"][0].tolist() == [1, 1, 2]


def test_groupby_integer_sum_overflow():
    # groupby.rs:1775 checked_group_sum
    k = np.array([1, 1, 2], np.int32)
    v = np.array([2 ** 63 - 1, 1, 5], np.int64)
    assert O.groupby([O.HostCol(k, O.I32)], [O.HostCol(v, O.I64)], [(O.SUM, 0)])["status"] == O.ERR_TYPE


# ---- INNER join ------------------------------------------------------------------------

def test_join_gpu_e2e_fixture():
    # tests/gpu_join_e2e.rs:68-142: build keys 0..4096 unique, probe i % 8192
    nb, npb = 4096, 8192
    bk = np.arange(nb, dtype=np.int32)
    pk = (np.arange(npb) % (2 * nb)).astype(np.int32)
    l, r = O.inner_join(O.HostCol(bk, O.I32), O.HostCol(pk, O.I32))
    assert len(l) == 4096
    assert np.all(bk[l] == pk[r])
    payload_b = 1000 + bk
    assert np.all(payload_b[l] == 1000 + pk[r])


def test_join_matches_bruteforce_with_duplicates_and_nulls():
    rng = np.random.default_rng(23)
    lk = rng.integers(0, 40, size=300).astype(np.int64)
    rk = rng.integers(0, 40, size=500).astype(np.int64)
    lv = rng.random(300) > 0.1
    rv = rng.random(500) > 0.1
    l, r = O.inner_join(O.HostCol(lk, O.I64, lv), O.HostCol(rk, O.I64, rv))
    want = sorted((i, j) for i in range(300) for j in range(500) if lv[i] and rv[j] and lk[i] == rk[j])
    assert sorted(zip(l.tolist(), r.tolist())) == want
    # emission order (join.rs:226-240): probe (right, larger) order, build rows in build order
    assert r.tolist() == sorted(r.tolist())


# ---- joins of every kind (src/exec/join.rs:132) ---------------------------------------------

NUL = 0xFFFFFFFF


def _outer_fixture_sides():
    # tests/gpu_join_e2e.rs:378-487: left keys 0..4096, right keys 2048..6144
    return (O.HostCol(np.arange(4096, dtype=np.int32), O.I32),
            O.HostCol(np.arange(2048, 2048 + 4096, dtype=np.int32), O.I32))


def test_join_left_outer_fixture():
    # tests/gpu_join_e2e.rs:378-417: every left row once, 2048 with the right side NULL
    lt, rt = _outer_fixture_sides()
    l, r = O.join([lt], [rt], "left")
    assert len(l) == 4096 and (r == NUL).sum() == 2048 and (l == NUL).sum() == 0
    assert sorted(l.tolist()) == list(range(4096))


def test_join_right_outer_fixture():
    # tests/gpu_join_e2e.rs:421-450: every right row once, 2048 with the left side NULL
    lt, rt = _outer_fixture_sides()
    l, r = O.join([lt], [rt], "right")
    assert len(l) == 4096 and (l == NUL).sum() == 2048 and (r == NUL).sum() == 0
    assert sorted(r.tolist()) == list(range(4096))


def test_join_full_outer_fixture():
    # tests/gpu_join_e2e.rs:454-487: 2048 matched + 2048 left-only + 2048 right-only
    lt, rt = _outer_fixture_sides()
    l, r = O.join([lt], [rt], "full")
    assert len(l) == 6144 and (l == NUL).sum() == 2048 and (r == NUL).sum() == 2048


def test_join_duplicate_build_keys_fixture():
    # tests/gpu_join_e2e.rs:491-524: 4 build x 8 probe rows per key, 1024 keys
    bk = (np.arange(4096) % 1024).astype(np.int32)
    pk = (np.arange(8192) % 1024).astype(np.int32)
    l, r = O.join([O.HostCol(bk, O.I32)], [O.HostCol(pk, O.I32)], "inner")
    assert len(l) == 32_768


def test_join_float_nan_fixture():
    # tests/gpu_join_e2e.rs:528-597: every NaN probe (a non-canonical NaN
    # pattern) meets the one NaN build row
    bk = np.arange(1, 4097, dtype=np.float64)
    bk[-1] = np.nan
    weird = np.array([np.nan]).view(np.uint64) ^ np.uint64(1)
    pk = np.where(np.arange(8192) < 4095, np.arange(1, 8193, dtype=np.float64), weird.view(np.float64)[0])
    assert np.isnan(pk[-1]) and pk[-1:].view(np.uint64)[0] != np.array([np.nan]).view(np.uint64)[0]
    l, r = O.join([O.HostCol(bk, O.F64)], [O.HostCol(pk, O.F64)], "inner")
    assert len(l) == (4096 - 1) + (8192 - (4096 - 1))


def test_join_two_key_fixture():
    # tests/gpu_join_e2e.rs:320-374: (i / 64, i % 64) build tuples, half the probe hits
    i = np.arange(4096, dtype=np.int32)
    j = np.arange(8192, dtype=np.int32)
    pa = np.where(j < 4096, j // 64, 1000 + j // 64).astype(np.int32)
    l, r = O.join([O.HostCol(i // 64, O.I32), O.HostCol(i % 64, O.I32)],
                  [O.HostCol(pa, O.I32), O.HostCol(j % 64, O.I32)], "inner")
    assert len(l) == 4096
    assert np.array_equal(np.sort(r), np.arange(4096))


def _brute_join(lvals, lvalid, rvals, rvalid, how):
    """Nested loops over SQL key equality (no hashing), in the reference's host order."""
    nl, nr = len(lvals), len(rvals)
    if how == "inner" and (nl == 0 or nr == 0):
        return []
    build_is_left = (nl <= nr) if how == "inner" else how == "right"
    bv, bok, pv, pok = (lvals, lvalid, rvals, rvalid) if build_is_left else (rvals, rvalid, lvals, lvalid)
    out, hit = [], [False] * len(bv)
    for j in range(len(pv)):
        ms = [i for i in range(len(bv)) if pok[j] and bok[i] and bv[i] == pv[j]]
        for i in ms:
            out.append((i, j))
            hit[i] = True
        if not ms and how != "inner":
            out.append((NUL, j))
    if how == "full":
        out += [(i, NUL) for i in range(len(bv)) if not hit[i]]
    return [(b, p) if build_is_left else (p, b) for b, p in out]


def _sql_key(v):
    if isinstance(v, float):
        return "nan" if math.isnan(v) else v + 0.0  # -0.0 == +0.0; NaN == NaN
    return v


@pytest.mark.parametrize("how", ["inner", "left", "right", "full"])
@pytest.mark.parametrize("kind", ["i64", "f64", "f32", "bool", "two_i32", "date32", "two_i64", "three_mixed"])
def test_join_kinds_match_bruteforce(how, kind):
    rng = np.random.default_rng(hash((how, kind)) & 0xFFFF)
    nl, nr = 70, 90
    valid_l, valid_r = rng.random(nl) > 0.15, rng.random(nr) > 0.15
    def col(n):
        if kind == "i64":
            return [rng.integers(-3, 9, n).astype(np.int64)], [O.I64]
        if kind == "date32":
            return [rng.integers(18000, 18012, n).astype(np.int32)], [O.DATE32]
        if kind == "f64":
            v = rng.choice(np.array([0.0, -0.0, 1.5, -2.25, np.nan, 7.0]), n)
            v.view(np.uint64)[rng.random(n) < 0.1] = np.uint64(0x7FF8000000000003)  # NaN, non-canonical bits
            return [v], [O.F64]
        if kind == "f32":
            return [rng.choice(np.array([0.0, -0.0, 1.5, np.nan, 3.0], np.float32), n)], [O.F32]
        if kind == "bool":
            return [rng.integers(0, 2, n).astype(np.uint8)], [O.BOOL]
        if kind == "two_i64":
            return [rng.integers(-2, 3, n).astype(np.int64) << 40, rng.integers(0, 4, n).astype(np.int64)], \
                [O.I64, O.I64]
        if kind == "three_mixed":
            return [rng.integers(0, 2, n).astype(np.int32), rng.integers(-1, 2, n).astype(np.int64) * (1 << 33),
                    rng.integers(18000, 18003, n).astype(np.int32)], [O.I32, O.I64, O.DATE32]
        return [rng.integers(-2, 3, n).astype(np.int32), rng.integers(0, 4, n).astype(np.int32)], [O.I32, O.I32]
    lv, lt = col(nl)
    rv, rt = col(nr)
    lcols = [O.HostCol(v, t, valid_l if c == 0 else None) for c, (v, t) in enumerate(zip(lv, lt))]
    rcols = [O.HostCol(v, t, valid_r if c == len(rv) - 1 else None) for c, (v, t) in enumerate(zip(rv, rt))]
    lnull = valid_l.copy()
    rnull = valid_r.copy()
    lkey = [tuple(_sql_key(v[i].item()) for v in lv) for i in range(nl)]
    rkey = [tuple(_sql_key(v[i].item()) for v in rv) for i in range(nr)]
    want = _brute_join(lkey, lnull, rkey, rnull, how)
    l, r = O.join(lcols, rcols, how)
    assert list(zip(l.tolist(), r.tolist())) == want


def test_join_lossy_two_i64_fixtures():
    # tests/gpu_join_e2e.rs:762-832: unique (i / 2, 3i + 1) build tuples, half the probe hits
    i = np.arange(4096, dtype=np.int64)
    j = np.arange(8192, dtype=np.int64)
    pa = np.where(j < 4096, j // 2, 99_999 + j)
    pb = np.where(j < 4096, j * 3 + 1, 88_888 + j)
    l, r = O.join([O.HostCol(i // 2, O.I64), O.HostCol(i * 3 + 1, O.I64)],
                  [O.HostCol(pa, O.I64), O.HostCol(pb, O.I64)], "inner")
    assert len(l) == 4096
    # tests/gpu_join_e2e.rs:1063-1162: LEFT over two Int64 keys, 2048 of 4096 matched
    n = np.arange(4096, dtype=np.int64)
    m = np.arange(2048, dtype=np.int64)
    l, r = O.join([O.HostCol(n, O.I64), O.HostCol(1_000_000 + n, O.I64)],
                  [O.HostCol(m, O.I64), O.HostCol(1_000_000 + m, O.I64)], "left")
    assert len(l) == 4096 and (r != NUL).sum() == 2048 and (r == NUL).sum() == 2048


def test_join_outer_empty_sides():
    # src/exec/join.rs:320-360: one side empty
    e = O.HostCol(np.zeros(0, np.int64), O.I64)
    x = O.HostCol(np.array([5, 6, 5], np.int64), O.I64)
    assert O.join([x], [e], "left")[0].tolist() == [0, 1, 2]
    assert O.join([x], [e], "full")[1].tolist() == [NUL] * 3
    assert len(O.join([x], [e], "right")[0]) == 0
    assert O.join([e], [x], "full")[1].tolist() == [0, 1, 2]
    assert len(O.join([e], [x], "inner")[0]) == 0


# ---- ORDER BY ----------------------------------------------------------------------------

def test_sort_matches_numpy_stable_lexsort():
    rng = np.random.default_rng(29)
    a = rng.integers(-5, 5, size=3000).astype(np.int32)
    b = rng.integers(-(2 ** 62), 2 ** 62, size=3000).astype(np.int64)
    got = O.sort_indices([O.HostCol(a, O.I32), O.HostCol(b, O.I64)], [False, True], [True, True])
    want = np.lexsort((-b, a))  # numpy lexsort is stable, last key primary
    assert np.array_equal(got, want)


def test_sort_float_nan_last_and_signed_zero():
    x = np.array([np.nan, 1.0, -0.0, 0.0, -np.inf, np.inf, -np.nan, -1.0])
    got = O.sort_indices([O.HostCol(x, O.F64)], [False], [True])
    # ascending, NaNs last and equal (stable), -0.0 == +0.0 (stable)
    assert got.tolist() == [4, 7, 2, 3, 1, 5, 0, 6]
    got = O.sort_indices([O.HostCol(x, O.F64)], [True], [True])
    assert got.tolist() == [0, 6, 5, 1, 2, 3, 7, 4]


def test_sort_decimal128_orders_as_signed_128_bit():
    """Decimal128 keys compare as i128: values straddling the 64-bit word
    boundary and negatives, checked against Python's integer order (a dropped
    high word, an unsigned high word or a signed low word each fail it)."""
    rng = np.random.default_rng(37)
    vals = [int(x) for x in rng.integers(-2 ** 62, 2 ** 62, 400)]
    vals += [2 ** 64 - 1, 2 ** 64, 2 ** 64 + 1, -(2 ** 64), -(2 ** 64) - 1, -1, 0, 1, 2 ** 127 - 1, -(2 ** 127)]
    vals += [v * 3 + (1 << 70) for v in vals[:50]]
    m = (1 << 64) - 1
    arr = np.array([[v & m, (v >> 64) & m] for v in vals], dtype=np.uint64).view(np.int64)
    for desc in (False, True):
        got = O.sort_indices([O.HostCol(arr, O.DEC128)], [desc], [True])
        want = sorted(range(len(vals)), key=lambda i: -vals[i] if desc else vals[i])  # stable
        assert got.tolist() == want


def test_limit_window_fixtures():
    # src/exec/limit.rs:54-131 unit tests (n, limit, offset) -> rows
    cases = [((10, 3, 0), [0, 1, 2]), ((10, 3, 2), [2, 3, 4]), ((5, 100, 0), [0, 1, 2, 3, 4]),
             ((10, 5, 7), [7, 8, 9]), ((4, 10, 4), []), ((4, 10, 999), []), ((10, 0, 0), []), ((0, 10, 0), [])]
    for (n, limit, offset), want in cases:
        a, b = O.limit_window(n, limit, offset)
        assert list(range(n))[a:b] == want
    v = np.arange(10, dtype=np.int32)
    got = O.sort_limit([O.HostCol(v, O.I32)], [False], [True], 3, 2)
    assert got.tolist() == [2, 3, 4]


def test_sort_nulls_first_last():
    x = np.array([3, 1, 2, 0], np.int64)
    valid = [1, 1, 1, 0]
    assert O.sort_indices([O.HostCol(x, O.I64, valid)], [False], [True]).tolist() == [3, 1, 2, 0]
    assert O.sort_indices([O.HostCol(x, O.I64, valid)], [False], [False]).tolist() == [1, 2, 0, 3]


# ---- fused filter + join + aggregate -----------------------------------------------------------

def test_filter_join_agg_matches_bruteforce():
    rng = np.random.default_rng(31)
    bk = rng.permutation(5000).astype(np.int64) * 7 - 1000
    bp = rng.integers(0, 1_000_000, size=5000).astype(np.int32)
    pk = np.where(rng.random(20000) < 0.5, rng.choice(bk, 20000), rng.integers(10 ** 9, 10 ** 10, 20000))
    pa = rng.uniform(0, 100, 20000)
    r = O.filter_join_agg(bk, bp, pk, pa, 10.0)
    sel = pa > 10.0
    pos = {int(k): i for i, k in enumerate(bk)}
    idx = np.array([pos.get(int(k), -1) for k in pk])
    m = sel & (idx >= 0)
    assert r["status"] == O.OK
    assert r["count"] == int(m.sum())
    assert r["sum_payload"] == int(bp[idx[m]].astype(np.int64).sum())
    assert r["sum_amount"] == math.fsum(pa[m])


# ---- per-row expression programs (src/plan/physical_plan.rs:81, src/jit/ptx_gen.rs) --------------

from craton_bolt_amd.ops import Program  # noqa: E402  (the op builder only; nothing runs on a device)

I32_MIN, I32_MAX = -(1 << 31), (1 << 31) - 1


def _run_prog(cols, prog, out_dtypes, out_validity=None):
    n = cols[0].col.n
    st, vals, valid = O.project(cols, prog.ops, out_dtypes, n, out_validity)
    assert st == O.OK
    return vals, valid


def test_project_int_arithmetic_wraps_and_guards():
    a = np.array([7, -7, I32_MIN, I32_MAX, 100, 0, I32_MIN, 5], np.int32)
    b = np.array([2, 2, -1, 2, 0, 3, 1, -1], np.int32)
    p = Program()
    ra, rb = p.load(0, O.I32), p.load(1, O.I32)
    p.store(p.binary("+", ra, rb, O.I32), 0)
    p.store(p.binary("*", ra, rb, O.I32), 1)
    p.store(p.binary("/", ra, rb, O.I32), 2)
    p.store(p.binary("%", ra, rb, O.I32), 3)
    p.store(p.binary("-", ra, rb, O.I32), 4)
    vals, _ = _run_prog([O.HostCol(a, O.I32), O.HostCol(b, O.I32)], p, [O.I32] * 5)
    with np.errstate(over="ignore"):
        assert np.array_equal(vals[0], a + b) and np.array_equal(vals[1], a * b) and np.array_equal(vals[4], a - b)
    # ptx_gen.rs:2231-2400: x / 0 = 0, x % 0 = 0, MIN / -1 = MIN, MIN % -1 = 0; C truncating division
    want_div = [3, -3, I32_MIN, I32_MAX // 2, 0, 0, I32_MIN, -5]
    want_mod = [1, -1, 0, 1, 0, 0, 0, 0]
    assert vals[2].tolist() == want_div and vals[3].tolist() == want_mod


def test_project_int64_guards_match_python():
    rng = np.random.default_rng(4)
    a = rng.integers(-(1 << 62), 1 << 62, 2000).astype(np.int64)
    b = rng.integers(-50, 50, 2000).astype(np.int64)
    a[:3] = [-(1 << 63), -(1 << 63), 12]
    b[:3] = [-1, 0, 0]
    p = Program()
    ra, rb = p.load(0, O.I64), p.load(1, O.I64)
    p.store(p.binary("/", ra, rb, O.I64), 0)
    p.store(p.binary("%", ra, rb, O.I64), 1)
    vals, _ = _run_prog([O.HostCol(a, O.I64), O.HostCol(b, O.I64)], p, [O.I64, O.I64])
    for x, y, q, r in zip(a.tolist(), b.tolist(), vals[0].tolist(), vals[1].tolist()):
        if y == 0:
            assert (q, r) == (0, 0)
        elif x == -(1 << 63) and y == -1:
            assert (q, r) == (x, 0)
        else:
            tq = abs(x) // abs(y) * (1 if (x >= 0) == (y >= 0) else -1)
            assert (q, r) == (tq, x - tq * y)


def test_project_float_ops_compare_and_cast():
    x = np.array([1.5, -2.5, np.nan, np.inf, -0.0, 3e10, -3e10, 2.9999999], np.float64)
    y = np.array([1.5, 2.0, np.nan, 1.0, 0.0, 1.0, np.nan, 3.0], np.float64)
    p = Program()
    rx, ry = p.load(0, O.F64), p.load(1, O.F64)
    p.store(p.binary("!=", rx, ry, O.F64), 0)
    p.store(p.binary("==", rx, ry, O.F64), 1)
    p.store(p.cast(rx, O.I32), 2)
    p.store(p.cast(rx, O.BOOL), 3)
    p.store(p.binary("/", rx, ry, O.F64), 4)
    p.store(p.cast(rx, O.F32), 5)
    vals, _ = _run_prog([O.HostCol(x, O.F64), O.HostCol(y, O.F64)], p, [O.BOOL, O.BOOL, O.I32, O.BOOL, O.F64, O.F32])
    ordered_ne = (x < y) | (x > y)  # NaN compares false, != included (setp.ne is ordered)
    assert np.array_equal(vals[0].astype(bool), ordered_ne) and np.array_equal(vals[1].astype(bool), x == y)
    sat = np.where(np.isnan(x), 0, np.clip(np.trunc(np.nan_to_num(x, posinf=1e300, neginf=-1e300)), I32_MIN, I32_MAX))
    assert vals[2].tolist() == sat.astype(np.int64).tolist()  # cvt.rzi: truncate, saturate, NaN -> 0
    assert vals[3].astype(bool).tolist() == [True, True, False, True, False, True, True, True]
    with np.errstate(invalid="ignore", divide="ignore"):
        assert np.array_equal(vals[4], x / y, equal_nan=True)
    assert np.array_equal(vals[5], x.astype(np.float32), equal_nan=True)


def test_project_float32_rounds_once_per_op():
    rng = np.random.default_rng(8)
    a = rng.standard_normal(5000).astype(np.float32)
    b = rng.standard_normal(5000).astype(np.float32)
    c = rng.standard_normal(5000).astype(np.float32)
    p = Program()
    ra, rb, rc = p.load(0, O.F32), p.load(1, O.F32), p.load(2, O.F32)
    p.store(p.binary("+", p.binary("*", ra, rb, O.F32), rc, O.F32), 0)  # no fused multiply-add
    vals, _ = _run_prog([O.HostCol(a, O.F32), O.HostCol(b, O.F32), O.HostCol(c, O.F32)], p, [O.F32])
    assert np.array_equal(vals[0], (a * b) + c)


def test_project_select_not_is_null_and_validity_fold():
    v = np.array([5, 6, 7, 8], np.int64)
    w = np.array([1, 2, 3, 4], np.int64)
    vv = np.array([True, False, True, True])
    wv = np.array([True, True, False, True])
    p = Program()
    rv, rw = p.load(0, O.I64), p.load(1, O.I64)
    isn = p.is_null(0)
    p.store(p.select(isn, p.const(-1, O.I64), rv), 0)  # CASE WHEN v IS NULL THEN -1 ELSE v END
    p.store(p.binary("+", rv, rw, O.I64), 1)           # depends on v and w
    p.store(p.not_(p.binary(">", rw, p.const(2, O.I64), O.I64)), 2)
    vals, valid = _run_prog([O.HostCol(v, O.I64, vv), O.HostCol(w, O.I64, wv)], p, [O.I64, O.I64, O.BOOL],
                            [True, True, True])
    assert vals[0].tolist() == [5, -1, 7, 8]
    assert vals[1].tolist() == [6, 8, 10, 12]  # values of NULL rows come from the stored bits
    assert vals[2].tolist() == [1, 1, 0, 0]
    # AND of the validity of the loaded inputs (ptx_gen.rs:474); IS NULL reads no value
    assert valid[0].tolist() == vv.tolist()
    assert valid[1].tolist() == (vv & wv).tolist()
    assert valid[2].tolist() == wv.tolist()


def test_project_decimal_and_date_arithmetic():
    m = (1 << 64) - 1
    xs = [10 ** 20 + 7, -(10 ** 19), 5, -(1 << 126)]
    ys = [3, 10 ** 18, -5, 4]
    def dec(vals):
        return np.array([[v & m, (v >> 64) & m] for v in vals], dtype=np.uint64).view(np.int64)
    p = Program()
    rx, ry = p.load(0, O.DEC128), p.load(1, O.DEC128)
    p.store(p.binary("+", rx, ry, O.DEC128), 0)
    p.store(p.binary("*", rx, ry, O.DEC128), 1)
    p.store(p.binary("<", rx, ry, O.DEC128), 2)
    vals, _ = _run_prog([O.HostCol(dec(xs), O.DEC128), O.HostCol(dec(ys), O.DEC128)], p, [O.DEC128, O.DEC128, O.BOOL])
    def back(a):
        u = [(int(lo) & m) | ((int(hi) & m) << 64) for lo, hi in a]
        return [x - (1 << 128) if x >> 127 else x for x in u]
    wrap = lambda z: ((z + (1 << 127)) % (1 << 128)) - (1 << 127)  # noqa: E731
    assert back(vals[0]) == [wrap(x + y) for x, y in zip(xs, ys)]
    assert back(vals[1]) == [wrap(x * y) for x, y in zip(xs, ys)]  # -(2^126) * 4 wraps to 0
    assert vals[2].tolist() == [int(x < y) for x, y in zip(xs, ys)]
    d1 = np.array([19000, 0, -5], np.int32)
    d2 = np.array([18000, 1, 5], np.int32)
    q = Program()
    q.store(q.binary("-", q.load(0, O.DATE32), q.load(1, O.DATE32), O.DATE32), 0)  # Date32 - Date32 -> Int32
    vals, _ = _run_prog([O.HostCol(d1, O.DATE32), O.HostCol(d2, O.DATE32)], q, [O.I32])
    assert vals[0].tolist() == [1000, -1, -10]


def test_project_type_errors():
    a = O.HostCol(np.arange(4, dtype=np.int32), O.I32)
    p = Program()
    ra = p.load(0, O.I32)
    p.store(p.binary("+", ra, p.const(1.0, O.F64), O.I32), 0)  # operands of different dtypes: cast first
    st, _, _ = O.project([a], p.ops, [O.I32], 4)
    assert st == O.ERR_TYPE
