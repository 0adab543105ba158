"""GPU parity: hash GROUP BY against the oracle's restatement of the
reference's classic path (src/exec/groupby.rs:295-833).

Group sets, keys, counts, integer / decimal sums and MIN / MAX are bit-exact
and in the reference's output order (ascending packed key, groupby.rs:800).
Float sums use hardware atomics, whose order is not fixed: they must lie
within 1e-12 * sum(|x|) of the oracle's compensated sum (the forward-error
scale of recursive summation), and AVG / VAR / STDDEV within REL_TOL = 1e-9."""
import math

import numpy as np
import pytest

import oracle as O
from craton_bolt_amd import _native as N
from craton_bolt_amd import datagen as D
from craton_bolt_amd.device import DeviceColumn
from craton_bolt_amd.ops import Term, execute_groupby

from gpu_util import both, dec_array

pytestmark = pytest.mark.gpu
REL_TOL = 1e-9


@pytest.fixture(autouse=True, params=["jit", "generic"])
def kernel_path(request, monkeypatch):
    """Every case runs through the LDS-tier kernel generated per query with
    hipRTC and through the precompiled generic one (BOLT_JIT=0)."""
    monkeypatch.setenv("BOLT_JIT", "0" if request.param == "generic" else "1")
    return request.param


def _packed_from_output(res, key_dtypes):
    """Rebuild the reference's packed i64 key from the output key columns."""
    cols = res["keys"]
    if len(cols) == 1:
        vals, _ = cols[0]
        dt = key_dtypes[0]
        if dt in (N.I32, N.DATE32, N.F32):
            return vals.view(np.uint32).astype(np.int64)
        return vals.view(np.int64)
    a = cols[0][0].view(np.uint32).astype(np.uint64)
    b = cols[1][0].view(np.uint32).astype(np.uint64)
    return ((a << np.uint64(32)) | b).view(np.int64)


def run_both(keys, inputs, aggs, key_dtypes, terms=None, flags=0, abs_inputs=None):
    """Runs the GPU operator and the oracle on the same (filtered) data and
    compares everything.  keys/inputs: lists of (values, dtype, valid)."""
    dk = [both(v, dt, va) for v, dt, va in keys]
    di = [both(v, dt, va) for v, dt, va in inputs]
    res = execute_groupby([d for d, _ in dk], [d for d, _ in di], aggs, terms, sorted_output=True, flags=flags)
    # the oracle sees the rows that pass the filter
    hk, hi = [h for _, h in dk], [h for _, h in di]
    n = len(keys[0][0])
    if terms:
        rows = O.filter_rows(hi, [(t.col, t.op, t.value) for t in terms], n)
        sub = lambda v, va, dt: (v[rows], None if va is None else np.asarray(va)[rows], dt)
        keys = [sub(v, va, dt) for v, dt, va in keys]
        inputs = [sub(v, va, dt) for v, dt, va in inputs]
        keys = [(v, dt, va) for v, va, dt in keys]
        inputs = [(v, dt, va) for v, va, dt in inputs]
        hk = [both(v, dt, va)[1] for v, dt, va in keys]
        hi = [both(v, dt, va)[1] for v, dt, va in inputs]
    want = O.groupby(hk, hi, aggs)
    assert want["status"] == O.OK
    g = want["n_groups"]
    assert res["n_groups"] == g
    got_keys = _packed_from_output(res, key_dtypes)
    kvalid = res["keys"][0][1]
    if kvalid is not None:
        assert np.array_equal(~kvalid, want["key_null"])
        got_keys = np.where(kvalid, got_keys, np.int64(2 ** 63 - 1))
    assert np.array_equal(got_keys, want["keys"])
    # per-group sum of |x| for float tolerances
    abs_sum = {}
    for ai, (op, col) in enumerate(aggs):
        vals, valid = res["aggs"][ai]
        null = want["null"][ai]
        if valid is not None:
            assert np.array_equal(~valid, null), f"agg {ai} null mask"
        else:
            assert not null.any(), f"agg {ai} has NULL groups but no bitmap"
        live = ~null
        dt_in = inputs[col][1] if op != N.COUNT_STAR else N.I64
        if op in (N.COUNT, N.COUNT_STAR):
            assert np.array_equal(vals, want["i64"][ai])
        elif op == N.SUM and dt_in == N.DEC128:
            got = [(int(h) << 64) | (int(l) & (2 ** 64 - 1)) for l, h in vals]
            exp = [(int(h) << 64) | (int(l) & (2 ** 64 - 1)) for l, h in zip(want["i64"][ai], want["hi"][ai])]
            assert [x for x, lv in zip(got, live) if lv] == [x for x, lv in zip(exp, live) if lv]
        elif op in (N.SUM, N.MIN, N.MAX) and dt_in in (N.I32, N.I64, N.DATE32):
            assert np.array_equal(vals.astype(np.int64)[live], want["i64"][ai][live])
        elif op in (N.MIN, N.MAX):
            g_ = vals.astype(np.float64)[live]
            w_ = want["f64"][ai][live]
            assert np.array_equal(g_, w_, equal_nan=True)
            assert np.array_equal(np.signbit(g_), np.signbit(w_))
        elif op == N.SUM:
            if col not in abs_sum:
                v, dt, va = inputs[col]
                ha = both(np.abs(v.astype(np.float64)), N.F64, va)[1]
                r = O.groupby(hk, [ha], [(O.SUM, 0)])
                abs_sum[col] = r["f64"][0]
            tol = 1e-12 * abs_sum[col] + 1e-300
            if dt_in == N.F32:  # f32 output: the final rounding dominates
                tol = np.maximum(tol, 2 ** -23 * np.abs(want["f64"][ai]))
            err = np.abs(vals.astype(np.float64) - want["f64"][ai])
            assert np.all(err[live] <= tol[live]), f"SUM max err {err[live].max()}"
        else:
            g_ = vals.astype(np.float64)[live]
            w_ = want["f64"][ai][live]
            assert np.allclose(g_, w_, rtol=REL_TOL, atol=1e-12), f"agg {ai} op {op}"
    return res, want


def test_plain_groupby_workload_shape(gpu):
    # k int32 = draw % 1000, v int64 = draw % 1000; SUM, COUNT(*), MIN, MAX
    k, v = D.groupby_small(1_000_000)
    run_both([(k, N.I32, None)], [(v, N.I64, None)],
             [(N.SUM, 0), (N.COUNT_STAR, 0), (N.MIN, 0), (N.MAX, 0)], [N.I32])


@pytest.mark.parametrize("card", [1, 16, 1000, 100_000, 2_000_000])
@pytest.mark.parametrize("flags", [0, N.GROUPBY_FORCE_GLOBAL, N.GROUPBY_FORCE_PARTITIONED])
def test_cardinality_sweep(gpu, card, flags):
    print(f"[case] groupby card={card} flags={flags}", flush=True)
    n = 3_000_000
    rng = np.random.default_rng(card)
    k = rng.integers(0, card, n).astype(np.int64) * 7919 - 5
    v = rng.integers(-10 ** 6, 10 ** 6, n).astype(np.int64)
    f = rng.normal(size=n)
    run_both([(k, N.I64, None)], [(v, N.I64, None), (f, N.F64, None)],
             [(N.SUM, 0), (N.COUNT_STAR, 0), (N.MIN, 1), (N.MAX, 0), (N.SUM, 1), (N.AVG, 1)], [N.I64],
             flags=flags)


@pytest.mark.parametrize("n", [1, 1023, 8191, 8193, 300_001])
@pytest.mark.parametrize("flags", [0, N.GROUPBY_FORCE_PARTITIONED])
def test_ragged_sizes_with_spills(gpu, n, flags):
    """Row counts around the 1024-row and 8192-row workgroup iterations, with
    about as many groups as rows so most rows spill out of the LDS tables."""
    print(f"[case] groupby ragged n={n}", flush=True)
    rng = np.random.default_rng(n)
    k = rng.integers(0, max(n, 2), n).astype(np.int64)
    v = rng.integers(-1000, 1000, n).astype(np.int64)
    kv = rng.random(n) > 0.05
    run_both([(k, N.I64, kv)], [(v, N.I64, None)], [(N.SUM, 0), (N.COUNT_STAR, 0), (N.MIN, 0)], [N.I64],
             flags=flags)


def test_global_table_overflow_reruns_larger(gpu, monkeypatch):
    """A global table sized far below the group count overflows; the
    operator detects it with the extraction's read-back and reruns passes B
    and C with a 4x larger table (three times here: 1024 -> 65536 slots)."""
    monkeypatch.setenv("BOLT_GB_INITIAL_CAP", "1024")
    n = 500_000
    rng = np.random.default_rng(21)
    k = rng.integers(0, 20_000, n).astype(np.int64) * 31 + 7
    v = rng.integers(-100, 100, n).astype(np.int64)
    run_both([(k, N.I64, None)], [(v, N.I64, None)], [(N.SUM, 0), (N.COUNT_STAR, 0), (N.MAX, 0)], [N.I64])


@pytest.mark.parametrize("kdt", [N.I32, N.I64, N.F32, N.F64, N.DATE32])
@pytest.mark.parametrize("flags", [0, N.GROUPBY_FORCE_PARTITIONED])
def test_key_dtypes_nulls_and_all_aggregates(gpu, kdt, flags):
    n = 400_000
    rng = np.random.default_rng(kdt)
    base = rng.integers(-300, 300, n)
    if kdt in (N.F32, N.F64):
        k = (base / 4.0).astype(np.float32 if kdt == N.F32 else np.float64)
        k[::97] = np.nan
        k[::89] = -0.0
    else:
        k = base.astype(np.int64 if kdt == N.I64 else np.int32)
    kv = rng.random(n) > 0.02
    x = rng.integers(-50_000, 50_000, n).astype(np.int32)
    xv = rng.random(n) > 0.1
    y = rng.normal(3.0, 2.0, n)
    yv = rng.random(n) > 0.1
    z = rng.normal(size=n).astype(np.float32)
    aggs = [(N.SUM, 0), (N.COUNT, 0), (N.COUNT_STAR, 0), (N.MIN, 0), (N.MAX, 0), (N.AVG, 1), (N.MIN, 1), (N.MAX, 1),
            (N.VAR_SAMP, 1), (N.STDDEV_POP, 1), (N.SUM, 2), (N.MIN, 2)]
    run_both([(k, kdt, kv)], [(x, N.I32, xv), (y, N.F64, yv), (z, N.F32, None)], aggs, [kdt], flags=flags)


def test_filtered_groupby(gpu):
    n = 1_000_000
    rng = np.random.default_rng(5)
    k = rng.integers(0, 5000, n).astype(np.int32)
    a = rng.uniform(0, 100, n)
    b = rng.integers(0, 10, n).astype(np.int64)
    run_both([(k, N.I32, None)], [(a, N.F64, None), (b, N.I64, None)], [(N.SUM, 1), (N.COUNT_STAR, 0), (N.MAX, 0)],
             [N.I32], terms=[Term(0, N.GT, 10.0), Term(1, N.NE, 3)])


def test_two_key_packing_edges(gpu):
    # tests/tier2_twokey_e2e.rs:96-160
    n = 200_000
    rng = np.random.default_rng(9)
    choices = np.array([-2 ** 31, 2 ** 31 - 1, -1, 0, 1], np.int32)
    a = choices[rng.integers(0, 5, n)]
    b = rng.integers(-2 ** 31, 2 ** 31, n).astype(np.int32) % 17
    v = rng.integers(0, 100, n).astype(np.int64)
    run_both([(a, N.I32, None), (b.astype(np.int32), N.I32, None)], [(v, N.I64, None)],
             [(N.SUM, 0), (N.COUNT_STAR, 0)], [N.I32, N.I32])


def test_int64_extreme_keys_are_ordinary(gpu):
    # the all-ones key and i64::MIN / MAX need no sentinel routing
    k = np.array([-1, np.iinfo(np.int64).min, np.iinfo(np.int64).max, -1, 5, np.iinfo(np.int64).min], np.int64)
    v = np.arange(6, dtype=np.int64)
    run_both([(k, N.I64, None)], [(v, N.I64, None)], [(N.SUM, 0), (N.COUNT_STAR, 0)], [N.I64])


def test_known_sums_per_group_n2(gpu):
    # tests/tier2_multi_sum_e2e.rs:392-440
    keys = np.array([10, 20, 10, 20, 10, 30, -7, -7], np.int32)
    v0 = np.array([1.0, 4.0, 2.0, 5.0, 3.0, 6.0, 7.0, 8.0])
    res, _ = run_both([(keys, N.I32, None)], [(v0, N.F64, None), (v0 * 10, N.F64, None)],
                      [(N.SUM, 0), (N.SUM, 1)], [N.I32])
    assert res["keys"][0][0].tolist() == [10, 20, 30, -7]
    assert res["aggs"][0][0].tolist() == [6.0, 9.0, 6.0, 15.0]
    assert res["aggs"][1][0].tolist() == [60.0, 90.0, 60.0, 150.0]


def test_integer_sum_overflow_is_an_error(gpu):
    k = np.array([1, 1, 2], np.int32)
    v = np.array([2 ** 63 - 1, 1, 5], np.int64)
    with pytest.raises(N.BoltTypeError, match="overflow"):
        execute_groupby([both(k, N.I32, None)[0]], [both(v, N.I64, None)[0]], [(N.SUM, 0)])


@pytest.mark.parametrize("flags", [0, N.GROUPBY_FORCE_GLOBAL, N.GROUPBY_FORCE_PARTITIONED])
def test_null_group_dropped_next_to_a_genuine_max_key(gpu, flags):
    """A single nullable Int64 key column holding a genuine i64::MAX: the
    reference cannot tell that key from its NULL-group sentinel and runs the
    valid-flag executor, which drops the NULL-key rows (groupby.rs:553-583).
    Without such a key, the NULL rows form one NULL group."""
    rng = np.random.default_rng(21)
    n = 300_000
    k = rng.integers(-50, 50, n).astype(np.int64)
    k[::997] = 2 ** 63 - 1
    valid = rng.random(n) > 0.1
    v = rng.integers(-1000, 1000, n).astype(np.int64)
    aggs = [(N.SUM, 0), (N.COUNT_STAR, 0), (N.MAX, 0)]
    res, want = run_both([(k, N.I64, valid)], [(v, N.I64, None)], aggs, [N.I64], flags=flags)
    kv = res["keys"][0][1]
    assert kv is None or kv.all(), "NULL group present next to a genuine i64::MAX key"
    assert (res["keys"][0][0] == 2 ** 63 - 1).sum() == 1
    k2 = np.where(k == 2 ** 63 - 1, 7, k)
    res2, _ = run_both([(k2, N.I64, valid)], [(v, N.I64, None)], aggs, [N.I64], flags=flags)
    assert res2["keys"][0][1] is not None and (~res2["keys"][0][1]).sum() == 1


@pytest.mark.parametrize("n", [50_000, 1_500_000])
@pytest.mark.parametrize("flags", [0, N.GROUPBY_FORCE_PARTITIONED])
def test_decimal_sum_by_group(gpu, n, flags):
    # 1.5M rows: enough LDS-tier workgroups that the partial records take the
    # two-level (LDS, then global) merge
    # (partitioned: 20K groups, so spilled Decimal rows become merge records)
    rng = np.random.default_rng(11)
    k = rng.integers(0, 20_000 if flags else 100, n).astype(np.int32)
    vals = [int(x) * 10 ** 18 + int(y) for x, y in zip(rng.integers(-10 ** 12, 10 ** 12, n), rng.integers(0, 10 ** 18, n))]
    run_both([(k, N.I32, None)], [(dec_array(vals), N.DEC128, None)], [(N.SUM, 0), (N.COUNT_STAR, 0)], [N.I32],
             flags=flags)


def test_zipf_high_cardinality_float_sum(gpu):
    """SUM(f64) + COUNT over Zipf(1.1) keys; reports the max relative error of
    the atomic float sums against the compensated-sum oracle."""
    n = 4_000_000
    k, v = D.zipf_groupby(n, n_ranks=1_000_000)
    res, want = run_both([(k, N.I64, None)], [(v, N.F64, None)], [(N.SUM, 0), (N.COUNT_STAR, 0)], [N.I64])
    got = res["aggs"][0][0]
    rel = np.abs(got - want["f64"][0]) / np.maximum(np.abs(want["f64"][0]), 1e-300)
    print(f"[zipf] groups={res['n_groups']} max_rel_err={rel.max():.3e} median={np.median(rel):.3e}", flush=True)


@pytest.mark.parametrize("query", ["q1", "q2", "q3", "q4", "q5"])
def test_gpu_h2o_queries_match_oracle(gpu, query):
    """h2o.ai groupby q1..q5 (benches/olap_benchmarks.rs:77-90) on the reference's generator, 2M rows."""
    from craton_bolt_amd import datagen as D
    from craton_bolt_amd.ops import execute_groupby
    keys, vals, aggs = {"q1": (["id1"], ["v1"], [(1, 0)]), "q2": (["id2"], ["v1", "v2"], [(1, 0), (1, 1)]),
                        "q3": (["id1", "id2"], ["v1"], [(1, 0)]),
                        "q4": (["id1"], ["v1", "v2", "v3"], [(6, 0), (6, 1), (6, 2)]),
                        "q5": (["id3"], ["v1"], [(1, 0)])}[query]
    print(f"[case] h2o {query}", flush=True)
    n = 2_000_000
    kd = [both(D.h2o_column(k, n), N.I32) for k in keys]
    vd = [both(D.h2o_column(v, n), N.F64) for v in vals]
    got = execute_groupby([d for d, _ in kd], [d for d, _ in vd], aggs)
    want = O.groupby([h for _, h in kd], [h for _, h in vd], aggs)
    assert got["n_groups"] == want["n_groups"]
    if len(keys) == 1:
        assert np.array_equal(got["keys"][0][0].astype(np.int64), want["keys"])
    for a in range(len(aggs)):
        g = got["aggs"][a][0]
        w = want["f64"][a]
        assert np.allclose(g, w, rtol=1e-12, atol=0), query


@pytest.mark.parametrize("n", [40_000, 2_000_000])
@pytest.mark.parametrize("flags", [0, N.GROUPBY_FORCE_GLOBAL, N.GROUPBY_FORCE_PARTITIONED])
def test_decimal128_keys(gpu, n, flags):
    """GROUP BY a Decimal128(12, 2) key (the served fact table's price type):
    grouped by its exact value, in value order, NULL group last, output key
    column Decimal128 with the input's scale; integer / decimal sums and
    counts exact.  A value beyond 64 bits declines (the host path takes it)."""
    print(f"[case] decimal keys n={n} flags={flags}", flush=True)
    rng = np.random.default_rng(n + flags)
    vals = [int(x) for x in rng.integers(-5000, 5000, n) * 100 + rng.integers(0, 3, n)]
    valid = rng.random(n) > 0.05
    v = rng.integers(-1000, 1000, n).astype(np.int64)
    price = dec_array([int(x) * 10 ** 9 for x in rng.integers(-10 ** 6, 10 ** 6, n)])
    dk = DeviceColumn.from_numpy(dec_array(vals), N.DEC128, valid=valid, scale=2)
    aggs = [(N.SUM, 0), (N.COUNT_STAR, 0), (N.SUM, 1), (N.MAX, 0)]
    got = execute_groupby([dk], [DeviceColumn.from_numpy(v), DeviceColumn.from_numpy(price, N.DEC128)], aggs,
                          flags=flags)
    want = O.groupby([O.HostCol(dec_array(vals), O.DEC128, valid)], [O.HostCol(v, O.I64), O.HostCol(price, O.DEC128)],
                     aggs)
    assert want["status"] == O.OK
    assert got["n_groups"] == want["n_groups"]
    kv, kvalid = got["keys"][0]
    assert kv.shape == (want["n_groups"], 2)
    assert np.array_equal(~kvalid, want["key_null"])
    live = ~want["key_null"]
    assert np.array_equal(kv[live, 0], want["keys"][live])
    assert np.array_equal(kv[live, 1], want["keys"][live] >> 63)  # sign-extended high word
    assert np.array_equal(got["aggs"][0][0], want["i64"][0])
    assert np.array_equal(got["aggs"][1][0], want["i64"][1])
    dsum = got["aggs"][2][0]
    assert np.array_equal(dsum[:, 0], want["i64"][2]) and np.array_equal(dsum[:, 1], want["hi"][2])
    assert np.array_equal(got["aggs"][3][0], want["i64"][3])
    wide = dec_array([1, 2 ** 70, 3])
    with pytest.raises(N.GpuCapacity):
        execute_groupby([DeviceColumn.from_numpy(wide, N.DEC128)], [DeviceColumn.from_numpy(v[:3])], [(N.SUM, 0)])
