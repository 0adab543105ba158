"""GPU parity: the batched GROUP BY (bolt_batch.h) against the oracle and
against the per-query operator on the same inputs.

Each query is `GROUP BY k, SUM(v_i64), SUM(v_f32), COUNT(*)` over seeded
Zipf(1.1) int32 keys (the batched workload's shape).  Group sets, keys,
integer sums and counts are bit-exact; SUM(Float32) is accumulated in f64 and
rounded to f32 once, so it must lie within 1e-12 * sum(|x|) + 2^-24 * |sum|
of the oracle's compensated f64 sum.  Groups come out in hash order and are
compared after sorting by key."""
import math

import numpy as np
import pytest

import oracle as O
from craton_bolt_amd import _native as N
from craton_bolt_amd import datagen as D
from craton_bolt_amd.device import DeviceColumn
from craton_bolt_amd.ops import GroupByBatch, execute_groupby

pytestmark = pytest.mark.gpu

AGGS = [(N.SUM, 0), (N.SUM, 1), (N.COUNT_STAR, 0)]


def _queries(shapes, seed0=D.SEED_BATCH):
    host, dev = [], []
    for q, (n, card) in enumerate(shapes):
        k, vi, vf = D.zipf_small(n, card, seed=seed0 + q)
        host.append((k, vi, vf))
        dev.append(([DeviceColumn.from_numpy(k)], [DeviceColumn.from_numpy(vi), DeviceColumn.from_numpy(vf)]))
    return host, dev


def _sorted(res):
    k = res["keys"][0][0]
    order = np.argsort(k, kind="stable")
    return k[order], [(v[order], None if va is None else va[order]) for v, va in res["aggs"]]


def _check_vs_oracle(res, k, vi, vf):
    want = O.groupby([O.HostCol(k, O.I32)], [O.HostCol(vi, O.I64), O.HostCol(vf, O.F32)], AGGS)
    assert want["status"] == O.OK
    g = want["n_groups"]
    assert res["n_groups"] == g
    keys, aggs = _sorted(res)  # ascending int32
    wkeys = want["keys"].astype(np.uint32).view(np.int32)  # packed int32 keys are zero-extended
    worder = np.argsort(wkeys, kind="stable")
    assert np.array_equal(keys, wkeys[worder])
    assert np.array_equal(aggs[0][0], want["i64"][0][worder])
    assert np.array_equal(aggs[2][0], want["i64"][2][worder])
    got_f = aggs[1][0].astype(np.float64)
    want_f = want["f64"][1][worder]
    uk, inv = np.unique(k, return_inverse=True)
    assert np.array_equal(uk, keys)
    sabs = np.bincount(inv, weights=np.abs(vf.astype(np.float64)), minlength=len(uk))
    tol = 1e-12 * sabs + 2.0 ** -24 * np.abs(want_f) + 1e-30
    assert np.all(np.abs(got_f - want_f) <= tol)


SHAPES = [(0, 16), (1, 16), (1000, 16), (50_000, 40), (200_000, 3000), (120_000, 100_000), (10_000, 9_000),
          (300_000, 1_000_000), (4096, 64), (65_537, 700), (20_000, 20_000), (7, 3)]


def test_batch_matches_oracle_per_query(gpu):
    host, dev = _queries(SHAPES)
    b = GroupByBatch(dev, AGGS)
    for rnd in range(2):  # the second run replays the graph on tables left clean by the first
        print(f"[case] batch run {rnd}", flush=True)
        b.run()
        for q, (k, vi, vf) in enumerate(host):
            _check_vs_oracle(b.result(q), k, vi, vf)
    b.release()


def test_batch_equals_plain_path(gpu):
    host, dev = _queries(SHAPES[:8], seed0=D.SEED_BATCH + 1000)
    b = GroupByBatch(dev, AGGS)
    b.run()
    for q in range(len(dev)):
        plain = execute_groupby(dev[q][0], dev[q][1], AGGS, sorted_output=False)
        got = b.result(q)
        assert got["n_groups"] == plain["n_groups"]
        gk, ga = _sorted(got)
        pk, pa = _sorted(plain)
        assert np.array_equal(gk, pk)
        assert np.array_equal(ga[0][0], pa[0][0]) and np.array_equal(ga[2][0], pa[2][0])
        sabs = float(np.abs(host[q][2].astype(np.float64)).sum())
        assert np.all(np.abs(ga[1][0].astype(np.float64) - pa[1][0].astype(np.float64))
                      <= 1e-12 * sabs + 2.0 ** -23 * np.abs(pa[1][0].astype(np.float64)) + 1e-30)


def test_batch_nullable_keys_and_values(gpu):
    rng = np.random.default_rng(5)
    host, dev = [], []
    for q, n in enumerate([0, 10, 5000, 100_000]):
        k = rng.integers(-50, 50, n).astype(np.int64)
        kv = rng.random(n) > 0.1
        v = rng.integers(-10 ** 12, 10 ** 12, n).astype(np.int64)
        vv = rng.random(n) > 0.3
        host.append((k, kv, v, vv))
        dev.append(([DeviceColumn.from_numpy(k, valid=kv)], [DeviceColumn.from_numpy(v, valid=vv)]))
    aggs = [(N.SUM, 0), (N.COUNT, 0), (N.COUNT_STAR, 0), (N.MIN, 0), (N.MAX, 0)]
    b = GroupByBatch(dev, aggs)
    b.run()
    for q, (k, kv, v, vv) in enumerate(host):
        want = O.groupby([O.HostCol(k, O.I64, kv)], [O.HostCol(v, O.I64, vv)], aggs)
        got = b.result(q)
        assert got["n_groups"] == want["n_groups"]
        keys, kvalid = got["keys"][0]
        packed = np.where(kvalid, keys, np.iinfo(np.int64).max) if kvalid is not None else keys
        order = np.argsort(packed, kind="stable")
        assert np.array_equal(packed[order], want["keys"])
        for a in range(len(aggs)):
            vals, valid = got["aggs"][a]
            vals = vals[order]
            if a in (1, 2):
                assert np.array_equal(vals, want["i64"][a])
            else:
                wnull = want["null"][a]
                assert np.array_equal(~valid[order], wnull)
                assert np.array_equal(vals[~wnull], want["i64"][a][~wnull])


def test_batch_replans_when_tables_fill_or_data_grows(gpu, monkeypatch):
    # a tiny size multiplier makes the first plan's tables overflow: the run
    # must detect it, grow the tables and still return exact results
    monkeypatch.setenv("BOLT_BATCH_EST_MULT", "0.05")
    host, dev = _queries([(50_000, 20_000), (3000, 30)], seed0=D.SEED_BATCH + 2000)
    b = GroupByBatch(dev, AGGS)
    b.run()
    for q, (k, vi, vf) in enumerate(host):
        _check_vs_oracle(b.result(q), k, vi, vf)
    # new contents of the same columns with many more groups: the device-side
    # plan outgrows the arena, which is re-sized
    k2 = np.arange(len(host[1][0]), dtype=np.int32) * 7
    dev[1][0][0].values.copy_from(k2)
    b.run()
    _check_vs_oracle(b.result(1), k2, host[1][1], host[1][2])
    b.release()


def test_batch_integer_overflow_is_an_error(gpu):
    k = np.zeros(4, np.int32)
    v = np.array([2 ** 62, 2 ** 62, 2 ** 62, 1], np.int64)
    f = np.zeros(4, np.float32)
    dev = [([DeviceColumn.from_numpy(k)], [DeviceColumn.from_numpy(v), DeviceColumn.from_numpy(f)])]
    b = GroupByBatch(dev, AGGS)
    with pytest.raises(N.BoltTypeError, match="overflow"):
        b.run()


def test_batch_rejects_mixed_dtypes(gpu):
    a = ([DeviceColumn.from_numpy(np.zeros(3, np.int32))], [DeviceColumn.from_numpy(np.zeros(3, np.int64)),
                                                          DeviceColumn.from_numpy(np.zeros(3, np.float32))])
    c = ([DeviceColumn.from_numpy(np.zeros(3, np.int64))], [DeviceColumn.from_numpy(np.zeros(3, np.int64)),
                                                          DeviceColumn.from_numpy(np.zeros(3, np.float32))])
    with pytest.raises(N.PlanError):
        GroupByBatch([a, c], AGGS)


def test_autotune_picks_a_valid_config_and_keeps_results(gpu):
    host, dev = _queries([(200_000, 500), (60_000, 50_000), (5000, 16)], seed0=D.SEED_BATCH + 3000)
    b = GroupByBatch(dev, AGGS)
    best, ms = b.autotune(reps=1)
    assert best["block"] in (256, 512, 1024) and best["tile_rows"] >= 4096 and ms > 0
    b.run()
    for q, (k, vi, vf) in enumerate(host):
        _check_vs_oracle(b.result(q), k, vi, vf)
    t, nbytes = b.tune()
    assert t["block"] == best["block"] and nbytes > 0
