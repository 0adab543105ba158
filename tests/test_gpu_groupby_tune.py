"""GPU: the partitioned GROUP BY tier at other fan-outs than its default, and
the fan-out autotuner (bolt_groupby_autotune_partitions).  Every fan-out must
give the oracle's groups (keys, counts and integer MIN exact, float sums
within 1e-12 * sum(|x|)); the tuner's choice must be one of the partition
counts it timed and must be what the next call of that shape uses."""
import numpy as np
import pytest

import oracle as O
from craton_bolt_amd import _native as N
from craton_bolt_amd import datagen as D
from craton_bolt_amd.device import DeviceColumn
from craton_bolt_amd.ops import autotune_groupby_partitions, execute_groupby, groupby_last_plan

pytestmark = pytest.mark.gpu

AGGS = [(N.SUM, 0), (N.COUNT_STAR, 0), (N.MIN, 1)]


def _inputs(n, ranks, seed):
    k, v = D.zipf_groupby(n, n_ranks=ranks, seed=seed)
    m = (np.random.default_rng(seed).integers(-10 ** 9, 10 ** 9, n)).astype(np.int64)
    return k, v, m


def _check(res, k, v, m):
    want = O.groupby([O.HostCol(k, O.I64)], [O.HostCol(v, O.F64), O.HostCol(m, O.I64), O.HostCol(np.abs(v), O.F64)],
                     AGGS + [(N.SUM, 2)], max_groups=len(k) + 1)
    assert want["status"] == O.OK
    assert res["n_groups"] == want["n_groups"]
    assert np.array_equal(res["keys"][0][0], want["keys"])
    assert np.array_equal(res["aggs"][1][0], want["i64"][1])
    assert np.array_equal(res["aggs"][2][0], want["i64"][2])
    err = np.abs(res["aggs"][0][0] - want["f64"][0])
    assert np.all(err <= 1e-12 * want["f64"][3] + 1e-300)


@pytest.mark.parametrize("delta", [-3, -1, 1, 3])
def test_partitioned_tier_at_other_fanouts_matches_oracle(gpu, monkeypatch, delta):
    n = 3_000_000
    print(f"[case] partition fan-out delta={delta}", flush=True)
    k, v, m = _inputs(n, 500_000, 11)
    monkeypatch.setenv("BOLT_GB_PART_BITS_DELTA", str(delta))
    res = execute_groupby([DeviceColumn.from_numpy(k)], [DeviceColumn.from_numpy(v), DeviceColumn.from_numpy(m)],
                          AGGS, flags=N.GROUPBY_FORCE_PARTITIONED)
    plan = groupby_last_plan()
    assert plan["tier"] == 2
    print(f"[tune] delta={delta} partitions={plan['partitions']}", flush=True)
    _check(res, k, v, m)


def test_autotuner_choice_is_used_and_correct(gpu, monkeypatch):
    monkeypatch.delenv("BOLT_GB_PART_BITS_DELTA", raising=False)
    n = 4_000_000
    print(f"[case] partition autotune n={n}", flush=True)
    k, v, m = _inputs(n, 1_000_000, 12)
    keys = [DeviceColumn.from_numpy(k)]
    ins = [DeviceColumn.from_numpy(v), DeviceColumn.from_numpy(m)]
    parts, ms = autotune_groupby_partitions(keys, ins, AGGS, reps=1)
    print(f"[tune] best partitions={parts} ms={ms:.3f}", flush=True)
    assert parts >= 1 and parts & (parts - 1) == 0 and parts <= 4096 and ms > 0
    res = execute_groupby(keys, ins, AGGS, flags=N.GROUPBY_FORCE_PARTITIONED)
    assert groupby_last_plan()["partitions"] == parts
    _check(res, k, v, m)
