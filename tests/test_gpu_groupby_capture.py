"""GPU parity of the partitioned GROUP BY tier on inputs above 2^25 rows,
where the LDS tier also writes each spilled row's key beside its row id and
the partitioned tier's histogram and first scatter read the keys from there
(and the values by row id).  Plans over each value width, two value columns
and a nullable column; every group against the oracle (integers, counts and
MIN / MAX exact, float sums within 1e-12 * sum(|x|))."""
import numpy as np
import pytest

import oracle as O
from craton_bolt_amd import _native as N
from craton_bolt_amd.device import DeviceColumn
from craton_bolt_amd.ops import execute_groupby, groupby_last_plan

pytestmark = pytest.mark.gpu

N_ROWS = (1 << 25) + 12_345


def _keys(rng, n):
    # about 3M groups with a Zipf-like head, so the LDS tier spills a large share of the rows
    hot = rng.integers(0, 64, n)
    cold = rng.integers(0, 3_000_000, n)
    return np.where(rng.random(n) < 0.3, hot, cold).astype(np.int64) * 7919 - 12_345_678


CASES = {
    "f64_sum_count": ([(N.SUM, 0), (N.COUNT_STAR, 0)], [N.F64], [False]),
    "i32_minmax_sum_count": ([(N.MIN, 0), (N.MAX, 0), (N.SUM, 0), (N.COUNT, 0)], [N.I32], [False]),
    "f32_min_avg": ([(N.MIN, 0), (N.AVG, 0), (N.COUNT_STAR, 0)], [N.F32], [False]),
    "i64_and_f64": ([(N.SUM, 0), (N.SUM, 1)], [N.I64, N.F64], [False, False]),
    "nullable": ([(N.SUM, 0), (N.COUNT, 0)], [N.I64], [True]),
}


def _col(rng, n, dt, nullable):
    if dt == N.F64:
        v = rng.normal(size=n)
    elif dt == N.F32:
        v = rng.normal(size=n).astype(np.float32)
    elif dt == N.I32:
        v = rng.integers(-1_000_000, 1_000_000, n).astype(np.int32)
    else:
        v = rng.integers(-(1 << 40), 1 << 40, n).astype(np.int64)
    valid = rng.random(n) > 0.1 if nullable else None
    return v, valid


@pytest.mark.parametrize("case", list(CASES))
def test_partitioned_tier_with_captured_spills_matches_oracle(gpu, case):
    aggs, dts, nulls = CASES[case]
    print(f"[case] captured spills {case} n={N_ROWS}", flush=True)
    rng = np.random.default_rng(len(case))
    k = _keys(rng, N_ROWS)
    cols = [_col(rng, N_ROWS, dt, nl) for dt, nl in zip(dts, nulls)]
    dev = [DeviceColumn.from_numpy(v, dt, valid) for (v, valid), dt in zip(cols, dts)]
    res = execute_groupby([DeviceColumn.from_numpy(k)], dev, aggs, flags=N.GROUPBY_FORCE_PARTITIONED)
    plan = groupby_last_plan()
    print(f"[plan] {plan}", flush=True)
    assert plan["tier"] == 2 and plan["spilled_rows"] > 0
    odt = {N.F64: O.F64, N.F32: O.F32, N.I32: O.I32, N.I64: O.I64}
    host = [O.HostCol(v, odt[dt], valid) for (v, valid), dt in zip(cols, dts)]
    want = O.groupby([O.HostCol(k, O.I64)], host, aggs, max_groups=N_ROWS)
    assert want["status"] == O.OK
    assert res["n_groups"] == want["n_groups"]
    assert np.array_equal(res["keys"][0][0], want["keys"])
    for a, (op, c) in enumerate(aggs):
        vals, valid = res["aggs"][a]
        null = want["null"][a]
        live = ~null
        if valid is not None:
            assert np.array_equal(~valid, null), (case, a)
        if op in (N.COUNT, N.COUNT_STAR):
            assert np.array_equal(vals, want["i64"][a]), (case, a)
        elif op in (N.MIN, N.MAX) or (op == N.SUM and dts[c] in (N.I32, N.I64)):
            if dts[c] in (N.F64, N.F32):
                assert np.array_equal(vals.astype(np.float64)[live], want["f64"][a][live]), (case, a)
            else:
                assert np.array_equal(vals.astype(np.int64)[live], want["i64"][a][live]), (case, a)
        else:
            v, va = cols[c]
            habs = O.HostCol(np.abs(v.astype(np.float64)), O.F64, va)
            sabs = O.groupby([O.HostCol(k, O.I64)], [habs], [(N.SUM, 0)], max_groups=N_ROWS)["f64"][0]
            got = vals.astype(np.float64)
            if op == N.AVG:
                cnt = want["i64"][2].astype(np.float64)  # COUNT(*) of the group (no NULLs in this case)
                tol = 1e-12 * sabs / np.maximum(cnt, 1) + 1e-300
            else:
                tol = 1e-12 * sabs + 1e-300
            err = np.abs(got - want["f64"][a])
            assert np.all(err[live] <= tol[live]), (case, a, float(np.max(err[live] / tol[live])))
