"""GPU parity: expression programs feeding GROUP BY (bolt_groupby_expr,
execute_groupby_with_pre, src/exec/groupby_with_pre.rs:222) and the fused
join probe (bolt_join_probe_aggregate_expr), against the oracle's
composition of the same steps: the oracle's program interpreter materialises
the values, the rows whose predicate value is true and not NULL are kept,
then the oracle's GROUP BY / host join fold runs over the kept values.
Integer, Decimal128 and count results bit-exact; float sums within
1e-12 * sum(|x|) (the atomics' order is not fixed)."""
import math

import numpy as np
import pytest

import oracle as O
from craton_bolt_amd import _native as N
from craton_bolt_amd import datagen as D
from craton_bolt_amd.device import DeviceColumn
from craton_bolt_amd.ops import JoinTable, Program, execute_groupby_expr

from gpu_util import both, dec_array

pytestmark = pytest.mark.gpu


def _star_inputs(n, seed):
    rng = np.random.default_rng(seed)
    cents = rng.integers(10_000, 100_000_000, n)
    cols = [
        both(rng.integers(0, 5000, n).astype(np.int32), N.I32, rng.random(n) > 0.05),      # 0 fk
        both(dec_array([int(c) for c in cents]), N.DEC128, None, scale=2),                 # 1 price
        both(rng.uniform(1, 50, n), N.F64, rng.random(n) > 0.02),                          # 2 qty
        both(rng.integers(10_957, 10_957 + 2557, n).astype(np.int32), N.DATE32),           # 3 date (2000-2006)
        both(rng.integers(0, 1 << 45, n), N.TIMESTAMP, None, scale=N.TS_MS),               # 4 ts
    ]
    return cols


def _program(two_keys=False):
    """values: 0 = EXTRACT(YEAR FROM date), 1 = DATE_TRUNC('month', ts) < const, 2 = CAST(price AS DOUBLE) * qty,
    3 = CAST(price AS Decimal128(38, 0)) (rescale down, truncating), 4 = fk, 5 = the predicate
    qty > 10 AND EXTRACT(MONTH FROM date) <= 6; with two_keys also 6 = CAST(year AS INT) (two GROUP BY keys
    must both be 32-bit)."""
    p = Program()
    fk, price, qty, date, ts = (p.load(0, N.I32), p.load(1, N.DEC128), p.load(2, N.F64), p.load(3, N.DATE32),
                                p.load(4, N.TIMESTAMP))
    year = p.extract("year", date)
    p.store(year, 0)
    p.store(p.binary("<", p.date_trunc("month", ts), p.const(1 << 44, N.TIMESTAMP, unit=N.TS_MS), N.TIMESTAMP), 1)
    p.store(p.binary("*", p.cast(price, N.F64, from_scale=2), qty, N.F64), 2)
    p.store(p.cast(price, N.DEC128, scale=0, from_scale=2), 3)
    p.store(fk, 4)
    pred = p.binary("and", p.binary(">", qty, p.const(10.0, N.F64), N.F64),
                    p.binary("<=", p.extract("month", date), p.const(6, N.I64), N.I64), N.BOOL)
    p.store(pred, 5)
    if two_keys:
        p.store(p.cast(year, N.I32), 6)
    return p


VALUE_DTYPES = [N.I64, N.BOOL, N.F64, N.DEC128, N.I32, N.BOOL]
_TO_O = {N.I64: O.I64, N.BOOL: O.BOOL, N.F64: O.F64, N.DEC128: O.DEC128, N.I32: O.I32}


@pytest.mark.parametrize("n", [10_000, 2_000_003])
@pytest.mark.parametrize("key_vals", [[0], [4], [4, 6]])
def test_groupby_over_a_program_matches_oracle(gpu, n, key_vals):
    print(f"[case] groupby_expr n={n} keys={key_vals}", flush=True)
    cols = _star_inputs(n, n + len(key_vals))
    two = len(key_vals) == 2
    prog = _program(two)
    nv = 7 if two else 6
    dts = VALUE_DTYPES + ([N.I32] if two else [])
    aggs = [(N.SUM, 2), (N.SUM, 3), (N.COUNT_STAR, 0), (N.COUNT, 2), (N.MAX, 4)]
    res = execute_groupby_expr([d for d, _ in cols], prog, nv, key_vals, aggs, filter_value=5)
    st, vals, valid = O.project([h for _, h in cols], prog.ops, dts, n, [True] * nv)
    assert st == O.OK
    keep = (vals[5] != 0) & valid[5]
    hv = [O.HostCol(np.ascontiguousarray(v[keep]), _TO_O[dt], va[keep]) for v, va, dt in zip(vals, valid, dts)]
    hk = [hv[k] for k in key_vals]
    want = O.groupby(hk, hv, aggs, max_groups=int(keep.sum()) + 1)
    assert want["status"] == O.OK
    assert res["n_groups"] == want["n_groups"]
    # keys, rebuilt as the oracle's packed key (NULL group = i64::MAX)
    if two:
        a = res["keys"][0][0].view(np.uint32).astype(np.uint64)
        b = res["keys"][1][0].view(np.uint32).astype(np.uint64)
        got = ((a << np.uint64(32)) | b).view(np.int64)
    else:
        kv, kvalid = res["keys"][0]
        got = kv.view(np.uint32).astype(np.int64) if kv.dtype == np.int32 else kv.view(np.int64)
        if kvalid is not None:
            got = np.where(kvalid, got, np.int64(2 ** 63 - 1))
    assert np.array_equal(got, want["keys"])
    null = [want["null"][i] for i in range(len(aggs))]
    # SUM(f64) within 1e-12 * sum(|x|) of the group
    habs = O.HostCol(np.abs(hv[2].values), O.F64, valid[2][keep])
    sabs = O.groupby(hk, [habs], [(N.SUM, 0)], max_groups=int(keep.sum()) + 1)["f64"][0]
    live = ~null[0]
    err = np.abs(res["aggs"][0][0] - want["f64"][0])
    assert np.all(err[live] <= 1e-12 * sabs[live] + 1e-300)
    dec = res["aggs"][1][0]
    assert np.array_equal(dec[:, 0], want["i64"][1]) and np.array_equal(dec[:, 1], want["hi"][1])
    assert np.array_equal(res["aggs"][2][0], want["i64"][2])
    assert np.array_equal(res["aggs"][3][0], want["i64"][3])
    live4 = ~null[4]
    assert np.array_equal(res["aggs"][4][0].astype(np.int64)[live4], want["i64"][4][live4])


@pytest.mark.parametrize("n_build,n_probe", [(1000, 20_000), (1 << 16, 3_000_001)])
def test_join_probe_over_a_program_matches_oracle(gpu, n_build, n_probe):
    """probe key = fk + 0 computed in the program, WHERE amount * 2 > 20 AND amount IS NOT NULL-safe,
    SUM(payload), SUM(amount * 1.5), COUNT(*), MIN(payload)."""
    print(f"[case] join_expr build={n_build} probe={n_probe}", flush=True)
    bk, bp = D.join_build(n_build)
    fk, amt = D.join_probe(n_build, n_probe)
    rng = np.random.default_rng(n_probe)
    avalid = rng.random(n_probe) > 0.03
    cols = [both(fk, N.I64), both(amt, N.F64, avalid)]
    p = Program()
    k, a = p.load(0, N.I64), p.load(1, N.F64)
    p.store(p.binary("+", k, p.const(0, N.I64), N.I64), 0)
    p.store(p.binary(">", p.binary("*", a, p.const(2.0, N.F64), N.F64), p.const(20.0, N.F64), N.F64), 1)
    p.store(p.binary("*", a, p.const(1.5, N.F64), N.F64), 2)
    t = JoinTable(DeviceColumn.from_numpy(bk), carry=DeviceColumn.from_numpy(bp))
    aggs = [(N.SUM, N.JOIN_CARRY), (N.SUM, 2), (N.COUNT_STAR, 0), (N.MIN, N.JOIN_CARRY)]
    got = t.probe_aggregate_expr([d for d, _ in cols], p, 3, 0, aggs, filter_value=1)
    st, vals, valid = O.project([h for _, h in cols], p.ops, [N.I64, N.BOOL, N.F64], n_probe, [True] * 3)
    assert st == O.OK
    keep = np.nonzero((vals[1] != 0) & valid[1])[0]
    ol, orr = O.inner_join(O.HostCol(bk, O.I64), O.HostCol(np.ascontiguousarray(vals[0][keep]), O.I64))
    assert got[2] == len(ol)
    assert got[0] == int(bp[ol].astype(np.int64).sum())
    assert got[3] == (int(bp[ol].min()) if len(ol) else None)
    v = vals[2][keep][orr]
    assert math.isclose(got[1], math.fsum(v), rel_tol=0, abs_tol=1e-12 * float(np.abs(v).sum()) + 1e-300)


def test_program_operator_errors(gpu):
    cols = _star_inputs(100, 1)
    prog = _program()
    with pytest.raises(N.BoltError, match="not a Bool value"):
        execute_groupby_expr([d for d, _ in cols], prog, 6, [0], [(N.COUNT_STAR, 0)], filter_value=2)
    with pytest.raises(N.BoltError, match="key value 9"):
        execute_groupby_expr([d for d, _ in cols], prog, 6, [9], [(N.COUNT_STAR, 0)])
    bk, bp = D.join_build(100)
    t = JoinTable(DeviceColumn.from_numpy(bk), carry=DeviceColumn.from_numpy(bp))
    with pytest.raises(N.BoltError, match="key value"):
        t.probe_aggregate_expr([d for d, _ in cols], prog, 6, 7, [(N.COUNT_STAR, 0)])
