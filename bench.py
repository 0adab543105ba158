#!/usr/bin/env python3
"""Benchmark of the GPU relational-operator path on one MI355X.

Default workload (one "step" = one query over device-resident tables):
  fused filter + hash join + aggregate —
    build  : 16M rows (unique int64 key, int32 payload uniform [0, 1e6))
    probe  : 256M rows (int64 fk, 50% build keys / 50% misses; float64 amount U[0,100))
    query  : WHERE amount > 10, INNER JOIN, SUM(payload), SUM(amount), COUNT(*)
  metric: probe rows/s; the dominant kernel is the fused probe.

The default run ("all") prints that line with a `sub_records` object beside
it: the other configurations of the benchmark, each with its rate, roofline
(operator-level and dominant kernel), a full-size correctness check against
the oracle, float error against the FP64 compensated sum where floats are
summed, and a single-threaded CPU baseline (median of 5 runs after one
warm-up, on a bounded sample, host recorded):
  groupby_1M / groupby_256M : plain GROUP BY (int32 key over 1000 values, int64 value)
  sort_256M                 : ORDER BY radix sort of (int64 key, u32 row id) pairs
  groupby_zipf_256M         : GROUP BY int64 Zipf(1.1) key over 10M ranks, SUM(float64), COUNT(*)
  batch_512                 : 512 batched GROUP BY queries vs 512 plain ones
  serve_200M                : served star queries over a 200M-row fact table

Prints one JSON line.  Inputs are generated on the host (seeded), uploaded
once, and resident in HBM before the timed region starts.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0


def _parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", default="all", choices=["all", "join", "sort", "topk", "expr", "h2o", "groupby",
                                                            "groupby-zipf", "calib", "batch", "serve"])
    p.add_argument("--limit", type=int, default=100, help="topk workload: LIMIT")
    p.add_argument("--key-dtype", default="int64", choices=["int64", "int32", "date32"],
                   help="topk workload: the ORDER BY key's type")
    p.add_argument("--fact-rows", type=int, default=200_000_000, help="serve workload: fact table rows")
    p.add_argument("--serve-queries", type=int, default=64, help="serve workload: queries per measurement")
    p.add_argument("--slots", type=int, default=8, help="serve workload: concurrent queries")
    p.add_argument("--queries", type=int, default=512, help="queries of the batch workload")
    p.add_argument("--no-autotune", action="store_true", help="batch workload: keep the default parameters")
    p.add_argument("--n", type=int, default=256 << 20, help="rows for the sort / groupby / calib workloads")
    p.add_argument("--n-build", type=int, default=16 << 20)
    p.add_argument("--n-probe", type=int, default=256 << 20)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-sample", type=int, default=8 << 20, help="probe rows timed on the CPU")
    p.add_argument("--no-check", action="store_true", help="skip the full-size checks against the oracle")
    return p.parse_args()


def _host_info() -> dict:
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def _cpu_median(fn, reps: int = 5) -> float:
    """Median wall time (s) of `reps` runs of fn after one warm-up run."""
    fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def _cpu_line(value: float, unit: str, sample: str, median_s: float) -> dict:
    return {"value": value, "unit": unit, "cores": 1, "kind": "port", "sample": sample, "median_s": median_s,
            "runs": "median of 5 after 1 warm-up", "host": _host_info()}


def cpu_baseline_join(bk, bp, fk, amt, n_probe_full: int, sample: int) -> dict:
    """The oracle's restatement of the reference host path (filter -> host
    hash join -> scalar aggregates), single-threaded, on a bounded sample:
    the full 16M-row build plus `sample` probe rows; the probe rate is
    extrapolated to the full probe side."""
    import oracle as O

    O.lib()
    t_build = _cpu_median(lambda: O.filter_join_agg(bk, bp, fk[:0], amt[:0], 10.0))
    t_both = _cpu_median(lambda: O.filter_join_agg(bk, bp, fk[:sample], amt[:sample], 10.0))
    t_probe = max(t_both - t_build, 1e-9)
    t_full = t_build + t_probe * (n_probe_full / sample)
    return _cpu_line(n_probe_full / t_full, "probe rows/s",
                     f"C restatement of join.rs:169 host join + aggregate.rs:154; full {len(bk)}-row build "
                     f"({t_build:.2f}s) + {sample} probe rows ({t_probe:.2f}s), extrapolated to {n_probe_full} rows",
                     t_both)


def bench_join(args) -> dict:
    from craton_bolt_amd import _native as N
    from craton_bolt_amd import datagen as D
    from craton_bolt_amd.device import DeviceColumn, Event, Stream
    from craton_bolt_amd.ops import JoinTable, Term

    nb, npb = args.n_build, args.n_probe
    t0 = time.perf_counter()
    bk, bp = D.join_build(nb)
    fk, amt = D.join_probe(nb, npb)
    print(f"[bench] generated build={nb} probe={npb} in {time.perf_counter() - t0:.1f}s", file=sys.stderr,
          flush=True)
    t0 = time.perf_counter()
    d_bk, d_bp = DeviceColumn.from_numpy(bk), DeviceColumn.from_numpy(bp)
    d_fk, d_amt = DeviceColumn.from_numpy(fk), DeviceColumn.from_numpy(amt)
    print(f"[bench] uploaded in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)

    stream = Stream()
    aggs = [(N.SUM, N.JOIN_CARRY), (N.SUM, 1), (N.COUNT_STAR, 0)]
    terms = [Term(1, N.GT, 10.0)]
    cols = [d_fk, d_amt]
    ev = [Event() for _ in range(3)]
    kev = [Event(), Event()]  # recorded by the library around the fused probe kernel itself

    def step(record: bool):
        if record:
            ev[0].record(stream)
        t = JoinTable(d_bk, carry=d_bp, stream=stream)
        if record:
            ev[1].record(stream)
        res = t.probe_aggregate(cols, 0, aggs, terms)  # synchronises the stream
        if record:
            ev[2].record(stream)
        t.release()
        return res

    for _ in range(args.warmup):
        res = step(False)
    stream.synchronize()
    build_ms, probe_ms, kernel_ms = [], [], []
    N.lib().bolt_profile_events(kev[0].handle, kev[1].handle)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = step(True)
        ev[2].synchronize()
        build_ms.append(ev[0].elapsed_ms(ev[1]))
        probe_ms.append(ev[1].elapsed_ms(ev[2]))
        kernel_ms.append(kev[0].elapsed_ms(kev[1]))
    stream.synchronize()
    wall = time.perf_counter() - t0
    N.lib().bolt_profile_events(None, None)
    ms_per_step = wall * 1000.0 / args.steps
    kern = float(np.mean(kernel_ms))
    # algorithmic bytes of the probe launch: fk (8 B) + amount (8 B) per probe row
    probe_bytes = npb * 16
    achieved = probe_bytes / (kern * 1e-3) / 1e9
    print(f"[bench] result sum_payload={res[0]} sum_amount={res[1]:.6f} count={res[2]} "
          f"build_ms={np.mean(build_ms):.3f} probe_call_ms={np.mean(probe_ms):.3f} probe_kernel_ms={kern:.3f}",
          file=sys.stderr, flush=True)
    traffic = _traffic("join", kern)

    line = {
        "metric": "rows/s + % HBM peak per operator; batched/served vs plain queries/s, p99 ms",
        "value": npb / (ms_per_step * 1e-3),
        "unit": "probe rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {
            "workload": "fused filter + hash join + aggregate: build 16M (unique int64 key, int32 payload), "
                        "probe 256M (int64 fk 50% hits, float64 amount); WHERE amount > 10, SUM(payload), "
                        "SUM(amount), COUNT(*); one step = build + probe",
            "n_build": nb,
            "n_probe": npb,
            "build_ms": float(np.mean(build_ms)),
            "probe_call_ms": float(np.mean(probe_ms)),
            "probe_kernel_ms": kern,
            "result": {"sum_payload": res[0], "sum_amount": res[1], "count": res[2]},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "bolt_jit_kernel (hipRTC fused filter + probe + aggregate), timed by HIP events on its stream",
            "bytes_per_launch": probe_bytes,
        },
    }
    if not args.no_check:
        import oracle as O

        t0 = time.perf_counter()
        want = O.filter_join_agg(bk, bp, fk, amt, 10.0)
        rel = abs(res[1] - want["sum_amount"]) / max(abs(want["sum_amount"]), 1e-300)
        line["check"] = {"ok": bool(want["status"] == 0 and res[0] == want["sum_payload"] and res[2] == want["count"]
                                    and rel <= 1e-12),
                         "oracle": "filter (compact.rs:146) -> host hash join (join.rs:169) -> SUM / SUM / COUNT "
                                   "(aggregate.rs:154), full size",
                         "sum_payload_exact": res[0] == want["sum_payload"], "count_exact": res[2] == want["count"],
                         "oracle_s": time.perf_counter() - t0}
        line["float_error"] = {"sum_amount_rel_err_vs_fp64_neumaier": rel, "tolerance": 1e-12}
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline_join(bk, bp, fk, amt, npb, min(args.cpu_sample, npb))
    return line


def _traffic(workload: str, kernel_ms: float):
    """HBM bytes per launch of the workload's dominant kernel, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs (profiles/traffic.json,
    counters already scaled by the copy-kernel calibration), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)[workload]
    except (OSError, KeyError, ValueError):
        return None
    return float(t["traffic_bytes"])


def _timed(stream, fn, steps, warmup):
    from craton_bolt_amd.device import Event

    for _ in range(warmup):
        fn()
    stream.synchronize()
    e0, e1 = Event(), Event()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    e1.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_ms(e1) / steps, wall * 1000.0 / steps


def _line(metric_value, unit, args, ms, dtype, config, roofline):
    return {
        "metric": "rows/s + % HBM peak per operator; batched/served vs plain queries/s, p99 ms",
        "value": metric_value, "unit": unit, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
        "data": "synthetic", "config": config, "roofline": roofline,
    }


def _roof(bytes_per_launch, kernel_ms, kernel):
    achieved = bytes_per_launch / (kernel_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": kernel, "bytes_per_launch": bytes_per_launch}


def bench_sort(args) -> dict:
    """256M (int64 key, u32 row id) pairs; 200 B/key model for 64-bit LSD."""
    from craton_bolt_amd import datagen as D
    from craton_bolt_amd.device import DeviceBuffer, Stream
    from craton_bolt_amd.ops import sort_pairs_i64

    n = args.n
    key, rid = D.sort_pairs(n)
    dk, dv = DeviceBuffer.from_numpy(key), DeviceBuffer.from_numpy(rid)
    s = Stream()
    from craton_bolt_amd import _native as N
    from craton_bolt_amd.device import Event

    kev = [Event(), Event()]
    kms = []

    def one():
        sort_pairs_i64(dk, dv, n, stream=s)
        kev[1].synchronize()
        kms.append(kev[0].elapsed_ms(kev[1]))

    N.lib().bolt_profile_events(kev[0].handle, kev[1].handle)
    ms, wall = _timed(s, one, args.steps, args.warmup)
    N.lib().bolt_profile_events(None, None)
    ms = float(np.mean(kms[args.warmup:]))  # the sort itself, on its stream
    model = n * 200
    return _line(n / (ms * 1e-3), "keys/s", args, ms, "int64",
                 {"workload": f"ORDER BY radix sort: {n} (int64 key uniform over the full range, u32 row id)",
                  "n": n, "compulsory_bytes": n * 24, "model_bytes": model,
                  "sort_ms_on_stream": ms, "wall_ms_per_call": wall},
                 _roof(model, ms, "hybrid radix sort (histogram, 2 onesweep passes, LDS bucket sorts), 200 B/key model"))


def bench_topk(args) -> dict:
    """ORDER BY key LIMIT k over 256M keys (int64 over the full range by default; int32 uniform, or
    Date32 over 20000 days): one bolt_sort_limit call per step (it synchronises), against one full
    bolt_sort_indices + slice."""
    from craton_bolt_amd import _native as N
    from craton_bolt_amd import datagen as D
    from craton_bolt_amd.device import DeviceColumn, Stream
    from craton_bolt_amd.ops import execute_sort, execute_sort_limit

    n, k = args.n, args.limit
    key, _ = D.sort_pairs(n)
    kd = args.key_dtype
    if kd == "int32":
        key = (key >> 32).astype(np.int32)
    elif kd == "date32":
        key = (key.view(np.uint64) % np.uint64(20_000)).astype(np.int32)
    dt = {"int64": N.I64, "int32": N.I32, "date32": N.DATE32}[kd]
    width = 8 if kd == "int64" else 4
    col = DeviceColumn.from_numpy(key, dt)
    s = Stream()
    ms, wall = _timed(s, lambda: execute_sort_limit([col], k, 0, stream=s), args.steps, args.warmup)
    ms = wall  # the call synchronises the host: its wall time is the step
    t0 = time.perf_counter()
    execute_sort([col], stream=s)
    full_ms = (time.perf_counter() - t0) * 1000.0
    what = {"int64": "int64 key (uniform over the full range)", "int32": "int32 key (uniform)",
            "date32": "Date32 key (20000 distinct days)"}[kd]
    return _line(n / (ms * 1e-3), "rows/s", args, ms, kd,
                 {"workload": f"ORDER BY {what} LIMIT {k}: {n} rows", "n": n, "limit": k,
                  "full_sort_then_slice_ms": full_ms},
                 _roof(n * width, ms, f"radix select (12-bit histogram passes) + candidate collect + candidate "
                                      f"sort; {width} B/row (one read of the key) is the compulsory traffic"))


def bench_expr(args) -> dict:
    """A TPC-H Q6 shape over 256M rows: SUM(price * disc) WHERE qty < 24 AND disc BETWEEN 0.05 AND 0.07
    AND price + qty > 100, one bolt_aggregate_expr call (one fused generated kernel) per step; the
    composed form (bolt_project of the two values, then bolt_aggregate on them) timed beside it."""
    from craton_bolt_amd import _native as N
    from craton_bolt_amd.device import DeviceColumn, Event, Stream
    from craton_bolt_amd.ops import Program, Term, execute_aggregate, execute_aggregate_expr, execute_projection

    n = args.n
    rng = np.random.default_rng(6)
    price = DeviceColumn.from_numpy(rng.uniform(1, 1000, n), N.F64)
    disc = DeviceColumn.from_numpy(np.round(rng.uniform(0, 0.1, n), 2), N.F64)
    qty = DeviceColumn.from_numpy(rng.integers(1, 50, n).astype(np.int32), N.I32)
    p = Program()
    pr, d, q = p.load(0, N.F64), p.load(1, N.F64), p.load(2, N.I32)
    p.store(p.binary("*", pr, d, N.F64), 0)
    c1 = p.binary("<", q, p.const(24, N.I32), N.I32)
    c2 = p.binary(">=", d, p.const(0.05, N.F64), N.F64)
    c3 = p.binary("<=", d, p.const(0.07, N.F64), N.F64)
    c4 = p.binary(">", p.binary("+", pr, p.cast(q, N.F64), N.F64), p.const(100.0, N.F64), N.F64)
    p.store(p.binary("and", p.binary("and", c1, c2, N.BOOL), p.binary("and", c3, c4, N.BOOL), N.BOOL), 1)
    cols = [price, disc, qty]
    aggs = [(1, 0), (3, 0)]
    s = Stream()
    kev = [Event(), Event()]
    kms = []

    def one():
        execute_aggregate_expr(cols, p, 2, aggs, 1, stream=s)
        kms.append(kev[0].elapsed_ms(kev[1]))

    N.lib().bolt_profile_events(kev[0].handle, kev[1].handle)
    _, wall = _timed(s, one, args.steps, args.warmup)
    N.lib().bolt_profile_events(None, None)
    kern = float(np.mean(kms[args.warmup:]))
    res = execute_aggregate_expr(cols, p, 2, aggs, 1, stream=s)
    for rep in range(2):  # the second call is timed (the first compiles the projection kernel)
        t0 = time.perf_counter()
        vals = execute_projection(cols, p, [N.F64, N.BOOL], stream=s)
        comp = execute_aggregate(vals, aggs, [Term(1, 5, 1)], stream=s)
        composed_ms = (time.perf_counter() - t0) * 1000.0
        del vals
    per_row = 8 + 8 + 4
    return _line(n / (wall * 1e-3), "rows/s", args, wall, "f64",
                 {"workload": f"TPC-H Q6 shape: {n} rows (price f64, disc f64, qty i32); SUM(price * disc), COUNT(*) "
                              "WHERE qty < 24 AND disc BETWEEN 0.05 AND 0.07 AND price + qty > 100",
                  "n": n, "fused_kernel_ms": kern, "wall_ms_per_call": wall,
                  "composed_project_then_aggregate_ms": composed_ms, "result": res,
                  "composed_agrees": comp[1] == res[1] and abs(comp[0] - res[0]) <= 1e-9 * abs(res[0])},
                 _roof(n * per_row, kern, "fused program + predicate + aggregate kernel (20 B/row read once)"))


# h2o.ai groupby q1..q5 (benches/olap_benchmarks.rs:77-90) and the reference's published times for them
# (docs/BENCHMARKS.md:28-32: end-to-end Engine::sql on an RTX 2060, host -> device upload included)
H2O_QUERIES = {
    "q1": (["id1"], ["v1"], [(1, 0)], 51.4),
    "q2": (["id2"], ["v1", "v2"], [(1, 0), (1, 1)], 384.0),
    "q3": (["id1", "id2"], ["v1"], [(1, 0)], 219.0),
    "q4": (["id1"], ["v1", "v2", "v3"], [(6, 0), (6, 1), (6, 2)], 70.5),
    "q5": (["id3"], ["v1"], [(1, 0)], 237.0),
}


def bench_h2o(args) -> dict:
    """The reference's published benchmark: h2o groupby q1..q5 over 10M rows.  Per query: the GROUP BY
    with its columns resident in HBM, and end to end as the reference measures it (upload of the query's
    columns from pageable host memory, the GROUP BY in key order, the result copied back)."""
    from craton_bolt_amd import _native as N
    from craton_bolt_amd import datagen as D
    from craton_bolt_amd.device import DeviceColumn, Stream
    from craton_bolt_amd.ops import execute_groupby

    n = 10_000_000
    host = {c: D.h2o_column(c, n) for c in ("id1", "id2", "id3", "v1", "v2", "v3")}
    dt = {c: (N.I32 if c.startswith("id") else N.F64) for c in host}
    dev = {c: DeviceColumn.from_numpy(host[c], dt[c]) for c in host}
    s = Stream()
    per = {}
    for q, (keys, vals, aggs, ref_ms) in H2O_QUERIES.items():
        def resident():
            execute_groupby([dev[k] for k in keys], [dev[v] for v in vals], aggs, stream=s)

        def end_to_end():
            cols = {c: DeviceColumn.from_numpy(host[c], dt[c]) for c in keys + vals}
            r = execute_groupby([cols[k] for k in keys], [cols[v] for v in vals], aggs, stream=s)
            return r["n_groups"]

        for _ in range(args.warmup):
            resident()
            end_to_end()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            resident()
        res_ms = (time.perf_counter() - t0) * 1000.0 / args.steps
        t0 = time.perf_counter()
        for _ in range(args.steps):
            g = end_to_end()
        e2e_ms = (time.perf_counter() - t0) * 1000.0 / args.steps
        per[q] = {"groups": g, "resident_ms": res_ms, "end_to_end_ms": e2e_ms, "reference_ms": ref_ms,
                  "speedup_end_to_end": ref_ms / e2e_ms}
    tot_res = sum(v["resident_ms"] for v in per.values())
    tot_e2e = sum(v["end_to_end_ms"] for v in per.values())
    tot_ref = sum(v["reference_ms"] for v in per.values())
    line = _line(5 * n / (tot_res * 1e-3), "rows/s", args, tot_res / 5, "f64",
                 {"workload": "h2o.ai groupby q1..q5 over 10M rows (benches/olap_benchmarks.rs generator), "
                              "output in key order; value = 5 x 10M rows / sum of resident query times",
                  "n": n, "queries": per, "sum_end_to_end_ms": tot_e2e, "sum_reference_ms": tot_ref,
                  "reference_setup": "end-to-end Engine::sql, RTX 2060 (docs/BENCHMARKS.md:12-32)"},
                 None)
    line["vs_baseline"] = tot_ref / tot_e2e  # published end-to-end times over ours, same protocol
    return line


def bench_groupby(args, zipf: bool) -> dict:
    """One bolt_groupby call per step, result left in HBM (the table is
    released, not downloaded).  The roofline is the LDS-tier pass (pass A),
    timed by the library's own events on the operator's stream."""
    import ctypes as C

    from craton_bolt_amd import _native as N
    from craton_bolt_amd import datagen as D
    from craton_bolt_amd.device import DeviceColumn, Event, Stream, columns_array
    from craton_bolt_amd.ops import _aggs

    n = args.n
    s = Stream()
    if zipf:
        k, v = D.zipf_groupby(n, n_ranks=10_000_000)
        aggs = [(N.SUM, 0), (N.COUNT_STAR, 0)]
        desc = f"high-cardinality GROUP BY: {n} rows, int64 key Zipf(1.1) over 10M ranks, SUM(float64), COUNT(*)"
        per_row, dtype, flags = 16, "f64", 0
    else:
        k, v = D.groupby_small(n)
        aggs = [(N.SUM, 0), (N.COUNT_STAR, 0), (N.MIN, 0), (N.MAX, 0)]
        desc = (f"plain GROUP BY: {n} rows, int32 key over 1000 values, int64 value; SUM, COUNT(*), MIN, MAX; "
                f"output ordered by key")
        per_row, dtype, flags = 12, "int64", N.GROUPBY_SORTED
    dk, dv = DeviceColumn.from_numpy(k), DeviceColumn.from_numpy(v)
    ka, ia, ag = columns_array([dk]), columns_array([dv]), _aggs(aggs)
    L = N.lib()
    ev = [Event(), Event()]
    kev = [Event(), Event()]  # recorded by the library around pass A
    groups = {}

    def one(record: bool):
        tbl = C.c_void_p()
        if record:
            ev[0].record(s)
        N.check(L.bolt_groupby(s.handle, ka, 1, ia, 1, ag, len(aggs), None, flags, C.byref(tbl)), "bolt_groupby")
        if record:
            ev[1].record(s)
        nr = C.c_int64()
        L.bolt_table_num_rows(tbl, C.byref(nr))
        groups["g"] = nr.value
        L.bolt_table_release(tbl)

    for _ in range(args.warmup):
        one(False)
    s.synchronize()
    op_ms, k_ms = [], []
    L.bolt_profile_events(kev[0].handle, kev[1].handle)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        one(True)
        ev[1].synchronize()
        op_ms.append(ev[0].elapsed_ms(ev[1]))
        k_ms.append(kev[0].elapsed_ms(kev[1]))
    s.synchronize()
    wall = (time.perf_counter() - t0) * 1000.0 / args.steps
    L.bolt_profile_events(None, None)
    ms, kern = float(np.mean(op_ms)), float(np.mean(k_ms))
    g = groups["g"]
    roof = _roof(n * per_row, kern, "GROUP BY LDS-tier pass (keys + inputs read once)")
    roof["traffic"] = _traffic("groupby-zipf" if zipf else "groupby", kern)
    return _line(n / (ms * 1e-3), "rows/s", args, ms, dtype,
                 {"workload": desc, "n": n, "groups": g, "operator_ms_on_stream": ms, "pass_a_kernel_ms": kern,
                  "wall_ms_per_call": wall}, roof)


def bench_calib(args) -> dict:
    """Known-size kernels for scaling the FETCH_SIZE / WRITE_SIZE counters:
    SUM over 2^28 float64 reads 2 GiB; an always-true compaction of the same
    column reads 2 GiB and writes 2 GiB."""
    from craton_bolt_amd import _native as N
    from craton_bolt_amd.device import DeviceColumn, Stream
    from craton_bolt_amd.ops import Term, compact_columns, execute_aggregate

    n = 1 << 28
    x = np.arange(n, dtype=np.float64)
    d = DeviceColumn.from_numpy(x)
    s = Stream()
    ms_sum, _ = _timed(s, lambda: execute_aggregate([d], [(N.SUM, 0)], stream=s), args.steps, args.warmup)
    ms_cp, _ = _timed(s, lambda: compact_columns([d], [Term(0, N.IS_NOT_NULL)], stream=s), args.steps, args.warmup)
    return _line(n * 8 / (ms_sum * 1e-3), "bytes/s", args, ms_sum, "f64",
                 {"workload": "calibration: SUM over 2^28 float64 (2 GiB read), always-true compaction (2 GiB read "
                              "+ 2 GiB write)", "sum_ms": ms_sum, "copy_ms": ms_cp, "bytes": n * 8},
                 _roof(n * 8, ms_sum, "scalar_agg_kernel"))


def bench_batch(args, keep: bool = False) -> dict:
    """512 independent GROUP BY k, SUM(v_i64), SUM(v_f32), COUNT(*) queries
    (rows log-uniform 1e4..4e6, Zipf(1.1) int32 keys over a log-uniform
    16..1e6 cardinality): one graph-captured persistent batch versus the same
    queries as 512 plain per-query bolt_groupby calls."""
    import ctypes as C

    from craton_bolt_amd import _native as N
    from craton_bolt_amd import datagen as D
    from craton_bolt_amd.device import DeviceColumn, Event, Stream, columns_array
    from craton_bolt_amd.ops import GroupByBatch, _aggs

    shapes = D.batch_shapes(args.queries)
    t0 = time.perf_counter()
    queries = []
    total_rows = 0
    for q, (n, card) in enumerate(shapes):
        k, vi, vf = D.zipf_small(n, card, seed=D.SEED_BATCH + q)
        queries.append(([DeviceColumn.from_numpy(k)], [DeviceColumn.from_numpy(vi), DeviceColumn.from_numpy(vf)]))
        total_rows += n
    print(f"[bench] generated + uploaded {args.queries} queries ({total_rows} rows) in {time.perf_counter() - t0:.1f}s",
          file=sys.stderr, flush=True)
    aggs = [(N.SUM, 0), (N.SUM, 1), (N.COUNT_STAR, 0)]
    s = Stream()

    # plain path: one bolt_groupby per query (each plans, sizes and extracts on its own)
    ag = _aggs(aggs)
    plain_args = [(columns_array(k), columns_array(i)) for k, i in queries]
    L = N.lib()

    def plain_all(collect=None):
        for q, (ka, ia) in enumerate(plain_args):
            tbl = C.c_void_p()
            N.check(L.bolt_groupby(s.handle, ka, 1, ia, 2, ag, len(aggs), None, 0, C.byref(tbl)), "bolt_groupby")
            if collect is not None:
                nr = C.c_int64()
                L.bolt_table_num_rows(tbl, C.byref(nr))
                collect.append(nr.value)
            L.bolt_table_release(tbl)
        s.synchronize()

    plain_groups = []
    plain_all(plain_groups)
    t0 = time.perf_counter()
    plain_reps = max(1, min(3, args.steps))
    for _ in range(plain_reps):
        plain_all()
    plain_ms = (time.perf_counter() - t0) * 1000.0 / plain_reps
    print(f"[bench] plain: {plain_ms:.1f} ms per {args.queries} queries", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    b = GroupByBatch(queries, aggs, stream=s)
    create_ms = (time.perf_counter() - t0) * 1000.0
    tuned = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        tuned, tuned_ms = b.autotune(reps=2)
        print(f"[bench] autotune: {tuned} {tuned_ms:.2f} ms ({time.perf_counter() - t0:.1f}s)", file=sys.stderr,
              flush=True)
    for _ in range(args.warmup):
        b.run()
    e0, e1 = Event(), Event()
    t0 = time.perf_counter()
    e0.record(s)
    for _ in range(args.steps):
        b.run()
    e1.record(s)
    e1.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1000.0 / args.steps
    dev_ms = e0.elapsed_ms(e1) / args.steps
    mism = sum(1 for q in range(args.queries) if b.num_groups(q) != plain_groups[q])
    tune, arena = b.tune()
    algo = total_rows * 16 + sum(plain_groups) * (4 + 8 + 4 + 8)
    line = _line(args.queries / (wall_ms * 1e-3), "queries/s", args, wall_ms, "int64",
                 {"workload": f"batched GROUP BY: {args.queries} queries x (GROUP BY int32 Zipf(1.1) key, SUM(int64), "
                              f"SUM(float32), COUNT(*)), rows log-uniform 1e4..4e6, cardinality log-uniform 16..1e6",
                  "total_rows": total_rows, "groups": int(sum(plain_groups)),
                  "plain_ms_per_batch": plain_ms, "plain_queries_per_s": args.queries / (plain_ms * 1e-3),
                  "speedup_vs_plain": plain_ms / wall_ms, "group_count_mismatches": mism,
                  "create_ms": create_ms, "tune": tune, "arena_bytes": arena},
                 _roof(algo, dev_ms, "whole batched pipeline (sketch, plan, aggregate, extract)"))
    if keep:
        line["_batch"], line["_queries"] = b, queries
    else:
        b.release()
    return line


def serve_specs(n: int, seed: int) -> list[dict]:
    """The served query mix (filter -> 1..3 joins -> GROUP BY 10K..100K groups ->
    ORDER BY + LIMIT 100) with seeded parameters; fact columns are
    fk1, fk2, fk3, price, qty, date."""
    from craton_bolt_amd import _native as N
    from craton_bolt_amd import datagen as D
    from craton_bolt_amd.ops import Term

    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        t = i % 3
        d0 = D.STAR_DATE0 + int(rng.integers(0, 2557 - 120))
        if t == 0:  # revenue by city, one quarter
            spec = dict(terms=[Term(5, N.GE, d0), Term(5, N.LT, d0 + 90)], joins=[(1, 1)], group=[(0, 0)],
                        aggs=[(N.SUM, 3), (N.SUM, 4), (N.COUNT_STAR, 0)], order_agg=0, limit=100, est_groups=10_000)
        elif t == 1:  # brand ranking, two months of expensive items
            p0 = int(rng.integers(100_000, 600_000))
            spec = dict(terms=[Term(5, N.GE, d0), Term(5, N.LT, d0 + 60), Term(3, N.GE, p0)], joins=[(0, 0)],
                        group=[(0, 0)], aggs=[(N.SUM, 3), (N.COUNT_STAR, 0)], order_agg=0, limit=100,
                        est_groups=100_000)
        else:  # category x zone over three joins, four months
            spec = dict(terms=[Term(5, N.GE, d0), Term(5, N.LT, d0 + 120)], joins=[(0, 0), (1, 1), (2, 2)],
                        group=[(2, 0), (1, 1)], aggs=[(N.COUNT_STAR, 0), (N.SUM, 4), (N.SUM, 3)], order_agg=2,
                        limit=100, est_groups=25_000)
        out.append(spec)
    return out


def serve_mix(srv, n: int, seed: int):
    return [srv.query(**spec) for spec in serve_specs(n, seed)]


def _agree(a, b, tol=1e-9):
    """Same groups and rows; floats within `tol` relative (atomic order)."""
    if a["status"] != b["status"] or a["n_groups"] != b["n_groups"] or len(a["rows"]) != len(b["rows"]):
        return False
    for ra, rb in zip(a["rows"], b["rows"]):
        for x, y in zip(ra, rb):
            if isinstance(y, float):
                if abs(x - y) > tol * max(abs(y), 1.0):
                    return False
            elif x != y:
                return False
    return True


def bench_serve(args, sweep_fracs=(0.25, 0.5, 0.75, 0.9), check: bool = False) -> dict:
    """Served form over a device-resident star schema: plain capacity (one
    query at a time, one launch per query) versus continuous batching with
    `slots` overlapping queries, plus a Poisson-arrival sweep with p50 / p99."""
    import threading

    from craton_bolt_amd import _native as N
    from craton_bolt_amd import datagen as D
    from craton_bolt_amd.device import DeviceColumn
    from craton_bolt_amd.ops import Server

    t0 = time.perf_counter()
    n = args.fact_rows
    f = D.star_fact(n)
    dts = [N.I32, N.I32, N.I32, N.DEC128, N.F64, N.DATE32]
    fact = [DeviceColumn.from_numpy(f[k], dt) for k, dt in zip(["fk1", "fk2", "fk3", "price", "qty", "date"], dts)]
    host = [f[k] for k in ["fk1", "fk2", "fk3", "price", "qty", "date"]] if check else None
    del f
    dims, dims_h = [], []
    for d, nd in enumerate(D.STAR_DIMS):
        k, attrs = D.star_dim(d, nd)
        dims_h.append((k, attrs))
        dims.append((DeviceColumn.from_numpy(k), [DeviceColumn.from_numpy(a) for a in attrs]))
    print(f"[bench] star schema {n} fact rows ready in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    nq = args.serve_queries

    def closed_loop(srv, queries):
        t = time.perf_counter()
        tickets = [srv.submit(q) for q in queries]
        res = [srv.wait(tk) for tk in tickets]
        return time.perf_counter() - t, res

    # plain: one query at a time through the operators (the reference's
    # execution of a star query: compaction, join + take per dimension,
    # GROUP BY, ORDER BY, LIMIT)
    from craton_bolt_amd.ops import execute_star_query

    for spec in serve_specs(3, seed=1):  # warm-up (module loads, pools)
        execute_star_query(fact, dims, **spec)
    n_plain = max(6, nq // 8)
    pspecs = serve_specs(n_plain, seed=D.SEED_SERVE)
    t = time.perf_counter()
    c_res = [execute_star_query(fact, dims, **spec) for spec in pspecs]
    c_t = time.perf_counter() - t
    plain_qps = n_plain / c_t
    print(f"[bench] plain (operator composition): {plain_qps:.2f} queries/s", file=sys.stderr, flush=True)

    # the fused single-query form: the server with one slot and one chunk
    fused = Server(fact, dims, slots=1, chunk_rows=n)
    closed_loop(fused, serve_mix(fused, 3, seed=1))
    fq = serve_mix(fused, max(6, nq // 4), seed=D.SEED_SERVE)
    f_t, f_res = closed_loop(fused, fq)
    fused_qps = len(fq) / f_t
    print(f"[bench] fused single-slot: {fused_qps:.2f} queries/s", file=sys.stderr, flush=True)
    fused.close()

    srv = Server(fact, dims, slots=args.slots)
    closed_loop(srv, serve_mix(srv, args.slots, seed=2))
    sq = serve_mix(srv, nq, seed=D.SEED_SERVE)
    steps0 = srv.stats()["steps"]
    s_t, s_res = closed_loop(srv, sq)
    laps = (srv.stats()["steps"] - steps0) / max(1, -(-n // (8 << 20)))
    served_qps = nq / s_t
    bad = sum(1 for r in s_res if r["status"] != 0)
    print(f"[bench] served: {served_qps:.2f} queries/s ({bad} errors)", file=sys.stderr, flush=True)
    # identical results: the same seeded queries through all three paths
    same = all(_agree(a, b) for a, b in zip(c_res, s_res[: len(c_res)])) and \
        all(_agree(a, b) for a, b in zip(f_res, s_res[: len(f_res)]))

    sweep = []
    rng = np.random.default_rng(D.SEED_SERVE + 7)
    for frac in sweep_fracs:
        qps = served_qps * frac
        queries = serve_mix(srv, nq, seed=D.SEED_SERVE + int(frac * 100))
        gaps = rng.exponential(1.0 / qps, size=nq)
        tickets = [None] * nq
        start = time.perf_counter()

        def feeder():
            at = start
            for i, q in enumerate(queries):
                at += gaps[i]
                d = at - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                tickets[i] = srv.submit(q)

        th = threading.Thread(target=feeder)
        th.start()
        th.join()
        res = [srv.wait(tk) for tk in tickets]
        lat = np.array([r["done_ms"] - r["submit_ms"] for r in res])
        span = (max(r["done_ms"] for r in res) - min(r["submit_ms"] for r in res)) / 1000.0
        sweep.append({"offered_qps": qps, "achieved_qps": nq / span, "p50_ms": float(np.percentile(lat, 50)),
                      "p99_ms": float(np.percentile(lat, 99))})
        print(f"[bench] offered {qps:.1f} qps: p50 {sweep[-1]['p50_ms']:.1f} ms p99 {sweep[-1]['p99_ms']:.1f} ms",
              file=sys.stderr, flush=True)
    st = srv.stats()
    srv.close()
    lat_closed = np.array([r["done_ms"] - r["submit_ms"] for r in s_res])
    fact_bytes = n * (3 * 4 + 16 + 8 + 4)
    line = _line(served_qps, "queries/s", args, s_t * 1000.0 / nq, "decimal128/int64/f64",
                 {"workload": f"served star queries (filter -> 1..3 joins -> GROUP BY 10K..100K groups -> ORDER BY "
                              f"+ LIMIT 100) over {n} fact rows, {args.slots} slots, continuous batching",
                  "fact_rows": n, "slots": args.slots,
                  "plain": "one query at a time through the operators (count + compaction, hash build + pair "
                           "probe + take per join, GROUP BY, ORDER BY, LIMIT), as the reference runs a query",
                  "plain_queries_per_s": plain_qps, "speedup_vs_plain": served_qps / plain_qps,
                  "fused_single_slot_queries_per_s": fused_qps, "speedup_vs_fused_single_slot": served_qps / fused_qps,
                  "saturation_p50_ms": float(np.percentile(lat_closed, 50)),
                  "saturation_p99_ms": float(np.percentile(lat_closed, 99)), "poisson_sweep": sweep,
                  "errors": bad, "plain_fused_and_served_agree": bool(same), "server_stats": st},
                 _roof(fact_bytes, s_t * 1000.0 / max(laps, 1e-9),
                       "shared scan step kernel: fact bytes per lap / mean lap time at saturation"))
    if check:
        import oracle as O

        specs = serve_specs(nq, seed=D.SEED_SERVE)[:3]  # one of each query template, as served above
        bad, worst = [], 0.0
        for i, spec in enumerate(specs):
            g, want, sabs = star_reference(host, dims_h, spec)
            got = s_res[i]
            ok = got["status"] == 0 and got["n_groups"] == g and len(got["rows"]) == len(want)
            nk = len(spec["group"])
            for a, b in zip(got["rows"], want) if ok else []:
                ok = ok and a[:nk] == b[:nk]
                for x, y in zip(a[nk:], b[nk:]):
                    if isinstance(y, float):
                        worst = max(worst, abs(x - y) / max(abs(y), 1e-300))
                        ok = ok and abs(x - y) <= 1e-12 * sabs + 1e-9 * abs(y)
                    else:
                        ok = ok and x == y
            if not ok:
                bad.append(i)
        line["check"] = {"ok": not bad and bool(same), "queries_checked": len(specs), "mismatches": bad,
                         "plain_fused_and_served_agree": bool(same),
                         "oracle": "one query of each template at the full 200M-row fact table and 1M / 100K / 2.5K "
                                   "dimensions against the oracle's filter, host joins, GROUP BY, ORDER BY, LIMIT "
                                   "(keys, counts, Decimal128 sums exact; float sums within 1e-12 * sum(|x|))"}
        line["float_error"] = {"sum_qty_max_rel_err_vs_fp64_neumaier": worst}
        m = min(n, 10_000_000)
        hs = [a[:m] for a in host]
        t = _cpu_median(lambda: [star_reference(hs, dims_h, sp) for sp in specs])
        t_q = t / len(specs) * (n / m)
        line["cpu_baseline"] = _cpu_line(1.0 / t_q, "queries/s",
                                         f"the oracle's composition (C filter, host joins, C GROUP BY fold, sort) of "
                                         f"one query per template over the first {m} fact rows, scaled by rows to "
                                         f"{n}", t)
    return line


# ---- sub-records of the default run ---------------------------------------------------------------

def _sub(value, unit, ms, steps, dtype, config, roofline, check, float_error=None, cpu=None) -> dict:
    return {"value": value, "unit": unit, "ms_per_step": ms, "steps": steps, "dtype": dtype, "config": config,
            "roofline": roofline, "check": check, "float_error": float_error, "cpu_baseline": cpu}


def _op_roof(bytes_per_call, op_ms, what, kernel_bytes=None, kernel_ms=None, kernel=None) -> dict:
    """Operator-level roofline (every kernel of the call, on its stream), plus
    the dominant kernel's where the library times it separately."""
    r = _roof(bytes_per_call, op_ms, what)
    r["level"] = "operator"
    if kernel_ms:
        k = _roof(kernel_bytes, kernel_ms, kernel)
        r["dominant_kernel"] = {"kernel": kernel, "ms": kernel_ms, "achieved": k["achieved"], "frac": k["frac"],
                                "bytes_per_launch": kernel_bytes}
    return r


def _time_groupby(dk, dv, aggs, flags, steps, warmup):
    """(operator ms on the stream, dominant-kernel ms (LDS-tier pass), groups)."""
    import ctypes as C

    from craton_bolt_amd import _native as N
    from craton_bolt_amd.device import Event, Stream, columns_array
    from craton_bolt_amd.ops import _aggs

    s = Stream()
    ka, ia, ag = columns_array([dk]), columns_array([dv]), _aggs(aggs)
    L = N.lib()
    ev, kev = [Event(), Event()], [Event(), Event()]
    groups = {}

    def one(record):
        tbl = C.c_void_p()
        if record:
            ev[0].record(s)
        N.check(L.bolt_groupby(s.handle, ka, 1, ia, 1, ag, len(aggs), None, flags, C.byref(tbl)), "bolt_groupby")
        if record:
            ev[1].record(s)
        nr = C.c_int64()
        L.bolt_table_num_rows(tbl, C.byref(nr))
        groups["g"] = nr.value
        L.bolt_table_release(tbl)

    for _ in range(warmup):
        one(False)
    s.synchronize()
    op, km = [], []
    L.bolt_profile_events(kev[0].handle, kev[1].handle)
    for _ in range(steps):
        one(True)
        ev[1].synchronize()
        op.append(ev[0].elapsed_ms(ev[1]))
        km.append(kev[0].elapsed_ms(kev[1]))
    L.bolt_profile_events(None, None)
    return float(np.mean(op)), float(np.mean(km)), groups["g"]


def sub_groupby_small(args, n: int) -> dict:
    """Config 1: SELECT k, SUM(v), COUNT(*), MIN(v), MAX(v) GROUP BY k over n rows, output in key order."""
    import oracle as O
    from craton_bolt_amd import _native as N
    from craton_bolt_amd import datagen as D
    from craton_bolt_amd.device import DeviceColumn
    from craton_bolt_amd.ops import execute_groupby

    k, v = D.groupby_small(n)
    dk, dv = DeviceColumn.from_numpy(k), DeviceColumn.from_numpy(v)
    aggs = [(N.SUM, 0), (N.COUNT_STAR, 0), (N.MIN, 0), (N.MAX, 0)]
    op_ms, k_ms, g = _time_groupby(dk, dv, aggs, N.GROUPBY_SORTED, args.steps, args.warmup)
    check = None
    if not args.no_check:
        res = execute_groupby([dk], [dv], aggs)
        want = O.groupby([O.HostCol(k, O.I32)], [O.HostCol(v, O.I64)], aggs, max_groups=4096)
        keys = res["keys"][0][0].astype(np.uint32).astype(np.int64)
        ok = want["status"] == O.OK and res["n_groups"] == want["n_groups"] and np.array_equal(keys, want["keys"])
        for a in range(4):
            ok = ok and np.array_equal(res["aggs"][a][0].astype(np.int64), want["i64"][a])
        check = {"ok": bool(ok), "groups": int(res["n_groups"]),
                 "oracle": "host GROUP BY fold (groupby.rs:295 classic semantics), full size, every group and aggregate"}
    sample = min(n, 1 << 24)
    ks, vs = [O.HostCol(k[:sample], O.I32)], [O.HostCol(v[:sample], O.I64)]
    t = _cpu_median(lambda: O.groupby(ks, vs, aggs, max_groups=4096))
    cpu = _cpu_line(sample / t, "rows/s", f"C restatement of the host GROUP BY fold (groupby_wide.rs:85 HashMap per "
                                          f"group, checked i64 SUM), first {sample} rows", t)
    byts = n * 12 + g * (4 + 8 + 8 + 8 + 8)
    return _sub(n / (op_ms * 1e-3), "rows/s", op_ms, args.steps, "int64",
                {"workload": f"plain GROUP BY: {n} rows, int32 key over 1000 values, int64 value; SUM, COUNT(*), "
                             f"MIN, MAX; output in key order", "n": n, "groups": g},
                _op_roof(byts, op_ms, "whole GROUP BY call (LDS tier, merge, extraction, key-order sort)",
                         n * 12, k_ms, "generated LDS-tier pass (keys + values read once)"), check, None, cpu)


def sub_sort(args) -> dict:
    """Config 3, sort: 256M (int64 key, u32 row id) pairs."""
    import oracle as O
    from craton_bolt_amd import _native as N
    from craton_bolt_amd import datagen as D
    from craton_bolt_amd.device import DeviceBuffer, Event, Stream
    from craton_bolt_amd.ops import sort_pairs_i64

    n = args.n
    key, rid = D.sort_pairs(n)
    dk, dv = DeviceBuffer.from_numpy(key), DeviceBuffer.from_numpy(rid)
    s = Stream()
    kev = [Event(), Event()]
    kms = []

    def one():
        sort_pairs_i64(dk, dv, n, stream=s)
        kev[1].synchronize()
        kms.append(kev[0].elapsed_ms(kev[1]))

    N.lib().bolt_profile_events(kev[0].handle, kev[1].handle)
    _timed(s, one, args.steps, args.warmup)
    N.lib().bolt_profile_events(None, None)
    ms = float(np.mean(kms[args.warmup:]))
    check = None
    if not args.no_check:
        ko, vo = sort_pairs_i64(dk, dv, n, stream=s)
        s.synchronize()
        k2, r2 = ko.to_numpy(np.int64, n), vo.to_numpy(np.uint32, n)
        del ko, vo
        inc = bool(np.all((k2[1:] > k2[:-1]) | ((k2[1:] == k2[:-1]) & (r2[1:] > r2[:-1]))))
        inr = bool(r2.max() < n) if n else True
        same = bool(inr and np.array_equal(key[r2], k2))
        check = {"ok": inc and same, "pairs_strictly_increasing": inc, "key_in_of_rid_equals_key_out": same,
                 "oracle": "property at full size: (key, row id) strictly increasing (sorted and stable, as "
                           "tests/sort_e2e.rs:57-93 checks order and permutation) and key_in[row id] == key_out"}
        del k2, r2
    sample = min(n, 1 << 22)
    hc = [O.HostCol(key[:sample], O.I64)]
    t = _cpu_median(lambda: O.sort_indices(hc, [False], [True]))
    t_full = t * (n * np.log2(max(n, 2))) / (sample * np.log2(max(sample, 2)))
    cpu = _cpu_line(n / t_full, "keys/s", f"C restatement of the host lexsort + take (sort.rs:222, a stable comparison "
                                          f"sort of row ids): {sample} keys in {t:.3f} s, extrapolated by n log n "
                                          f"to {n} keys", t)
    return _sub(n / (ms * 1e-3), "keys/s", ms, args.steps, "int64",
                {"workload": f"ORDER BY radix sort: {n} (int64 key uniform over the full range, u32 row id)",
                 "n": n, "compulsory_bytes": n * 24, "model_bytes": n * 200},
                _op_roof(n * 200, ms, "hybrid radix sort (histogram, 2 onesweep passes, LDS bucket sorts), "
                                      "200 B/key pass model"), check, None, cpu)


def sub_zipf(args) -> dict:
    """Config 3, GROUP BY: 256M rows, int64 key Zipf(1.1) over 10M ranks, SUM(float64), COUNT(*)."""
    import oracle as O
    from craton_bolt_amd import _native as N
    from craton_bolt_amd import datagen as D
    from craton_bolt_amd.device import DeviceColumn
    from craton_bolt_amd.ops import autotune_groupby_partitions, execute_groupby, groupby_last_plan

    n = args.n
    k, v = D.zipf_groupby(n, n_ranks=10_000_000)
    dk, dv = DeviceColumn.from_numpy(k), DeviceColumn.from_numpy(v)
    aggs = [(N.SUM, 0), (N.COUNT_STAR, 0)]
    # the partitioned tier's fan-out, tuned on this input before the timed calls (which then use it)
    tuned_parts, tuned_ms = autotune_groupby_partitions([dk], [dv], aggs, reps=1)
    op_ms, k_ms, g = _time_groupby(dk, dv, aggs, 0, args.steps, args.warmup)
    plan = groupby_last_plan()
    check, ferr = None, None
    if not args.no_check:
        res = execute_groupby([dk], [dv], aggs, sorted_output=False)
        gk, gs, gc = res["keys"][0][0], res["aggs"][0][0], res["aggs"][1][0]
        # every group whose key has its low 6 bits clear (keys are scrambled ranks, so about 1/64 of the
        # groups and rows), against the oracle run on exactly those rows
        rows = (k & 63) == 0
        ks, vs = k[rows], v[rows]
        want = O.groupby([O.HostCol(ks, O.I64)], [O.HostCol(vs, O.F64)], aggs)
        wabs = O.groupby([O.HostCol(ks, O.I64)], [O.HostCol(np.abs(vs), O.F64)], [(N.SUM, 0)])
        sel = (gk & 63) == 0
        order = np.argsort(gk[sel], kind="stable")
        sk, ss, sc = gk[sel][order], gs[sel][order], gc[sel][order]
        same_keys = want["status"] == O.OK and np.array_equal(sk, want["keys"])
        cnt_ok = same_keys and np.array_equal(sc, want["i64"][1])
        err = np.abs(ss - want["f64"][0]) if same_keys else np.array([np.inf])
        scale = wabs["f64"][0] if same_keys else np.array([1.0])
        within = bool(np.all(err <= 1e-12 * scale + 1e-300))
        big = np.abs(want["f64"][0]) > 1e-6 * scale if same_keys else np.array([False])
        total_gpu = float(np.sum(gs.astype(np.float64)))
        total_want = O.neumaier_sum(v)
        check = {"ok": bool(same_keys and cnt_ok and within and int(gc.sum()) == n), "groups": int(g),
                 "checked_groups": int(sel.sum()), "checked_rows": int(rows.sum()),
                 "count_sum_equals_rows": int(gc.sum()) == n, "keys_and_counts_exact": bool(cnt_ok),
                 "oracle": "every group with key & 63 == 0 (about 1/64 of groups and rows) against the host GROUP "
                           "BY fold of exactly those rows, full-size run; all counts sum to n"}
        ferr = {"sum_max_abs_err_over_sum_abs": float(np.max(err / scale)) if len(err) else 0.0,
                "sum_max_rel_err": float(np.max(err[big] / np.abs(want["f64"][0][big]))) if np.any(big) else 0.0,
                "tolerance": "1e-12 * sum(|x|) per group",
                "total_rel_err_vs_fp64_neumaier": abs(total_gpu - total_want) / max(abs(total_want), 1e-300)}
    sample = min(n, 1 << 23)
    kc, vc = [O.HostCol(k[:sample], O.I64)], [O.HostCol(v[:sample], O.F64)]
    t = _cpu_median(lambda: O.groupby(kc, vc, aggs, max_groups=sample))
    cpu = _cpu_line(sample / t, "rows/s", f"C restatement of the host GROUP BY fold (HashMap per group, sequential "
                                          f"f64 SUM): first {sample} rows (a smaller group table than the full run)",
                    t)
    byts = n * 16 + g * (8 + 8 + 8)
    return _sub(n / (op_ms * 1e-3), "rows/s", op_ms, args.steps, "f64",
                {"workload": f"high-cardinality GROUP BY: {n} rows, int64 key Zipf(1.1) over 10M ranks, "
                             f"SUM(float64), COUNT(*)", "n": n, "groups": g, "tier": plan["tier"],
                 "partitions": plan["partitions"], "spilled_rows": plan["spilled_rows"],
                 "partial_records": plan["partial_records"],
                 "autotuned_partitions": {"chosen": tuned_parts, "ms_per_call_while_tuning": tuned_ms}},
                _op_roof(byts, op_ms, "whole GROUP BY call (LDS tier, partitioned tier, extraction)",
                         n * 16, k_ms, "generated LDS-tier pass"), check, ferr, cpu)


def sub_batch(args) -> dict:
    """Config 4: the batched form, every query's values against the plain path, a sample against the oracle."""
    import oracle as O
    from craton_bolt_amd import _native as N
    from craton_bolt_amd import datagen as D
    from craton_bolt_amd.ops import execute_groupby

    line = bench_batch(args, keep=True)
    b, queries = line.pop("_batch"), line.pop("_queries")
    aggs = [(N.SUM, 0), (N.SUM, 1), (N.COUNT_STAR, 0)]
    check, ferr = None, None
    if not args.no_check:
        bad, worst = [], 0.0
        for q, (kc, ic) in enumerate(queries):
            rb = b.result(q)
            rp = execute_groupby(kc, ic, aggs)
            # both in ascending int32 key order (the plain path's packed-key order puts negative keys last)
            o = np.argsort(rb["keys"][0][0], kind="stable")
            op = np.argsort(rp["keys"][0][0], kind="stable")
            ok = rb["n_groups"] == rp["n_groups"] and np.array_equal(rb["keys"][0][0][o], rp["keys"][0][0][op])
            ok = ok and np.array_equal(rb["aggs"][0][0][o], rp["aggs"][0][0][op]) and \
                np.array_equal(rb["aggs"][2][0][o], rp["aggs"][2][0][op])
            if ok:
                fb, fp = rb["aggs"][1][0][o].astype(np.float64), rp["aggs"][1][0][op].astype(np.float64)
                d = np.abs(fb - fp) / np.maximum(np.abs(fp), 1e-30)
                worst = max(worst, float(d.max()) if len(d) else 0.0)
            if not ok:
                bad.append(q)
        # a sample of queries against the oracle (keys, integer sums, counts exact; the f32 sums' error)
        shapes = D.batch_shapes(args.queries)
        sample_q = list(range(0, args.queries, max(1, args.queries // 8)))[:8]
        o_bad, o_err, rows_s, t_s = [], 0.0, 0, 0.0
        for q in sample_q:
            nq, card = shapes[q]
            k, vi, vf = D.zipf_small(nq, card, seed=D.SEED_BATCH + q)
            t0 = time.perf_counter()
            want = O.groupby([O.HostCol(k, O.I32)], [O.HostCol(vi, O.I64), O.HostCol(vf, O.F32)], aggs,
                             max_groups=card + 1)
            t_s += time.perf_counter() - t0
            rows_s += nq
            rb = b.result(q)
            o = np.argsort(rb["keys"][0][0], kind="stable")
            wk = want["keys"].astype(np.uint32).view(np.int32)
            wo = np.argsort(wk, kind="stable")
            ok = np.array_equal(rb["keys"][0][0][o], wk[wo]) and np.array_equal(rb["aggs"][0][0][o], want["i64"][0][wo]) \
                and np.array_equal(rb["aggs"][2][0][o], want["i64"][2][wo])
            if ok and len(wo):
                wf = want["f64"][1][wo]
                o_err = max(o_err, float(np.max(np.abs(rb["aggs"][1][0][o].astype(np.float64) - wf) /
                                                np.maximum(np.abs(wf), 1e-30))))
            if not ok:
                o_bad.append(q)
        check = {"ok": not bad and not o_bad, "queries_equal_to_plain": args.queries - len(bad),
                 "mismatched_queries": bad[:10], "oracle_sample_queries": sample_q, "oracle_mismatches": o_bad,
                 "oracle": "all queries: groups, keys, SUM(int64), COUNT exact against the plain per-query "
                           "operator; the sampled queries against the host GROUP BY fold"}
        ferr = {"sum_f32_max_rel_diff_batch_vs_plain": worst,
                "sum_f32_max_rel_err_vs_fp64_neumaier_sample": o_err,
                "note": "SUM(Float32) is accumulated in f64 and rounded to f32 once"}
        # the CPU baseline: the sampled queries' host GROUP BYs, scaled by rows to the whole batch
        total_rows = line["config"]["total_rows"]
        t_all = t_s * total_rows / max(rows_s, 1)
        line["cpu_baseline"] = _cpu_line(args.queries / t_all, "queries/s",
                                         f"C restatement of the host GROUP BY fold, one query after another: "
                                         f"{len(sample_q)} sampled queries ({rows_s} rows) in {t_s:.2f} s (one run), "
                                         f"scaled by rows to the {args.queries} queries", t_s)
    b.release()
    line["check"], line["float_error"] = check, ferr
    return _sub(line["value"], line["unit"], line["ms_per_step"], args.steps, line["dtype"], line["config"],
                line["roofline"], check, ferr, line.get("cpu_baseline"))


def star_reference(host, dims_h, spec):
    """The oracle's composition of a star query: filter, host joins against the dimension keys, GROUP BY,
    then ORDER BY the chosen aggregate DESC, key ASC, LIMIT."""
    import oracle as O
    from craton_bolt_amd import _native as N

    dts = [N.I32, N.I32, N.I32, N.DEC128, N.F64, N.DATE32]
    odt = [{N.I32: O.I32, N.DEC128: O.DEC128, N.F64: O.F64, N.DATE32: O.DATE32}[dt] for dt in dts]
    n = len(host[0])
    hc = [O.HostCol(a, o) for a, o in zip(host, odt)]
    rows = O.filter_rows(hc, [(t.col, t.op, t.value) for t in spec["terms"]], n) if spec["terms"] else np.arange(n)
    drow = []
    for fk, d in spec["joins"]:
        key = dims_h[d][0]
        order = np.argsort(key)
        pos = np.minimum(np.searchsorted(key[order], host[fk][rows]), len(key) - 1)
        hit = key[order][pos] == host[fk][rows]
        rows = rows[hit]
        drow = [r[hit] for r in drow] + [order[pos[hit]]]
    gk = [dims_h[spec["joins"][j][1]][1][a][drow[j]] for j, a in spec["group"]]
    inputs = [O.HostCol(np.ascontiguousarray(host[c][rows]), odt[c]) for c in range(len(host))]
    want = O.groupby([O.HostCol(np.ascontiguousarray(x), O.I32) for x in gk], inputs, spec["aggs"],
                     max_groups=len(rows) + 1)
    keys = want["keys"]
    kcols = [keys.astype(np.uint32).view(np.int32).astype(np.int64)] if len(gk) == 1 else \
        [(keys >> 32).astype(np.int64), (keys & 0xFFFFFFFF).astype(np.int64)]
    vals = []
    for a, (op, col) in enumerate(spec["aggs"]):
        if op in (N.COUNT_STAR, N.COUNT):
            vals.append([int(x) for x in want["i64"][a]])
        elif dts[col] == N.DEC128:
            vals.append([(int(h) << 64) | (int(l) & 0xFFFFFFFFFFFFFFFF) for l, h in zip(want["i64"][a], want["hi"][a])])
        else:
            vals.append([float(x) for x in want["f64"][a]])
    out = [tuple(int(k[i]) for k in kcols) + tuple(v[i] for v in vals) for i in range(want["n_groups"])]
    oa = len(kcols) + spec["order_agg"]
    out.sort(key=lambda r: tuple(r[:len(kcols)]))
    out.sort(key=lambda r: r[oa], reverse=True)
    return want["n_groups"], out[: spec["limit"]], float(np.abs(host[4][rows]).sum()) if len(rows) else 0.0


def sub_serve(args) -> dict:
    """Config 5, trimmed to fit the default run: 32 queries at saturation, a two-point Poisson sweep."""
    line = bench_serve(args, sweep_fracs=(0.5, 0.9), check=not args.no_check)
    return _sub(line["value"], line["unit"], line["ms_per_step"], args.serve_queries, line["dtype"], line["config"],
                line["roofline"], line.get("check"), line.get("float_error"), line.get("cpu_baseline"))


def bench_all(args) -> dict:
    """The headline line (config 2) with every other configuration as a sub-record."""
    line = bench_join(args)
    subs = {}
    steps, warmup = args.steps, args.warmup
    plan = [("groupby_1M", lambda: sub_groupby_small(args, 1 << 20)),
            ("groupby_256M", lambda: sub_groupby_small(args, args.n)),
            ("sort_256M", lambda: sub_sort(args)),
            ("groupby_zipf_256M", lambda: sub_zipf(args)),
            ("batch_512", lambda: sub_batch(args)),
            ("serve_200M", lambda: sub_serve(args))]
    for name, fn in plan:
        t0 = time.perf_counter()
        print(f"[bench] sub-record {name} ...", file=sys.stderr, flush=True)
        args.steps, args.warmup = min(steps, 10), min(warmup, 2)
        try:
            subs[name] = fn()
        except Exception as e:  # a failing configuration is reported, not hidden
            subs[name] = {"error": f"{type(e).__name__}: {e}"}
        args.steps, args.warmup = steps, warmup
        subs[name]["bench_s"] = time.perf_counter() - t0
        print(f"[bench] sub-record {name} done in {subs[name]['bench_s']:.1f}s", file=sys.stderr, flush=True)
    line["sub_records"] = subs
    return line


def main():
    args = _parse()
    if args.workload == "all" and args.serve_queries == 64:
        args.serve_queries = 32
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each box has one MI355X); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "all":
        line = bench_all(args)
    elif args.workload == "join":
        line = bench_join(args)
    elif args.workload == "sort":
        line = bench_sort(args)
    elif args.workload == "topk":
        line = bench_topk(args)
    elif args.workload == "expr":
        line = bench_expr(args)
    elif args.workload == "h2o":
        line = bench_h2o(args)
    elif args.workload == "groupby":
        line = bench_groupby(args, zipf=False)
    elif args.workload == "groupby-zipf":
        line = bench_groupby(args, zipf=True)
    elif args.workload == "batch":
        line = bench_batch(args)
    elif args.workload == "serve":
        line = bench_serve(args)
    else:
        line = bench_calib(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
