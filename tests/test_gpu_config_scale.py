"""GPU parity at the benchmark configurations' full sizes.

The other test files compare against the oracle at sizes it finishes in
seconds; these reach the code paths that only the full-size inputs take:
  * the 256M-pair radix sort (buckets near the LDS capacity), checked by a
    size-independent property: the output (key, row id) pairs are strictly
    increasing (sorted and stable, the check of tests/sort_e2e.rs:57-93) and
    key_in[row id] == key_out (a permutation of the input);
  * the high-cardinality GROUP BY at 128M Zipf(1.1) rows over 10M ranks, where
    the default tier choice takes the radix-partitioned tier with thousands of
    partitions, against the oracle's host fold of every group;
  * all 512 batched queries of the batched configuration against 512 plain
    per-query GROUP BYs (every group and value), and a sample against the
    oracle;
  * served star queries at the configuration's dimension sizes (1M / 100K /
    2.5K rows) over 20M fact rows against the oracle's composition.
Integer results are bit-exact; float sums within 1e-12 * sum(|x|) of the
oracle's compensated sum (the atomics' order is not fixed)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from craton_bolt_amd import _native as N
from craton_bolt_amd import datagen as D
from craton_bolt_amd.device import DeviceBuffer, DeviceColumn
from craton_bolt_amd.ops import GroupByBatch, Server, execute_groupby, sort_pairs_i64

pytestmark = pytest.mark.gpu


def test_sort_256m_pairs_sorted_stable_and_a_permutation(gpu):
    n = 256 << 20
    print(f"[case] sort_pairs n={n}", flush=True)
    key, rid = D.sort_pairs(n)
    ko, vo = sort_pairs_i64(DeviceBuffer.from_numpy(key), DeviceBuffer.from_numpy(rid), n)
    k2, r2 = ko.to_numpy(np.int64, n), vo.to_numpy(np.uint32, n)
    del ko, vo
    assert int(r2.max()) < n
    assert np.all((k2[1:] > k2[:-1]) | ((k2[1:] == k2[:-1]) & (r2[1:] > r2[:-1])))
    assert np.array_equal(key[r2], k2)


def test_zipf_groupby_128m_partitioned_tier_matches_oracle(gpu):
    n = 128 << 20
    print(f"[case] zipf groupby n={n}", flush=True)
    k, v = D.zipf_groupby(n, n_ranks=10_000_000)
    aggs = [(N.SUM, 0), (N.COUNT_STAR, 0)]
    res = execute_groupby([DeviceColumn.from_numpy(k)], [DeviceColumn.from_numpy(v)], aggs)
    tier, spilled, partials, parts = C.c_int32(), C.c_int64(), C.c_int64(), C.c_int32()
    N.lib().bolt_groupby_last_plan(C.byref(tier), C.byref(spilled), C.byref(partials), C.byref(parts))
    print(f"[zipf] tier={tier.value} spilled={spilled.value} partials={partials.value} partitions={parts.value}",
          flush=True)
    assert tier.value == 2 and parts.value >= 1024  # the default choice at this size
    # one fold for the sums, the counts and the tolerance scale sum(|x|) per group
    want = O.groupby([O.HostCol(k, O.I64)], [O.HostCol(v, O.F64), O.HostCol(np.abs(v), O.F64)],
                     aggs + [(N.SUM, 1)], max_groups=10_000_000)
    assert want["status"] == O.OK
    assert res["n_groups"] == want["n_groups"]
    assert np.array_equal(res["keys"][0][0], want["keys"])  # output in key order
    assert np.array_equal(res["aggs"][1][0], want["i64"][1])
    sabs = want["f64"][2]
    err = np.abs(res["aggs"][0][0] - want["f64"][0])
    assert np.all(err <= 1e-12 * sabs + 1e-300), f"max err/sum|x| {np.max(err / sabs):.3e}"
    print(f"[zipf] groups={want['n_groups']} max err/sum|x|={np.max(err / sabs):.3e}", flush=True)


def _by_key(res):
    o = np.argsort(res["keys"][0][0], kind="stable")
    return res["keys"][0][0][o], [a[0][o] for a in res["aggs"]]


def test_batch_512_queries_equal_plain_and_oracle(gpu):
    shapes = D.batch_shapes(512)
    aggs = [(N.SUM, 0), (N.SUM, 1), (N.COUNT_STAR, 0)]
    queries = []
    for q, (nq, card) in enumerate(shapes):
        k, vi, vf = D.zipf_small(nq, card, seed=D.SEED_BATCH + q)
        queries.append(([DeviceColumn.from_numpy(k)], [DeviceColumn.from_numpy(vi), DeviceColumn.from_numpy(vf)]))
    print(f"[case] batch 512 queries, {sum(n for n, _ in shapes)} rows", flush=True)
    b = GroupByBatch(queries, aggs)
    b.run()
    for q, (kc, ic) in enumerate(queries):
        bk, ba = _by_key(b.result(q))
        pk, pa = _by_key(execute_groupby(kc, ic, aggs))
        assert np.array_equal(bk, pk), f"query {q} keys"
        assert np.array_equal(ba[0], pa[0]) and np.array_equal(ba[2], pa[2]), f"query {q} SUM(int64) / COUNT"
        # both fold in f64 with atomics (order not fixed) and round to f32 once
        assert np.all(np.abs(ba[1] - pa[1]) <= 2.0 ** -23 * np.abs(pa[1]) + 1e-30), f"query {q} SUM(float32)"
    for q in range(0, 512, 64):
        nq, card = shapes[q]
        k, vi, vf = D.zipf_small(nq, card, seed=D.SEED_BATCH + q)
        want = O.groupby([O.HostCol(k, O.I32)], [O.HostCol(vi, O.I64), O.HostCol(vf, O.F32)], aggs,
                         max_groups=card + 1)
        bk, ba = _by_key(b.result(q))
        wk = want["keys"].astype(np.uint32).view(np.int32)
        wo = np.argsort(wk, kind="stable")
        assert np.array_equal(bk, wk[wo])
        assert np.array_equal(ba[0], want["i64"][0][wo]) and np.array_equal(ba[2], want["i64"][2][wo])
        wf = want["f64"][1][wo]
        uk, inv = np.unique(k, return_inverse=True)
        sabs = np.bincount(inv, weights=np.abs(vf.astype(np.float64)), minlength=len(uk))
        assert np.all(np.abs(ba[1].astype(np.float64) - wf) <= 1e-12 * sabs + 2.0 ** -24 * np.abs(wf) + 1e-30)
    b.release()


FK1, FK2, FK3, PRICE, QTY, DATE = range(6)
DTS = [N.I32, N.I32, N.I32, N.DEC128, N.F64, N.DATE32]


def _star_reference(host, dims_h, spec):
    odt = [{N.I32: O.I32, N.DEC128: O.DEC128, N.F64: O.F64, N.DATE32: O.DATE32}[dt] for dt in DTS]
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
    assert want["status"] == O.OK
    keys = want["keys"]
    kcols = [keys.astype(np.uint32).view(np.int32).astype(np.int64)] if len(gk) == 1 else \
        [(keys >> 32).astype(np.int64), (keys & 0xFFFFFFFF).astype(np.int64)]
    vals = []
    for a, (op, col) in enumerate(spec["aggs"]):
        if op in (N.COUNT_STAR, N.COUNT):
            vals.append([int(x) for x in want["i64"][a]])
        elif DTS[col] == N.DEC128:
            vals.append([(int(h) << 64) | (int(l) & 0xFFFFFFFFFFFFFFFF) for l, h in zip(want["i64"][a], want["hi"][a])])
        else:
            vals.append([float(x) for x in want["f64"][a]])
    out = [tuple(int(kk[i]) for kk in kcols) + tuple(v[i] for v in vals) for i in range(want["n_groups"])]
    oa = len(kcols) + spec["order_agg"]
    out.sort(key=lambda r: tuple(r[:len(kcols)]))
    out.sort(key=lambda r: r[oa], reverse=True)
    return want["n_groups"], out[: spec["limit"]], float(np.abs(host[QTY][rows]).sum()) if len(rows) else 0.0


def test_served_star_queries_at_config_dimensions(gpu):
    import bench

    n = 20_000_000
    print(f"[case] served star queries over {n} fact rows, dims {D.STAR_DIMS}", flush=True)
    f = D.star_fact(n)
    host = [f[c] for c in ("fk1", "fk2", "fk3", "price", "qty", "date")]
    dev = [DeviceColumn.from_numpy(a, dt) for a, dt in zip(host, DTS)]
    dims_h = [D.star_dim(d, nd) for d, nd in enumerate(D.STAR_DIMS)]
    dims_d = [(DeviceColumn.from_numpy(k), [DeviceColumn.from_numpy(a) for a in attrs]) for k, attrs in dims_h]
    specs = bench.serve_specs(6, seed=D.SEED_SERVE + 1)  # two of each template
    srv = Server(dev, dims_d, slots=8, chunk_rows=4 << 20)
    tickets = [srv.submit(srv.query(**sp)) for sp in specs]  # all in flight together
    res = [srv.wait(t) for t in tickets]
    st = srv.stats()
    srv.close()
    assert st["peak_active"] > 1
    for i, (sp, r) in enumerate(zip(specs, res)):
        assert r["status"] == N.BOLT_OK, r["error"]
        g, want, sabs = _star_reference(host, dims_h, sp)
        assert r["n_groups"] == g, f"query {i}"
        nk = len(sp["group"])
        assert len(r["rows"]) == len(want)
        for a, w in zip(r["rows"], want):
            assert a[:nk] == w[:nk], (i, a, w)
            for x, y in zip(a[nk:], w[nk:]):
                if isinstance(y, float):
                    assert abs(x - y) <= 1e-12 * sabs + 1e-9 * abs(y), (i, a, w)
                else:
                    assert x == y, (i, a, w)
