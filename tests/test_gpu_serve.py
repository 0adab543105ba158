"""GPU parity: the served form (bolt_serve.h).

Star queries (filter -> 1..3 joins -> GROUP BY -> ORDER BY aggregate DESC,
key ASC -> LIMIT) over a seeded star schema are checked against the oracle
(its filter, host join semantics and GROUP BY, composed on the host), and
every query's result in a mix of concurrently admitted queries must equal its
result when run alone.  Keys, counts and Decimal128 sums are exact; float
sums within 1e-12 * sum(|x|) (atomic order differs between runs)."""
import numpy as np
import pytest

import oracle as O
from craton_bolt_amd import _native as N
from craton_bolt_amd import datagen as D
from craton_bolt_amd.device import DeviceColumn
from craton_bolt_amd.ops import Server, Term, execute_star_query

pytestmark = pytest.mark.gpu

N_FACT = 400_000
DIMS = (5000, 1000, 100)
# fact columns
FK1, FK2, FK3, PRICE, QTY, DATE = range(6)


@pytest.fixture(scope="module")
def star(gpu):
    f = D.star_fact(N_FACT, DIMS)
    host = [f["fk1"], f["fk2"], f["fk3"], f["price"], f["qty"], f["date"]]
    dts = [N.I32, N.I32, N.I32, N.DEC128, N.F64, N.DATE32]
    dev = [DeviceColumn.from_numpy(a, dt) for a, dt in zip(host, dts)]
    dims_h = [D.star_dim(d, n) for d, n in enumerate(DIMS)]
    dims_d = [(DeviceColumn.from_numpy(k), [DeviceColumn.from_numpy(a) for a in attrs]) for k, attrs in dims_h]
    return {"host": host, "dts": dts, "dev": dev, "dims_h": dims_h, "dims_d": dims_d}


def templates(rng):
    d0 = D.STAR_DATE0 + int(rng.integers(0, 2000))
    p0 = int(rng.integers(100_000, 900_000))
    return [
        # revenue by city over a date window
        dict(terms=[Term(DATE, N.GE, d0), Term(DATE, N.LT, d0 + 200)], joins=[(FK2, 1)], group=[(0, 0)],
             aggs=[(N.SUM, PRICE), (N.SUM, QTY), (N.COUNT_STAR, 0)], order_agg=0, limit=50, est_groups=1000),
        # brand ranking among expensive items
        dict(terms=[Term(PRICE, N.GE, p0)], joins=[(FK1, 0)], group=[(0, 0)],
             aggs=[(N.SUM, PRICE), (N.COUNT_STAR, 0)], order_agg=1, limit=100, est_groups=5000),
        # category x zone over three joins
        dict(terms=[Term(DATE, N.GE, d0), Term(DATE, N.LT, d0 + 400)], joins=[(FK1, 0), (FK2, 1), (FK3, 2)],
             group=[(2, 0), (1, 1)], aggs=[(N.COUNT_STAR, 0), (N.SUM, QTY), (N.MAX, QTY)], order_agg=0, limit=100,
             est_groups=25_000),
        # quantity by sub-category, no filter
        dict(terms=[], joins=[(FK3, 2)], group=[(0, 1)], aggs=[(N.SUM, QTY), (N.COUNT_STAR, 0)], order_agg=1,
             limit=20, est_groups=250),
    ]


def reference(star, spec):
    """Oracle filter + host INNER join + oracle GROUP BY, then ORDER BY / LIMIT."""
    host, dts = star["host"], star["dts"]
    n = len(host[0])
    odt = [{N.I32: O.I32, N.DEC128: O.DEC128, N.F64: O.F64, N.DATE32: O.DATE32}[dt] for dt in dts]
    hc = [O.HostCol(a, o) for a, o in zip(host, odt)]
    rows = O.filter_rows(hc, [(t.col, t.op, t.value) for t in spec["terms"]], n) if spec["terms"] else np.arange(n)
    drow = []
    for fk, d in spec["joins"]:
        key = star["dims_h"][d][0]
        order = np.argsort(key)
        pos = np.searchsorted(key[order], host[fk][rows])
        pos = np.minimum(pos, len(key) - 1)
        hit = key[order][pos] == host[fk][rows]
        rows = rows[hit]
        drow = [r[hit] for r in drow] + [order[pos[hit]]]
    gk = []
    for j, a in spec["group"]:
        d = spec["joins"][j][1]
        gk.append(star["dims_h"][d][1][a][drow[j]])
    inputs = [O.HostCol(np.ascontiguousarray(host[c][rows]), odt[c]) for c in range(len(host))]
    want = O.groupby([O.HostCol(np.ascontiguousarray(k), O.I32) for k in gk], inputs, spec["aggs"])
    assert want["status"] == O.OK
    g = want["n_groups"]
    keys = want["keys"]
    if len(gk) == 1:
        kcols = [keys.astype(np.uint32).view(np.int32).astype(np.int64)]
    else:
        kcols = [(keys >> 32).astype(np.int64), (keys & 0xFFFFFFFF).astype(np.int64)]
    vals = []
    for a, (op, col) in enumerate(spec["aggs"]):
        if op in (N.COUNT_STAR, N.COUNT):
            vals.append([int(v) for v in want["i64"][a]])
        elif dts[col] == N.DEC128:
            vals.append([(int(h) << 64) | (int(l) & 0xFFFFFFFFFFFFFFFF) for l, h in zip(want["i64"][a], want["hi"][a])])
        else:
            vals.append([float(v) for v in want["f64"][a]])
    rows_out = [tuple(int(k[i]) for k in kcols) + tuple(v[i] for v in vals) for i in range(g)]
    oa = len(kcols) + spec["order_agg"]
    rows_out.sort(key=lambda r: tuple(r[:len(kcols)]))
    rows_out.sort(key=lambda r: r[oa], reverse=True)  # stable: ties stay in key order
    sabs = float(np.abs(host[QTY][rows]).sum()) if len(rows) else 0.0
    return g, rows_out[: spec["limit"]], sabs


def same_rows(got, want, n_keys, sabs):
    assert len(got) == len(want)
    for a, b in zip(got, want):
        assert a[:n_keys] == b[:n_keys], (a, b)
        for x, y in zip(a[n_keys:], b[n_keys:]):
            if isinstance(y, float):
                assert abs(x - y) <= 1e-12 * sabs + 1e-9 * abs(y), (a, b)
            else:
                assert x == y, (a, b)


def make_server(star, **kw):
    return Server(star["dev"], star["dims_d"], **kw)


def run_query(srv, spec):
    t = srv.submit(srv.query(**spec))
    r = srv.wait(t)
    return r


def test_serve_queries_match_oracle(star):
    srv = make_server(star, slots=1, chunk_rows=100_000)
    rng = np.random.default_rng(11)
    for i, spec in enumerate(templates(rng)):
        print(f"[case] serve template {i}", flush=True)
        r = run_query(srv, spec)
        assert r["status"] == N.BOLT_OK, r["error"]
        g, want, sabs = reference(star, spec)
        assert r["n_groups"] == g
        same_rows(r["rows"], want, len(spec["group"]), sabs)
    srv.close()


def test_serve_predicate_forms(star):
    """Float, Decimal and Date32 intervals (strict and inclusive bounds,
    equality, an empty range) and the terms that are not intervals (NE,
    IS NOT NULL), evaluated per term: all against the oracle, served
    together in one mix."""
    d0 = D.STAR_DATE0 + 700
    base = dict(joins=[(FK3, 2)], group=[(0, 1)], aggs=[(N.SUM, QTY), (N.COUNT_STAR, 0), (N.SUM, PRICE)],
                order_agg=1, limit=30, est_groups=250)
    cases = [
        [Term(QTY, N.GT, 25.5), Term(QTY, N.LE, 40.0)],
        [Term(QTY, N.GE, 10.0), Term(PRICE, N.LT, 500_000), Term(DATE, N.GE, d0), Term(DATE, N.LE, d0 + 300)],
        [Term(DATE, N.EQ, d0 + 5)],
        [Term(QTY, N.GT, 30.0), Term(DATE, N.NE, d0)],
        [Term(PRICE, N.IS_NOT_NULL), Term(PRICE, N.GT, 900_000)],
        [Term(QTY, N.GT, 40.0), Term(QTY, N.LT, 20.0)],
        # four preloaded columns (the Int32 image of fk1 too)
        [Term(FK1, N.LT, 2500), Term(DATE, N.GE, d0), Term(PRICE, N.GE, 200_000), Term(QTY, N.LE, 45.0)],
        # a fifth predicate column across the mix (the earlier queries preload four):
        # fk2 is read per term, term-by-term evaluation
        [Term(FK2, N.GE, 100), Term(DATE, N.LT, d0 + 900), Term(QTY, N.GT, 2.0)],
    ]
    srv = make_server(star, slots=8, chunk_rows=100_000)
    tickets = [srv.submit(srv.query(terms=t, **base)) for t in cases]
    got = [srv.wait(tk) for tk in tickets]
    srv.close()
    for i, (t, r) in enumerate(zip(cases, got)):
        print(f"[case] predicate form {i}", flush=True)
        assert r["status"] == N.BOLT_OK, r["error"]
        g, want, sabs = reference(star, dict(base, terms=t))
        assert r["n_groups"] == g
        same_rows(r["rows"], want, 1, sabs)


def test_serve_mix_equals_alone_and_overlaps(star):
    rng = np.random.default_rng(12)
    specs = []
    for _ in range(4):
        specs += templates(rng)
    mix = make_server(star, slots=8, chunk_rows=50_000)
    tickets = [mix.submit(mix.query(**s)) for s in specs]
    mixed = [mix.wait(t) for t in tickets]
    st = mix.stats()
    mix.close()
    assert st["completed"] == len(specs)
    assert st["peak_active"] > 1  # queries really shared the device
    alone = make_server(star, slots=1, chunk_rows=N_FACT)
    for s, r in zip(specs, mixed):
        assert r["status"] == N.BOLT_OK, r["error"]
        a = run_query(alone, s)
        assert a["n_groups"] == r["n_groups"]
        sabs = float(np.abs(star["host"][QTY]).sum())
        same_rows(r["rows"], a["rows"], len(s["group"]), sabs)
    alone.close()


def test_serve_admission_respects_the_hbm_budget(star):
    spec = templates(np.random.default_rng(13))[1]
    probe = make_server(star, slots=1, chunk_rows=100_000)
    need = run_query(probe, spec)["hbm_bytes"]
    probe.close()
    assert need > 0
    budget = int(2.5 * need)  # room for two of these queries, not three
    srv = make_server(star, slots=8, chunk_rows=50_000, hbm_budget=budget)
    tickets = [srv.submit(srv.query(**spec)) for _ in range(6)]
    res = [srv.wait(t) for t in tickets]
    st = srv.stats()
    srv.close()
    assert all(r["status"] == N.BOLT_OK for r in res)
    assert st["peak_bytes"] <= budget
    assert st["peak_active"] == 2
    # FIFO admission: later requests are admitted no earlier than earlier ones
    admits = [r["admit_ms"] for r in res]
    assert admits == sorted(admits)
    assert all(r["rows"] == res[0]["rows"] for r in res)


def test_serve_declines_a_query_larger_than_the_budget(star):
    srv = make_server(star, slots=2, hbm_budget=1 << 20)
    spec = templates(np.random.default_rng(1))[2]
    spec["est_groups"] = 10_000_000
    r = run_query(srv, spec)
    assert r["status"] == N.BOLT_DECLINE
    srv.close()


def test_serve_rejects_bad_queries_and_regrows_small_tables(star):
    srv = make_server(star, slots=2, chunk_rows=100_000)
    spec = templates(np.random.default_rng(2))[0]
    bad = dict(spec, group=[(0, 7)])
    r = run_query(srv, bad)
    assert r["status"] == N.BOLT_ERR_PLAN and "attribute" in r["error"]
    # far too small an estimate: the first lap overflows the table and the
    # query is re-run with a larger one, transparently
    small = dict(templates(np.random.default_rng(3))[1], est_groups=1)
    r = run_query(srv, small)
    assert r["status"] == N.BOLT_OK, r["error"]
    g, want, sabs = reference(star, small)
    assert r["n_groups"] == g
    same_rows(r["rows"], want, 1, sabs)
    srv.close()


def test_composed_operator_path_matches_oracle_and_served(star):
    """The one-query-at-a-time operator composition (count, compaction, hash
    build + pair probe + take per join, GROUP BY, ORDER BY, LIMIT) returns the
    oracle's rows and the served form's."""
    srv = make_server(star, slots=1, chunk_rows=100_000)
    rng = np.random.default_rng(11)
    for i, spec in enumerate(templates(rng)):
        print(f"[case] composed template {i}", flush=True)
        r = execute_star_query(star["dev"], star["dims_d"], **spec)
        g, want, sabs = reference(star, spec)
        assert r["n_groups"] == g
        same_rows(r["rows"], want, len(spec["group"]), sabs)
        sv = run_query(srv, spec)
        same_rows(r["rows"], sv["rows"], len(spec["group"]), sabs)
    srv.close()


def test_serve_declines_a_regrown_table_over_the_budget(star):
    """A query whose table overflows is re-planned 4x larger; when that no
    longer fits the whole budget it completes with BOLT_DECLINE instead of
    waiting at the queue head for ever, the memory its first attempt held is
    returned exactly, and later queries still run within the budget."""
    fits = dict(terms=[], joins=[(FK3, 2)], group=[(0, 1)], aggs=[(N.SUM, QTY), (N.COUNT_STAR, 0)], order_agg=1,
                limit=20, est_groups=1)  # <= 100 groups: no overflow at the smallest table
    grows = dict(terms=[], joins=[(FK1, 0)], group=[(0, 0)], aggs=[(N.SUM, QTY), (N.COUNT_STAR, 0)], order_agg=1,
                 limit=20, est_groups=1)  # thousands of groups: overflows the smallest table
    probe = make_server(star, slots=1, chunk_rows=100_000)
    need = run_query(probe, fits)["hbm_bytes"]
    probe.close()
    srv = make_server(star, slots=2, chunk_rows=100_000, hbm_budget=need)
    tickets = [srv.submit(srv.query(**grows))] + [srv.submit(srv.query(**fits)) for _ in range(3)]
    res = [srv.wait(t) for t in tickets]
    st = srv.stats()
    srv.close()
    assert res[0]["status"] == N.BOLT_DECLINE, res[0]
    assert all(r["status"] == N.BOLT_OK for r in res[1:])
    assert st["peak_bytes"] <= need
    g, want, sabs = reference(star, fits)
    for r in res[2:]:  # SUM(qty) is a float sum: equal within the tolerance, not bit for bit
        same_rows(r["rows"], res[1]["rows"], 1, sabs)
    assert res[1]["n_groups"] == g
    same_rows(res[1]["rows"], want, 1, sabs)
