"""GPU parity: hash join build / probe / fused filter+join+aggregate.

Join cardinalities and row sets are bit-exact against the oracle's host
INNER join (compared after canonical ordering of the pairs); integer sums and
counts exact; the float sum within REL_TOL of the oracle's Neumaier sum."""
import math

import numpy as np
import pytest

import oracle as O
from craton_bolt_amd import _native as N
from craton_bolt_amd import datagen as D
from craton_bolt_amd.device import DeviceColumn
from craton_bolt_amd.ops import JoinTable, Term, execute_inner_join

from gpu_util import both, rel_close

REL_TOL = 1e-9
pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, params=["jit", "generic", "jit+bloom", "partitioned"])
def kernel_path(request, monkeypatch):
    """Every case runs through the per-query hipRTC kernel, the precompiled
    generic kernel (BOLT_JIT=0), the hipRTC kernel behind the build side's
    Bloom filter (BOLT_JOIN_BLOOM=1), and with the fused probe + aggregate
    forced onto the radix-partitioned form (BOLT_JOIN_PARTITIONED=1; the
    random form otherwise, chosen only for large tables)."""
    monkeypatch.setenv("BOLT_JIT", "0" if request.param == "generic" else "1")
    monkeypatch.setenv("BOLT_JOIN_BLOOM", "1" if request.param == "jit+bloom" else "0")
    monkeypatch.setenv("BOLT_JOIN_PARTITIONED", "1" if request.param == "partitioned" else "0")
    return request.param


def _canon(a, b):
    order = np.lexsort((b, a))
    return np.stack([a[order], b[order]])


def test_gpu_join_e2e_fixture(gpu):
    # tests/gpu_join_e2e.rs:68-142: 4K unique build keys, probe i % 8192
    bk = np.arange(4096, dtype=np.int32)
    pk = (np.arange(8192) % 8192).astype(np.int32)
    db, hb = both(bk, N.I32)
    dp, hp = both(pk, N.I32)
    l, r = execute_inner_join(db, dp)
    assert len(l) == 4096
    assert np.all(bk[l] == pk[r])
    assert np.all(1000 + bk[l] == 1000 + pk[r])
    ol, orr = O.inner_join(hb, hp)
    assert np.array_equal(_canon(l, r), _canon(ol, orr))


@pytest.mark.parametrize("n_left,n_right,card", [(1, 1, 1), (300, 500, 40), (20_000, 5_000, 3_000),
                                                 (100_000, 1_000_000, 200_000), (0, 10, 5)])
def test_inner_join_matches_oracle_with_duplicates_and_nulls(gpu, n_left, n_right, card):
    print(f"[case] inner_join {n_left}x{n_right} card={card}", flush=True)
    rng = np.random.default_rng(n_left + n_right)
    lk = rng.integers(0, card, n_left).astype(np.int64) * 1_000_003 - 7
    rk = rng.integers(0, card, n_right).astype(np.int64) * 1_000_003 - 7
    lv = rng.random(n_left) > 0.05
    rv = rng.random(n_right) > 0.05
    dl, hl = both(lk, N.I64, lv)
    dr, hr = both(rk, N.I64, rv)
    l, r = execute_inner_join(dl, dr)
    ol, orr = O.inner_join(hl, hr)
    assert len(l) == len(ol)
    assert np.array_equal(_canon(l, r), _canon(ol, orr))


def test_join_keys_at_int64_extremes(gpu):
    # no sentinel: i64::MIN and i64::MAX are ordinary keys
    bk = np.array([np.iinfo(np.int64).min, np.iinfo(np.int64).max, 0, -1], np.int64)
    pk = np.array([0, np.iinfo(np.int64).min, 5, np.iinfo(np.int64).max, np.iinfo(np.int64).min], np.int64)
    db, hb = both(bk, N.I64)
    dp, hp = both(pk, N.I64)
    l, r = execute_inner_join(db, dp)
    ol, orr = O.inner_join(hb, hp)
    assert np.array_equal(_canon(l, r), _canon(ol, orr))
    assert len(l) == 4


def test_build_uniqueness_flag(gpu):
    t = JoinTable(DeviceColumn.from_numpy(np.arange(100_000, dtype=np.int64)))
    assert t.unique()
    t = JoinTable(DeviceColumn.from_numpy(np.array([1, 2, 3, 2], np.int64)))
    assert not t.unique()


def test_probe_capacity_declines(gpu):
    t = JoinTable(DeviceColumn.from_numpy(np.zeros(10, np.int64)))
    with pytest.raises(N.GpuCapacity):
        t.probe([DeviceColumn.from_numpy(np.zeros(10, np.int64))], 0, capacity=50)


@pytest.mark.parametrize("n_probe", [2047, 2048, 2049, 300_001])
def test_probe_pairs_in_probe_row_order(gpu, n_probe):
    """Pairs come out sorted by probe row across the 2048-row tiles, with
    duplicate build keys, NULL probe keys and a filter."""
    print(f"[case] probe order n_probe={n_probe}", flush=True)
    rng = np.random.default_rng(n_probe)
    bk = rng.integers(0, 900, 3000).astype(np.int64)
    pk = rng.integers(0, 1200, n_probe).astype(np.int64)
    pv = rng.random(n_probe) > 0.1
    amt = rng.uniform(0, 100, n_probe)
    t = JoinTable(DeviceColumn.from_numpy(bk))
    pr, br = t.probe([DeviceColumn.from_numpy(pk, valid=pv), DeviceColumn.from_numpy(amt)], 0,
                     [Term(1, N.GT, 25.0)], capacity=8 * n_probe)
    assert np.all(np.diff(pr.astype(np.int64)) >= 0)
    keep = pv & (amt > 25.0)
    ol, orr = O.inner_join(O.HostCol(bk, O.I64), O.HostCol(pk, O.I64, pv))
    m = keep[orr]
    assert np.array_equal(_canon(pr, br), _canon(orr[m].astype(np.uint32), ol[m].astype(np.uint32)))


@pytest.mark.parametrize("n_build,n_probe", [(1, 1), (1000, 10_000), (100_000, 2_000_000), (1 << 20, 3_000_001)])
def test_filter_join_aggregate_matches_oracle(gpu, n_build, n_probe):
    """The star-query shape: WHERE amount > 10, INNER JOIN, SUM(payload), SUM(amount), COUNT(*)."""
    print(f"[case] filter_join_agg build={n_build} probe={n_probe}", flush=True)
    bk, bp = D.join_build(n_build)
    fk, amt = D.join_probe(n_build, n_probe)
    want = O.filter_join_agg(bk, bp, fk, amt, 10.0)
    t = JoinTable(DeviceColumn.from_numpy(bk), carry=DeviceColumn.from_numpy(bp))
    assert t.unique()
    cols = [DeviceColumn.from_numpy(fk), DeviceColumn.from_numpy(amt)]
    got = t.probe_aggregate(cols, 0, [(N.SUM, N.JOIN_CARRY), (N.SUM, 1), (N.COUNT_STAR, 0)],
                            [Term(1, N.GT, 10.0)])
    assert got[0] == want["sum_payload"]
    assert got[2] == want["count"]
    if want["count"]:
        assert rel_close(got[1], want["sum_amount"], REL_TOL)
    else:
        assert got[1] is None
    # the pair-emitting probe agrees on the row set
    t2 = JoinTable(DeviceColumn.from_numpy(bk))
    pr, br = t2.probe(cols, 0, [Term(1, N.GT, 10.0)])
    assert len(pr) == want["count"]
    assert np.all(bk[br] == fk[pr]) and np.all(amt[pr] > 10.0)
    assert len(np.unique(pr)) == len(pr)


def test_join_aggregate_with_duplicate_build_keys(gpu):
    rng = np.random.default_rng(77)
    bk = rng.integers(0, 500, 5000).astype(np.int64)
    bp = rng.integers(0, 100, 5000).astype(np.int32)
    pk = rng.integers(0, 700, 50_000).astype(np.int64)
    amt = rng.uniform(0, 100, 50_000)
    t = JoinTable(DeviceColumn.from_numpy(bk), carry=DeviceColumn.from_numpy(bp))
    assert not t.unique()
    got = t.probe_aggregate([DeviceColumn.from_numpy(pk), DeviceColumn.from_numpy(amt)], 0,
                            [(N.SUM, N.JOIN_CARRY), (N.COUNT_STAR, 0), (N.MIN, N.JOIN_CARRY), (N.SUM, 1)])
    ol, orr = O.inner_join(O.HostCol(bk, O.I64), O.HostCol(pk, O.I64))
    assert got[0] == int(bp[ol].astype(np.int64).sum())
    assert got[1] == len(ol)
    assert got[2] == int(bp[ol].min())
    assert rel_close(got[3], math.fsum(amt[orr]), REL_TOL)


@pytest.mark.parametrize("load", ["0.15", "0.9"])
def test_region_build_matches_atomic_build(gpu, monkeypatch, load, kernel_path):
    """The region build (tables of 2^20 slots and more) and the global-claim
    build give the same join: duplicates, NULL keys, and at load 0.9 many rows
    that overflow their region."""
    if kernel_path != "jit":
        pytest.skip("the build does not depend on the probe kernel")
    print(f"[case] region build load={load}", flush=True)
    monkeypatch.setenv("BOLT_JOIN_LOAD_FACTOR", load)
    rng = np.random.default_rng(99)
    nb, npr = 1_500_000, 3_000_000
    bk = rng.integers(0, 1_200_000, nb).astype(np.int64) * 7 - 3
    pk = rng.integers(0, 1_400_000, npr).astype(np.int64) * 7 - 3
    bv = rng.random(nb) > 0.02
    db, _ = both(bk, N.I64, bv)
    dp, _ = both(pk, N.I64)
    pairs = {}
    for how in ("region", "atomic"):
        monkeypatch.setenv("BOLT_JOIN_BUILD", how)
        t = JoinTable(db)
        assert not t.unique()
        pr, br = t.probe([dp], 0, capacity=4 * npr)
        pairs[how] = _canon(pr, br)
        t.release()
    assert np.array_equal(pairs["region"], pairs["atomic"])
    # unique keys: the region build's flag says unique
    monkeypatch.setenv("BOLT_JOIN_BUILD", "region")
    du, _ = both(rng.permutation(nb).astype(np.int64), N.I64)
    t = JoinTable(du)
    assert t.unique()
    t.release()


@pytest.mark.parametrize("jit", ["1", "0"])
@pytest.mark.parametrize("key_dt", [N.I64, N.I32])
@pytest.mark.parametrize("part_bits", ["3", "6", "10"])
def test_partitioned_probe_aggregate_matches_oracle(gpu, kernel_path, monkeypatch, key_dt, part_bits, jit):
    """The radix-partitioned probe + aggregate against the oracle's host join:
    duplicate build keys, NULL probe keys, three filter terms over two
    columns, two probe columns carried with the key (Float64, Int32), the
    carried build column, and five aggregates (one generated probe, or two
    launches of the precompiled one), through the hipRTC-generated scatter
    and probe and through the precompiled kernels (BOLT_JIT=0)."""
    if kernel_path != "partitioned":
        pytest.skip("a test of the partitioned form")
    monkeypatch.setenv("BOLT_JIT", jit)
    monkeypatch.setenv("BOLT_JOIN_PART_BITS", part_bits)
    print(f"[case] partitioned key={key_dt} bits={part_bits} jit={jit}", flush=True)
    rng = np.random.default_rng(int(part_bits) * 10 + key_dt)
    nb, npr = 400_000, 2_500_000
    np_dt = np.int64 if key_dt == N.I64 else np.int32
    bk = (rng.integers(0, 300_000, nb) * 13 - 1000).astype(np_dt)
    bp = rng.integers(-50, 1000, nb).astype(np.int32)
    pk = (rng.integers(0, 450_000, npr) * 13 - 1000).astype(np_dt)
    pv = rng.random(npr) > 0.03
    amt = rng.uniform(0, 100, npr)
    q = rng.integers(-20, 20, npr).astype(np.int32)
    t = JoinTable(DeviceColumn.from_numpy(bk), carry=DeviceColumn.from_numpy(bp))
    assert not t.unique()
    cols = [DeviceColumn.from_numpy(pk, valid=pv), DeviceColumn.from_numpy(amt), DeviceColumn.from_numpy(q)]
    aggs = [(N.SUM, N.JOIN_CARRY), (N.SUM, 1), (N.COUNT_STAR, 0), (N.MAX, 2), (N.SUM, 2)]
    terms = [Term(1, N.GT, 10.0), Term(1, N.LE, 95.0), Term(2, N.NE, 0)]
    got = t.probe_aggregate(cols, 0, aggs, terms)
    okd = O.I64 if key_dt == N.I64 else O.I32
    ol, orr = O.inner_join(O.HostCol(bk, okd), O.HostCol(pk, okd, pv))
    m = (amt[orr] > 10.0) & (amt[orr] <= 95.0) & (q[orr] != 0)
    ol, orr = ol[m], orr[m]
    assert got[0] == int(bp[ol].astype(np.int64).sum())
    assert rel_close(got[1], math.fsum(amt[orr]), REL_TOL)
    assert got[2] == len(ol)
    assert got[3] == int(q[orr].max())
    assert got[4] == int(q[orr].astype(np.int64).sum())


@pytest.mark.parametrize("jit", ["1", "0"])
def test_partitioned_probe_aggregate_without_probe_columns(gpu, kernel_path, monkeypatch, jit):
    """The partitioned form when no aggregate reads a probe column, so nothing
    is carried with the key: COUNT(*), SUM and MAX of the build side's carried
    column, behind a filter on a probe column; duplicate build keys."""
    if kernel_path != "partitioned":
        pytest.skip("a test of the partitioned form")
    monkeypatch.setenv("BOLT_JIT", jit)
    print(f"[case] partitioned, no carried probe column, jit={jit}", flush=True)
    rng = np.random.default_rng(5)
    nb, npr = 200_000, 1_500_000
    bk = (rng.integers(0, 150_000, nb) * 7 + 3).astype(np.int64)
    bp = rng.integers(-100, 100, nb).astype(np.int32)
    pk = (rng.integers(0, 200_000, npr) * 7 + 3).astype(np.int64)
    amt = rng.uniform(0, 100, npr)
    t = JoinTable(DeviceColumn.from_numpy(bk), carry=DeviceColumn.from_numpy(bp))
    got = t.probe_aggregate([DeviceColumn.from_numpy(pk), DeviceColumn.from_numpy(amt)], 0,
                            [(N.COUNT_STAR, 0), (N.SUM, N.JOIN_CARRY), (N.MAX, N.JOIN_CARRY)],
                            [Term(1, N.GT, 30.0)])
    ol, orr = O.inner_join(O.HostCol(bk, O.I64), O.HostCol(pk, O.I64))
    ol = ol[amt[orr] > 30.0]
    assert got[0] == len(ol)
    assert got[1] == int(bp[ol].astype(np.int64).sum())
    assert got[2] == int(bp[ol].max())
