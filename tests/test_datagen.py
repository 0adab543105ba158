"""The parallel table generators reproduce the sequential Xorshift64Star
stream of the oracle exactly (GF(2) jump-ahead), and the workloads have the
shapes their specifications state."""
import numpy as np

import oracle as O
from craton_bolt_amd import datagen as D


def test_parallel_draws_equal_sequential_stream():
    seq = O.rng_draws(0xB0170003, 100_000)
    assert np.array_equal(D.draws(0xB0170003, 100_000), seq)
    # jumping to an offset lands on the same draw
    assert np.array_equal(D.draws(0xB0170003, 1000, offset=54_321), seq[54_321:55_321])


def test_groupby_small_rows():
    k, v = D.groupby_small(5000)
    d = O.rng_draws(D.SEED_GROUPBY_SMALL, 10_000)
    assert np.array_equal(k, (d[0::2] % np.uint64(1000)).astype(np.int32))
    assert np.array_equal(v, (d[1::2] % np.uint64(1000)).astype(np.int64))


def test_join_tables_shape():
    nb, npb = 10_000, 50_000
    bk, bp = D.join_build(nb)
    assert len(np.unique(bk)) == nb
    assert bp.min() >= 0 and bp.max() < 1_000_000
    d = O.rng_draws(D.SEED_JOIN, nb)
    assert np.array_equal(bp, (d % np.uint64(1_000_000)).astype(np.int32))
    fk, amt = D.join_probe(nb, npb)
    hit = np.isin(fk, bk)
    assert 0.47 < hit.mean() < 0.53
    assert amt.min() >= 0 and amt.max() < 100
    # the hit decision is the top bit of the first draw of each row
    d = O.rng_draws(D.SEED_JOIN, nb + 3 * npb)[nb:]
    assert np.array_equal(hit, (d[0::3] >> np.uint64(63)).astype(bool))


def test_sort_pairs_and_zipf_shape():
    key, rid = D.sort_pairs(1000)
    assert np.array_equal(key.view(np.uint64), O.rng_draws(D.SEED_SORT, 1000))
    assert np.array_equal(rid, np.arange(1000, dtype=np.uint32))
    k, v = D.zipf_groupby(200_000, n_ranks=10_000)
    _, counts = np.unique(k, return_counts=True)
    top = np.sort(counts)[::-1]
    # Zipf(1.1): rank-1 share = 1 / H(10^4, 1.1)
    h = np.sum(np.arange(1, 10_001) ** -1.1)
    assert abs(top[0] / 200_000 - 1 / h) < 0.01
    assert abs(v.mean()) < 0.02 and abs(v.std() - 1) < 0.02


def test_h2o_columns_follow_the_reference_generator():
    """benches/olap_benchmarks.rs:106-123, restated with Python integers: a
    64-bit wrapping product, its low 32 bits as a signed int, then rem_euclid."""
    import numpy as np

    from craton_bolt_amd import datagen as D

    def low32(v):
        v &= (1 << 32) - 1
        return v - (1 << 32) if v >> 31 else v

    def rem_euclid(a, m):
        return a % m  # Python's % is Euclidean for a positive modulus

    rows = [0, 1, 2, 3, 4095, 65_537, 1_000_003, 9_999_999]
    n = rows[-1] + 1
    cols = {c: D.h2o_column(c, n) for c in ("id1", "id2", "id3", "v1", "v2", "v3")}
    m64 = (1 << 64) - 1
    for i in rows:
        assert cols["id1"][i] == rem_euclid(low32((i * 2_654_435_761) & m64), 100)
        assert cols["id2"][i] == rem_euclid(low32((i * 40_503) & m64), 10_000)
        assert cols["id3"][i] == rem_euclid(low32((i * 11_400_714_819_323_198_485) & m64), 1_000_000)
        assert cols["v1"][i] == float(rem_euclid(low32(i * 7), 5) + 1)
        assert cols["v2"][i] == float(rem_euclid(low32(i * 13), 15) + 1)
        assert cols["v3"][i] == rem_euclid(low32(i * 17), 10_000) / 100.0
    assert cols["id1"].dtype == np.int32 and cols["v3"].dtype == np.float64
    assert cols["id3"].min() >= 0 and cols["id3"].max() < 1_000_000
