"""GPU parity: filter + stream compaction and the mask prefix scan, against
the oracle's filter (src/exec/compact.rs:146 semantics).  Row sets and values
are bit-exact, order preserved."""
import numpy as np
import pytest

import oracle as O
from craton_bolt_amd import _native as N
from craton_bolt_amd.device import DeviceBuffer
from craton_bolt_amd.ops import Term, compact_columns, prefix_scan_mask

from gpu_util import both, dec_array

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 100, 4096, 4097, 1_000_003, 9_000_000])
def test_compact_matches_oracle(gpu, n):
    print(f"[case] compact n={n}", flush=True)
    rng = np.random.default_rng(n)
    a = rng.uniform(0, 100, n)
    b = rng.integers(-1000, 1000, n).astype(np.int32)
    c = rng.integers(-2 ** 62, 2 ** 62, n).astype(np.int64)
    vb = rng.random(n) > 0.1
    da, ha = both(a, N.F64)
    db, hb = both(b, N.I32, vb)
    dc, hc = both(c, N.I64)
    outs, k = compact_columns([da, db, dc], [Term(0, N.GT, 10.0), Term(1, N.LT, 500)])
    rows = O.filter_rows([ha, hb, hc], [(0, O.GT, 10.0), (1, O.LT, 500)], n)
    assert k == len(rows)
    assert np.array_equal(outs[0].to_numpy(), a[rows])
    assert np.array_equal(outs[1].to_numpy(), b[rows])
    assert np.array_equal(outs[1].valid_numpy(), vb[rows])
    assert np.array_equal(outs[2].to_numpy(), c[rows])


def test_compact_with_mask_and_nulls(gpu):
    rng = np.random.default_rng(3)
    n = 300_001
    x = rng.normal(size=n).astype(np.float32)
    vx = rng.random(n) > 0.3
    mask = (rng.random(n) > 0.5).astype(np.uint8)
    dx, hx = both(x, N.F32, vx)
    outs, k = compact_columns([dx], [Term(0, N.IS_NOT_NULL)], mask=DeviceBuffer.from_numpy(mask))
    want = np.nonzero(vx & (mask != 0))[0]
    assert k == len(want)
    assert np.array_equal(outs[0].to_numpy(), x[want])
    assert outs[0].valid_numpy().all()
    outs, k = compact_columns([dx], [Term(0, N.IS_NULL)])
    assert k == int((~vx).sum())


def test_compact_decimal_and_none_pass(gpu):
    vals = [10 ** 30, -5, 7, 10 ** 20 + 3]
    dd, _ = both(dec_array(vals), N.DEC128)
    outs, k = compact_columns([dd], [Term(0, N.GT, 6)])
    got = [(int(h) << 64) | (int(l) & (2 ** 64 - 1)) for l, h in outs[0].to_numpy(k)]
    assert got == [10 ** 30, 7, 10 ** 20 + 3]
    outs, k = compact_columns([dd], [Term(0, N.GT, 10 ** 31)])
    assert k == 0


@pytest.mark.parametrize("n", [1, 4096, 5000, 10_000_000])
def test_prefix_scan_mask(gpu, n):
    rng = np.random.default_rng(n + 1)
    mask = (rng.random(n) > 0.3).astype(np.uint8)
    off, total = prefix_scan_mask(DeviceBuffer.from_numpy(mask), n)
    want = np.concatenate([[0], np.cumsum(mask, dtype=np.int64)[:-1]]).astype(np.uint32)
    assert total == int(mask.sum())
    assert np.array_equal(off.to_numpy(np.uint32, n), want)
