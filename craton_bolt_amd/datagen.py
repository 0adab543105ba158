"""Seeded synthetic tables (numpy) from libbolt_datagen.so.

Each generator consumes a fixed number of Xorshift64Star draws per row, so
row r depends only on (seed, r) and the tables are filled by all host
threads (see csrc/datagen.c).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libbolt_datagen.so"
_lib = None

# seeds of the operator workloads
SEED_GROUPBY_SMALL = 0xB0170001
SEED_JOIN = 0xB0170002
SEED_SORT = 0xB0170003
SEED_ZIPF = 0xB0170004
SEED_BATCH = 0xB0170005
SEED_SERVE = 0xB0170006

_u64, _i64, _p = C.c_uint64, C.c_int64, C.c_void_p


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m craton_bolt_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        L.bolt_gen_state_at.restype = _u64
        L.bolt_gen_state_at.argtypes = [_u64, _u64]
        L.bolt_gen_draws.argtypes = [_u64, _u64, _i64, _p]
        L.bolt_scramble64.restype = _u64
        L.bolt_scramble64.argtypes = [_u64]
        L.bolt_scramble32.restype = C.c_uint32
        L.bolt_scramble32.argtypes = [C.c_uint32]
        L.bolt_gen_groupby_small.argtypes = [_u64, _i64, _u64, _u64, _p, _p]
        L.bolt_gen_join_build.argtypes = [_u64, _i64, _u64, _p, _p]
        L.bolt_gen_join_probe.argtypes = [_u64, _i64, _i64, C.c_double, _p, _p]
        L.bolt_gen_sort_pairs.argtypes = [_u64, _i64, _p, _p]
        L.bolt_gen_zipf_groupby.argtypes = [_u64, _i64, _i64, C.c_double, _p, _p]
        L.bolt_gen_zipf_small.argtypes = [_u64, _i64, _i64, C.c_double, _p, _p, _p]
        L.bolt_gen_star_fact.argtypes = [_u64, _i64, _i64, _i64, _i64, C.c_int32, _p, _p, _p, _p, _p, _p]
        _lib = L
    return _lib


def draws(seed: int, n: int, offset: int = 0) -> np.ndarray:
    out = np.empty(n, dtype=np.uint64)
    lib().bolt_gen_draws(seed, offset, n, out.ctypes.data)
    return out


def groupby_small(n: int, card: int = 1000, vmod: int = 1000, seed: int = SEED_GROUPBY_SMALL):
    """k int32 = draw % card, v int64 = draw % vmod (two draws per row)."""
    k = np.empty(n, np.int32)
    v = np.empty(n, np.int64)
    lib().bolt_gen_groupby_small(seed, n, card, vmod, k.ctypes.data, v.ctypes.data)
    return k, v


def join_build(n_build: int, pmod: int = 1_000_000, seed: int = SEED_JOIN):
    key = np.empty(n_build, np.int64)
    payload = np.empty(n_build, np.int32)
    lib().bolt_gen_join_build(seed, n_build, pmod, key.ctypes.data, payload.ctypes.data)
    return key, payload


def join_probe(n_build: int, n_probe: int, amount_scale: float = 100.0, seed: int = SEED_JOIN,
               out_fk: np.ndarray | None = None, out_amount: np.ndarray | None = None):
    fk = out_fk if out_fk is not None else np.empty(n_probe, np.int64)
    amount = out_amount if out_amount is not None else np.empty(n_probe, np.float64)
    lib().bolt_gen_join_probe(seed, n_build, n_probe, amount_scale, fk.ctypes.data, amount.ctypes.data)
    return fk, amount


def sort_pairs(n: int, seed: int = SEED_SORT):
    key = np.empty(n, np.int64)
    rid = np.empty(n, np.uint32)
    lib().bolt_gen_sort_pairs(seed, n, key.ctypes.data, rid.ctypes.data)
    return key, rid


def zipf_groupby(n: int, n_ranks: int = 10_000_000, s: float = 1.1, seed: int = SEED_ZIPF):
    key = np.empty(n, np.int64)
    val = np.empty(n, np.float64)
    lib().bolt_gen_zipf_groupby(seed, n, n_ranks, s, key.ctypes.data, val.ctypes.data)
    return key, val


def zipf_small(n: int, card: int, s: float = 1.1, seed: int = SEED_BATCH):
    key = np.empty(n, np.int32)
    vi = np.empty(n, np.int64)
    vf = np.empty(n, np.float32)
    lib().bolt_gen_zipf_small(seed, n, card, s, key.ctypes.data, vi.ctypes.data, vf.ctypes.data)
    return key, vi, vf


def batch_shapes(n_queries: int = 512, seed: int = SEED_BATCH) -> list[tuple[int, int]]:
    """(rows, key cardinality) of each batched query: both log-uniform, rows over
    [1e4, 4e6] and cardinality over [16, 1e6], from two draws of seed + q."""
    import math

    out = []
    for q in range(n_queries):
        u = (draws(seed + q, 2) >> np.uint64(11)).astype(np.float64) / 2.0 ** 53
        n = int(round(math.exp(math.log(1e4) + u[0] * (math.log(4e6) - math.log(1e4)))))
        c = int(round(math.exp(math.log(16) + u[1] * (math.log(1e6) - math.log(16)))))
        out.append((n, c))
    return out


STAR_DIMS = (1_000_000, 100_000, 2_500)
STAR_DATE0 = 17532  # 2018-01-01 as days since the epoch


def star_fact(n: int, dims=STAR_DIMS, seed: int = SEED_SERVE) -> dict:
    """Fact table of the served workload: fk1/fk2/fk3 int32, price Decimal128(12,2)
    as an (n, 2) int64 [lo, hi] array of cents, qty float64, date Date32."""
    out = {"fk1": np.empty(n, np.int32), "fk2": np.empty(n, np.int32), "fk3": np.empty(n, np.int32),
           "price": np.empty((n, 2), np.int64), "qty": np.empty(n, np.float64), "date": np.empty(n, np.int32)}
    lib().bolt_gen_star_fact(seed, n, dims[0], dims[1], dims[2], STAR_DATE0, out["fk1"].ctypes.data,
                             out["fk2"].ctypes.data, out["fk3"].ctypes.data, out["price"].ctypes.data,
                             out["qty"].ctypes.data, out["date"].ctypes.data)
    return out


STAR_ATTRS = (("brand", 100_000, "a1000", 1000), ("city", 10_000, "zone", 1000, "region", 50),
              ("cat", 25, "sub", 250))


def star_dim(d: int, n: int) -> tuple[np.ndarray, list[np.ndarray]]:
    """Dimension d (0, 1, 2) with n rows: int32 keys are the permutation
    (r * 1000003 + 12345) mod n of the row ids (so the join is a real key
    lookup), attributes are key mod a fixed modulus (dim 0: brand = key % 100000,
    key % 1000; dim 1: city = key % 10000, zone = key % 1000, region = key % 50;
    dim 2: cat = key % 25, sub = key % 250)."""
    r = np.arange(n, dtype=np.int64)
    key = ((r * 1000003 + 12345) % n).astype(np.int32)
    mods = STAR_ATTRS[d][1::2]
    return key, [(key.astype(np.int64) % m).astype(np.int32) for m in mods]


def scramble64(x: int) -> int:
    return int(lib().bolt_scramble64(x))


# ---- h2o.ai groupby "small" columns (benches/olap_benchmarks.rs:63-123) -----------------------------
#
# Row i's values are fixed functions of i (no seed): a 64-bit wrapping product,
# truncated to its low 32 bits as a signed int, then a Euclidean remainder.

H2O_CARD = {"id1": 100, "id2": 10_000, "id3": 1_000_000}
_H2O_MUL = {"id1": 2_654_435_761, "id2": 40_503, "id3": 11_400_714_819_323_198_485, "v1": 7, "v2": 13, "v3": 17}


def _h2o_i32(i: np.ndarray, mult: int, modulus: int) -> np.ndarray:
    low = (i * np.uint64(mult)).astype(np.uint32).view(np.int32).astype(np.int64)
    return low % modulus  # non-negative for a positive modulus (rem_euclid)


def h2o_column(name: str, n: int) -> np.ndarray:
    """Column `name` (id1..id3 Int32, v1..v3 Float64) of the h2o groupby table of n rows."""
    i = np.arange(n, dtype=np.uint64)
    if name in H2O_CARD:
        return _h2o_i32(i, _H2O_MUL[name], H2O_CARD[name]).astype(np.int32)
    if name == "v1":
        return (_h2o_i32(i, 7, 5) + 1).astype(np.float64)
    if name == "v2":
        return (_h2o_i32(i, 13, 15) + 1).astype(np.float64)
    return _h2o_i32(i, 17, 10_000).astype(np.float64) / 100.0
