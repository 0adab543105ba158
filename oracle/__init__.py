"""CPU oracle (test infrastructure): ctypes binding of oracle/bolt_oracle.c.

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
use this package, and only as the checker.  Each wrapper names the
reference file:line its C function restates (see bolt_oracle.h).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "build" / "libbolt_oracle.so"

BOOL, I32, I64, F32, F64, DEC128, DATE32, TIMESTAMP = 1, 2, 3, 4, 5, 6, 7, 8
SUM, COUNT, COUNT_STAR, MIN, MAX, AVG, VAR_POP, VAR_SAMP, STDDEV_POP, STDDEV_SAMP = range(1, 11)
LT, LE, GT, GE, EQ, NE, IS_NULL, IS_NOT_NULL = range(1, 9)
OK, ERR_PLAN, ERR_TYPE, ERR_CAP = 0, 1, 2, 3


class Col(C.Structure):
    _fields_ = [("data", C.c_void_p), ("validity", C.c_void_p), ("n", C.c_int64), ("dtype", C.c_int32),
                ("scale", C.c_int32)]


class Scalar(C.Structure):
    _fields_ = [("status", C.c_int32), ("is_null", C.c_int32), ("dtype", C.c_int32), ("pad", C.c_int32),
                ("i64", C.c_int64), ("f64", C.c_double), ("hi", C.c_int64)]


class TermS(C.Structure):
    _fields_ = [("col", C.c_int32), ("op", C.c_int32), ("lit_i64", C.c_int64), ("lit_f64", C.c_double),
                ("lit_hi", C.c_int64)]


_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        L.bo_rng_seed.restype = C.c_uint64
        L.bo_rng_seed.argtypes = [C.c_uint64]
        L.bo_rng_next.restype = C.c_uint64
        L.bo_rng_next.argtypes = [C.POINTER(C.c_uint64)]
        L.bo_rng_unit.restype = C.c_double
        L.bo_rng_unit.argtypes = [C.POINTER(C.c_uint64)]
        L.bo_rng_fill.argtypes = [C.c_uint64, C.c_void_p, C.c_int64]
        L.bo_shuffle_u32.argtypes = [C.c_void_p, C.c_int64, C.c_uint64]
        L.bo_neumaier_sum.restype = C.c_double
        L.bo_neumaier_sum.argtypes = [C.c_void_p, C.c_int64]
        L.bo_scalar_agg.restype = Scalar
        L.bo_scalar_agg.argtypes = [C.POINTER(Col), C.c_int32]
        L.bo_filter_rows.restype = C.c_int64
        L.bo_filter_rows.argtypes = [C.POINTER(Col), C.c_int32, C.POINTER(TermS), C.c_int32, C.c_int64, C.c_void_p]
        L.bo_groupby.restype = C.c_int32
        L.bo_groupby.argtypes = [C.POINTER(Col), C.c_int32, C.POINTER(Col), C.POINTER(C.c_int32),
                                 C.POINTER(C.c_int32), C.c_int32, C.POINTER(C.c_int64), C.c_void_p, C.c_void_p,
                                 C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                 C.POINTER(C.c_void_p), C.c_int64]
        L.bo_inner_join.restype = C.c_int64
        L.bo_inner_join.argtypes = [C.POINTER(Col), C.POINTER(Col), C.c_void_p, C.c_void_p, C.c_int64,
                                    C.POINTER(C.c_int64)]
        L.bo_project.restype = C.c_int32
        L.bo_project.argtypes = [C.POINTER(Col), C.c_int32, C.POINTER(ExprOp), C.c_int32, C.POINTER(Col), C.c_int32,
                                 C.c_int64]
        L.bo_join.restype = C.c_int64
        L.bo_join.argtypes = [C.POINTER(Col), C.POINTER(Col), C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                              C.c_int64, C.POINTER(C.c_int64)]
        L.bo_sort_indices.argtypes = [C.POINTER(Col), C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                      C.c_int64, C.c_void_p]
        L.bo_filter_join_agg.restype = C.c_int32
        L.bo_filter_join_agg.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int64,
                                         C.c_double, C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                         C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        _lib = L
    return _lib


# ---- host column helper ---------------------------------------------------------

class ExprOp(C.Structure):
    """bo_expr_op (same layout as the product's bolt_expr_op_t)."""
    _fields_ = [
        ("kind", C.c_int32), ("dst", C.c_int32), ("a", C.c_int32), ("b", C.c_int32), ("c", C.c_int32),
        ("col", C.c_int32), ("op", C.c_int32), ("dtype", C.c_int32),
        ("lit_i64", C.c_int64), ("lit_f64", C.c_double), ("lit_hi", C.c_int64),
    ]


class HostCol:
    """Keeps numpy buffers alive behind a `Col` descriptor."""

    def __init__(self, values: np.ndarray, dtype: int, valid: np.ndarray | None = None, scale: int = 0):
        self.values = np.ascontiguousarray(values)
        self.bitmap = np.packbits(np.asarray(valid, bool), bitorder="little") if valid is not None else None
        n = self.values.shape[0]
        self.col = Col(self.values.ctypes.data, self.bitmap.ctypes.data if self.bitmap is not None else None, n,
                       dtype, scale)


def _cols(cols: list[HostCol]):
    arr = (Col * max(len(cols), 1))()
    for i, c in enumerate(cols):
        arr[i] = c.col
    return arr


# ---- wrappers ---------------------------------------------------------------------

def rng_draws(seed: int, n: int) -> np.ndarray:
    """tests/common/mod.rs:46-90 Xorshift64Star, n successive next_u64()."""
    out = np.empty(n, np.uint64)
    lib().bo_rng_fill(seed, out.ctypes.data, n)
    return out


def rng_unit(seed: int, n: int) -> np.ndarray:
    s = C.c_uint64(lib().bo_rng_seed(seed))
    return np.array([lib().bo_rng_unit(C.byref(s)) for _ in range(n)])


def shuffle_u32(xs: np.ndarray, seed: int) -> np.ndarray:
    """tests/common/mod.rs:102-110."""
    a = np.ascontiguousarray(xs, dtype=np.uint32).copy()
    lib().bo_shuffle_u32(a.ctypes.data, a.shape[0], seed)
    return a


def neumaier_sum(x: np.ndarray) -> float:
    """src/exec/aggregate.rs:2001."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    return float(lib().bo_neumaier_sum(x.ctypes.data, x.shape[0]))


def scalar_value(s: Scalar):
    if s.status != OK:
        raise ValueError(f"oracle status {s.status}")
    if s.is_null:
        return None
    if s.dtype in (F32, F64):
        return float(s.f64)
    if s.dtype == DEC128:
        return (int(s.hi) << 64) | (int(s.i64) & 0xFFFFFFFFFFFFFFFF)
    return int(s.i64)


def scalar_agg(col: HostCol, op: int) -> Scalar:
    """src/exec/aggregate.rs:154,2032-2262 (raw struct: check .status)."""
    return lib().bo_scalar_agg(C.byref(col.col), op)


def filter_rows(cols: list[HostCol], terms: list[tuple], n: int) -> np.ndarray:
    """src/exec/compact.rs:146; terms are (col, op, literal)."""
    ta = (TermS * max(len(terms), 1))()
    for i, (c, op, lit) in enumerate(terms):
        t = TermS(c, op, 0, 0.0, 0)
        if cols[c].col.dtype in (F32, F64):
            t.lit_f64 = float(lit)
        else:
            t.lit_i64 = int(lit)
        ta[i] = t
    out = np.empty(max(n, 1), np.uint32)
    k = lib().bo_filter_rows(_cols(cols), len(cols), ta, len(terms), n, out.ctypes.data)
    return out[:k]


def groupby(keys: list[HostCol], inputs: list[HostCol], aggs: list[tuple[int, int]], max_groups: int | None = None):
    """src/exec/groupby.rs:295 classic semantics. Returns dict or raises on error status.
    Outputs are sized for `max_groups` groups (default: one per row); more
    groups than that costs one more fold with outputs of the exact size."""
    n = keys[0].col.n
    cap = max(min(n, max_groups) if max_groups else n, 1)
    na = len(aggs)
    ops = (C.c_int32 * max(na, 1))(*[a[0] for a in aggs])
    colsi = (C.c_int32 * max(na, 1))(*[a[1] for a in aggs])
    out_keys = np.empty(cap, np.int64)
    out_kn = np.empty(cap, np.uint8)
    oi = [np.empty(cap, np.int64) for _ in range(na)]
    of = [np.empty(cap, np.float64) for _ in range(na)]
    oh = [np.empty(cap, np.int64) for _ in range(na)]
    on = [np.empty(cap, np.uint8) for _ in range(na)]

    def ptrs(lst):
        arr = (C.c_void_p * max(na, 1))()
        for i, a in enumerate(lst):
            arr[i] = a.ctypes.data
        return arr

    ng = C.c_int64()
    st = lib().bo_groupby(_cols(keys), len(keys), _cols(inputs), ops, colsi, na, C.byref(ng), out_keys.ctypes.data,
                          out_kn.ctypes.data, ptrs(oi), ptrs(of), ptrs(oh), ptrs(on), cap)
    if st == ERR_CAP:
        return groupby(keys, inputs, aggs, max_groups=ng.value)
    if st != OK:
        return {"status": st}
    g = ng.value
    return {"status": OK, "n_groups": g, "keys": out_keys[:g], "key_null": out_kn[:g].astype(bool),
            "i64": [a[:g] for a in oi], "f64": [a[:g] for a in of], "hi": [a[:g] for a in oh],
            "null": [a[:g].astype(bool) for a in on]}


def inner_join(left: HostCol, right: HostCol) -> tuple[np.ndarray, np.ndarray]:
    """src/exec/join.rs:169 (pairs in the reference's emission order)."""
    need = C.c_int64()
    lib().bo_inner_join(C.byref(left.col), C.byref(right.col), None, None, 0, C.byref(need))
    k = need.value
    ol, orr = np.empty(max(k, 1), np.uint32), np.empty(max(k, 1), np.uint32)
    lib().bo_inner_join(C.byref(left.col), C.byref(right.col), ol.ctypes.data, orr.ctypes.data, k, C.byref(need))
    return ol[:k], orr[:k]


JOIN_TYPES = {"inner": 0, "left": 1, "right": 2, "full": 3}


def join(left: list[HostCol], right: list[HostCol], how: str = "inner") -> tuple[np.ndarray, np.ndarray]:
    """src/exec/join.rs:132 (INNER :169, OUTER :283): (left_row, right_row) pairs in the reference's
    host order, 0xFFFFFFFF on the NULL-padded side."""
    lc, rc = _cols(left), _cols(right)
    need = C.c_int64()
    lib().bo_join(lc, rc, len(left), JOIN_TYPES[how], None, None, 0, C.byref(need))
    k = need.value
    ol, orr = np.empty(max(k, 1), np.uint32), np.empty(max(k, 1), np.uint32)
    lib().bo_join(lc, rc, len(left), JOIN_TYPES[how], ol.ctypes.data, orr.ctypes.data, k, C.byref(need))
    return ol[:k], orr[:k]


def sort_indices(keys: list[HostCol], descending: list[bool], nulls_first: list[bool]) -> np.ndarray:
    """src/exec/sort.rs:161 + gpu_sort.rs:1070-1180 transforms; stable."""
    k = len(keys)
    n = keys[0].col.n
    out = np.empty(max(n, 1), np.uint32)
    d = (C.c_int32 * k)(*[int(x) for x in descending])
    nf = (C.c_int32 * k)(*[int(x) for x in nulls_first])
    lib().bo_sort_indices(_cols(keys), k, d, nf, n, out.ctypes.data)
    return out[:n]


_NP = {BOOL: np.uint8, I32: np.int32, I64: np.int64, F32: np.float32, F64: np.float64, DATE32: np.int32,
       TIMESTAMP: np.int64}


def project(inputs: list[HostCol], ops: list[dict], out_dtypes: list[int], n: int,
            out_validity: list[bool] | None = None, out_scales: list[int] | None = None):
    """Expression program (src/plan/physical_plan.rs:81, semantics of src/jit/ptx_gen.rs) over n rows.
    Returns (status, [values], [validity or None]); Decimal128 values as (n, 2) int64 [lo, hi]."""
    arr = (ExprOp * max(len(ops), 1))()
    for i, op in enumerate(ops):
        for k, v in op.items():
            setattr(arr[i], k, v)
    vals, bms, outs = [], [], []
    for i, dt in enumerate(out_dtypes):
        v = np.zeros((n, 2), np.int64) if dt == DEC128 else np.zeros(n, _NP[dt])
        bm = np.zeros((n + 7) // 8 + 1, np.uint8) if out_validity and out_validity[i] else None
        vals.append(v)
        bms.append(bm)
        outs.append(Col(v.ctypes.data, bm.ctypes.data if bm is not None else None, n, dt,
                        out_scales[i] if out_scales else 0))
    oc = (Col * max(len(outs), 1))(*outs)
    st = lib().bo_project(_cols(inputs), len(inputs), arr, len(ops), oc, len(outs), n)
    valid = [np.unpackbits(b, bitorder="little")[:n].astype(bool) if b is not None else None for b in bms]
    return st, vals, valid


def aggregate_expr(inputs: list[HostCol], ops: list[dict], value_dtypes: list[int], filter_value: int,
                   aggs: list[tuple[int, int]], n: int) -> list:
    """agg_with_pre::try_execute_resident (src/exec/agg_with_pre.rs:157): the pre kernel materialises the
    program's values (project above), the WHERE value selects rows, then each aggregate runs over its value
    (src/exec/aggregate.rs:154).  Returns one Scalar per aggregate (COUNT(*) as an int)."""
    st, vals, valid = project(inputs, ops, value_dtypes, n, [True] * len(value_dtypes))
    if st != OK:
        raise ValueError(f"program status {st}")
    keep = np.ones(n, bool) if filter_value < 0 else (vals[filter_value] != 0) & valid[filter_value]
    out = []
    for op, col in aggs:
        if op == 3:  # COUNT(*)
            out.append(int(keep.sum()))
            continue
        v = vals[col][keep]
        out.append(scalar_agg(HostCol(v, value_dtypes[col], valid[col][keep]), op))
    return out


def limit_window(n: int, limit: int, offset: int) -> tuple[int, int]:
    """execute_limit (src/exec/limit.rs:18): the [start, end) window of LIMIT/OFFSET over n rows."""
    if offset >= n:
        return n, n
    return offset, offset + min(limit, n - offset)


def sort_limit(keys: list[HostCol], descending: list[bool], nulls_first: list[bool], limit: int,
               offset: int = 0) -> np.ndarray:
    """ORDER BY + LIMIT as the reference runs it: the full stable sort (sort.rs:161), then the slice."""
    perm = sort_indices(keys, descending, nulls_first)
    a, b = limit_window(len(perm), limit, offset)
    return perm[a:b]


def filter_join_agg(build_key, build_payload, probe_key, probe_amount, amount_gt: float):
    """filter (compact.rs:146) -> host join (join.rs:169) -> SUM/SUM/COUNT (aggregate.rs:154)."""
    bk = np.ascontiguousarray(build_key, np.int64)
    bp = np.ascontiguousarray(build_payload, np.int32)
    pk = np.ascontiguousarray(probe_key, np.int64)
    pa = np.ascontiguousarray(probe_amount, np.float64)
    sp, sn, sc, cnt = C.c_int64(), C.c_double(), C.c_double(), C.c_int64()
    st = lib().bo_filter_join_agg(bk.ctypes.data, bp.ctypes.data, bk.shape[0], pk.ctypes.data, pa.ctypes.data,
                                  pk.shape[0], amount_gt, C.byref(sp), C.byref(sn), C.byref(sc), C.byref(cnt))
    return {"status": st, "sum_payload": sp.value, "sum_amount_naive": sn.value, "sum_amount": sc.value,
            "count": cnt.value}
