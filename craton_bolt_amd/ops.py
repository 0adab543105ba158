"""Host-side operator layer over the C ABI (include/bolt_ops.h).

Function names and argument meaning follow the reference's executors
(src/exec/*.rs): execute_aggregate, compact_columns, execute_groupby,
execute_inner_join, execute_sort, plus the join build / probe pair of
src/exec/gpu_join.rs.  Inputs and outputs are device columns; results that
the reference returns as scalars are returned as Python values, with SQL NULL
as None.  Errors surface as the typed exceptions of `_native` (GpuCapacity is
the decline the caller answers by running its host path).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .device import DeviceBuffer, DeviceColumn, Stream, columns_array


# ---- predicates -------------------------------------------------------------

@dataclass
class Term:
    """`col <op> literal` (or IS [NOT] NULL) over the operator's input columns."""

    col: int
    op: int
    value: float | int = 0

    def c(self, dtype: int) -> N.BoltPredTerm:
        t = N.BoltPredTerm(self.col, self.op, 0, 0.0, 0)
        if dtype in (N.F32, N.F64):
            t.lit_f64 = float(self.value)
        elif dtype == N.DEC128:
            v = int(self.value)
            t.lit_i64 = C.c_int64(v & 0xFFFFFFFFFFFFFFFF).value
            t.lit_hi = C.c_int64((v >> 64) & 0xFFFFFFFFFFFFFFFF).value
        else:
            t.lit_i64 = int(self.value)
        return t


def _filter(terms: list[Term] | None, cols: list[DeviceColumn]):
    if not terms:
        return None, None
    arr = (N.BoltPredTerm * len(terms))()
    for i, t in enumerate(terms):
        arr[i] = t.c(cols[t.col].dtype)
    f = N.BoltFilter(arr, len(terms))
    return f, arr  # keep `arr` alive


def _stream(s: Stream | None):
    return s.handle if s is not None else None


def scalar_value(s: N.BoltScalar):
    """Python value of a bolt_scalar_t (None for SQL NULL)."""
    if s.is_null:
        return None
    if s.dtype in (N.F32, N.F64):
        return float(s.f64)
    if s.dtype == N.DEC128:
        return (int(s.hi) << 64) | (int(s.i64) & 0xFFFFFFFFFFFFFFFF)
    return int(s.i64)


def _aggs(aggs: list[tuple[int, int]]):
    arr = (N.BoltAgg * max(len(aggs), 1))()
    for i, (op, col) in enumerate(aggs):
        arr[i] = N.BoltAgg(op, col)
    return arr


# ---- scalar aggregates (src/exec/aggregate.rs:154) ----------------------------

def execute_aggregate(cols: list[DeviceColumn], aggs: list[tuple[int, int]], terms: list[Term] | None = None,
                      stream: Stream | None = None) -> list:
    """SELECT agg(col), ... FROM cols [WHERE terms]; one value per aggregate."""
    ca = columns_array(cols)
    f, keep = _filter(terms, cols)
    res = (N.BoltScalar * max(len(aggs), 1))()
    st = N.lib().bolt_aggregate(_stream(stream), ca, len(cols), _aggs(aggs), len(aggs),
                                C.byref(f) if f else None, res)
    N.check(st, "bolt_aggregate")
    return [scalar_value(res[i]) for i in range(len(aggs))]


# ---- filter + compaction (src/exec/gpu_compact.rs:1005) -----------------------

def compact_columns(cols: list[DeviceColumn], terms: list[Term] | None = None, mask: DeviceBuffer | None = None,
                    stream: Stream | None = None) -> tuple[list[DeviceColumn], int]:
    """Rows passing `terms` (and `mask`, a u8 per row) gathered in input order."""
    n = cols[0].length
    outs = [DeviceColumn.empty(n, c.dtype, with_validity=c.validity is not None) for c in cols]
    ca, oa = columns_array(cols), columns_array(outs)
    f, keep = _filter(terms, cols)
    n_out = C.c_int64()
    st = N.lib().bolt_filter_compact(_stream(stream), ca, len(cols), C.byref(f) if f else None,
                                     mask.ptr if mask else None, oa, C.byref(n_out))
    N.check(st, "bolt_filter_compact")
    for o in outs:
        o.length = n_out.value
    return outs, n_out.value


def prefix_scan_mask(mask: DeviceBuffer, n: int, stream: Stream | None = None) -> tuple[DeviceBuffer, int]:
    out = DeviceBuffer(max(n, 1) * 4)
    total = C.c_int64()
    N.check(N.lib().bolt_prefix_scan_mask(_stream(stream), mask.ptr, n, out.ptr, C.byref(total)),
            "bolt_prefix_scan_mask")
    return out, total.value


# ---- GROUP BY (src/exec/groupby.rs:295) -----------------------------------------

def execute_groupby(keys: list[DeviceColumn], inputs: list[DeviceColumn], aggs: list[tuple[int, int]],
                    terms: list[Term] | None = None, sorted_output: bool = True, flags: int = 0,
                    stream: Stream | None = None) -> dict:
    """Returns {'n_groups', 'keys': [(values, valid)], 'aggs': [(values, valid)]} as numpy arrays."""
    ka, ia = columns_array(keys), columns_array(inputs)
    f, keep = _filter(terms, inputs)
    tbl = C.c_void_p()
    fl = flags | (N.GROUPBY_SORTED if sorted_output else 0)
    st = N.lib().bolt_groupby(_stream(stream), ka, len(keys), ia, len(inputs), _aggs(aggs), len(aggs),
                              C.byref(f) if f else None, fl, C.byref(tbl))
    N.check(st, "bolt_groupby")
    return _table_result(tbl, len(keys))


def groupby_last_plan() -> dict:
    """How the calling thread's last GROUP BY ran (bolt_groupby_last_plan)."""
    tier, spilled, partials, parts = C.c_int32(), C.c_int64(), C.c_int64(), C.c_int32()
    N.check(N.lib().bolt_groupby_last_plan(C.byref(tier), C.byref(spilled), C.byref(partials), C.byref(parts)))
    return {"tier": tier.value, "spilled_rows": spilled.value, "partial_records": partials.value,
            "partitions": parts.value}


def autotune_groupby_partitions(keys: list[DeviceColumn], inputs: list[DeviceColumn], aggs: list[tuple[int, int]],
                                reps: int = 2, stream: Stream | None = None) -> tuple[int, float]:
    """Times the partitioned tier's fan-out candidates on these inputs and keeps the fastest for inputs of
    this shape (bolt_groupby_autotune_partitions); returns (partitions, ms per call)."""
    parts, ms = C.c_int32(), C.c_float()
    st = N.lib().bolt_groupby_autotune_partitions(_stream(stream), columns_array(keys), len(keys), columns_array(inputs),
                                                  len(inputs), _aggs(aggs), len(aggs), reps, C.byref(parts),
                                                  C.byref(ms))
    N.check(st, "bolt_groupby_autotune_partitions")
    return parts.value, float(ms.value)


def _table_result(tbl: C.c_void_p, n_keys: int) -> dict:
    """Downloads and releases a GROUP BY result table."""
    try:
        nr = C.c_int64()
        nc = C.c_int32()
        N.check(N.lib().bolt_table_num_rows(tbl, C.byref(nr)))
        N.check(N.lib().bolt_table_num_columns(tbl, C.byref(nc)))
        cols = []
        for i in range(nc.value):
            bc = N.BoltColumn()
            N.check(N.lib().bolt_table_column(tbl, i, C.byref(bc)))
            cols.append(_download(bc, nr.value))
    finally:
        N.lib().bolt_table_release(tbl)
    return {"n_groups": nr.value, "keys": cols[:n_keys], "aggs": cols[n_keys:]}


def _download(bc: N.BoltColumn, n: int):
    from .device import _NP_DTYPE

    if bc.dtype == N.DEC128:
        vals = np.empty((n, 2), dtype=np.int64)
    else:
        vals = np.empty(n, dtype=_NP_DTYPE[bc.dtype])
    if n:
        N.check_rt(N.lib().bolt_hip_memcpy_dtoh(vals.ctypes.data, bc.data, vals.nbytes), "dtoh")
    valid = None
    if bc.validity:
        bits = np.empty((n + 7) // 8, dtype=np.uint8)
        if n:
            N.check_rt(N.lib().bolt_hip_memcpy_dtoh(bits.ctypes.data, bc.validity, bits.nbytes), "dtoh")
        valid = np.unpackbits(bits, bitorder="little")[:n].astype(bool)
    return vals, valid


# ---- batched GROUP BY (bolt_batch.h) ----------------------------------------------------

class GroupByBatch:
    """Many GROUP BY queries with one aggregate list, run as one graph-captured
    persistent pipeline.  queries: list of (key_columns, input_columns)."""

    def __init__(self, queries: list[tuple[list[DeviceColumn], list[DeviceColumn]]], aggs: list[tuple[int, int]],
                 tune: dict | None = None, stream: Stream | None = None):
        self._h = C.c_void_p()
        self.n_queries = len(queries)
        self.n_keys = len(queries[0][0])
        self.n_inputs = len(queries[0][1])
        self.n_aggs = len(aggs)
        self._cols = queries  # keep the device columns alive
        self._stream = stream  # and the stream: the batch runs on it and frees its buffers on it
        ka = (N.BoltColumn * (self.n_queries * self.n_keys))()
        ia = (N.BoltColumn * max(1, self.n_queries * self.n_inputs))()
        for q, (ks, ins) in enumerate(queries):
            if len(ks) != self.n_keys or len(ins) != self.n_inputs:
                raise ValueError("every query needs the same number of key and input columns")
            for j, c in enumerate(ks):
                ka[q * self.n_keys + j] = c.c()
            for j, c in enumerate(ins):
                ia[q * self.n_inputs + j] = c.c()
        t = None
        if tune:
            t = N.BoltGbTune(tune.get("block", 0), tune.get("tile_rows", 0), tune.get("lds_slots", 0),
                             tune.get("grid", 0))
        st = N.lib().bolt_gb_batch_create(_stream(stream), ka, self.n_keys, ia, self.n_inputs, self.n_queries,
                                          _aggs(aggs), len(aggs), C.byref(t) if t else None, C.byref(self._h))
        N.check(st, "bolt_gb_batch_create")

    def run(self) -> None:
        N.check(N.lib().bolt_gb_batch_run(self._h), "bolt_gb_batch_run")

    def num_groups(self, q: int) -> int:
        n = C.c_int64()
        N.check(N.lib().bolt_gb_batch_num_groups(self._h, q, C.byref(n)))
        return n.value

    def result(self, q: int) -> dict:
        """Query q's groups (hash order), in execute_groupby's dict layout."""
        n = self.num_groups(q)
        cols = []
        for i in range(self.n_keys + self.n_aggs):
            bc = N.BoltColumn()
            N.check(N.lib().bolt_gb_batch_column(self._h, q, i, C.byref(bc)))
            cols.append(_download(bc, n))
        return {"n_groups": n, "keys": cols[: self.n_keys], "aggs": cols[self.n_keys:]}

    def tune(self) -> tuple[dict, int]:
        t = N.BoltGbTune()
        nb = C.c_int64()
        N.check(N.lib().bolt_gb_batch_get_tune(self._h, C.byref(t), C.byref(nb)))
        return {"block": t.block, "tile_rows": t.tile_rows, "lds_slots": t.lds_slots, "grid": t.grid}, nb.value

    def autotune(self, reps: int = 2) -> tuple[dict, float]:
        t = N.BoltGbTune()
        ms = C.c_float()
        N.check(N.lib().bolt_gb_batch_autotune(self._h, reps, C.byref(t), C.byref(ms)), "bolt_gb_batch_autotune")
        return {"block": t.block, "tile_rows": t.tile_rows, "lds_slots": t.lds_slots, "grid": t.grid}, ms.value

    def release(self) -> None:
        if self._h:
            N.lib().bolt_gb_batch_release(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


# ---- served form (bolt_serve.h) ------------------------------------------------------

class Server:
    """Request loop over a device-resident star schema (continuous batching,
    HBM-budget admission).  dims: list of (key column, [attribute columns])."""

    def __init__(self, fact: list[DeviceColumn], dims: list[tuple[DeviceColumn, list[DeviceColumn]]],
                 slots: int = 8, hbm_budget: int = 4 << 30, chunk_rows: int = 0):
        self._h = C.c_void_p()
        self._keep = (fact, dims)
        fa = columns_array(fact)
        self._attr_arrays = [columns_array(attrs) for _, attrs in dims]
        da = (N.BoltServeDim * len(dims))()
        for i, (key, attrs) in enumerate(dims):
            da[i] = N.BoltServeDim(key.c(), C.cast(self._attr_arrays[i], C.POINTER(N.BoltColumn)), len(attrs))
        cfg = N.BoltServeConfig(slots, hbm_budget, chunk_rows)
        N.check(N.lib().bolt_server_create(None, fa, len(fact), da, len(dims), C.byref(cfg), C.byref(self._h)),
                "bolt_server_create")

    def query(self, terms: list[Term], joins: list[tuple[int, int]], group: list[tuple[int, int]],
              aggs: list[tuple[int, int]], order_agg: int = 0, limit: int = 100, est_groups: int = 0):
        """joins: (fact fk column, dimension); group: (join index, attribute index)."""
        q = N.BoltServeQuery()
        q.n_terms = len(terms)
        for i, t in enumerate(terms):
            q.terms[i] = t.c(self._keep[0][t.col].dtype)
        q.n_joins = len(joins)
        for i, (fk, dim) in enumerate(joins):
            q.join_fk[i] = fk
            q.join_dim[i] = dim
        q.n_group = len(group)
        for i, (j, a) in enumerate(group):
            q.group_dim[i] = j
            q.group_attr[i] = a
        q.n_aggs = len(aggs)
        for i, (op, col) in enumerate(aggs):
            q.aggs[i] = N.BoltAgg(op, col)
        q.order_agg = order_agg
        q.limit = limit
        q.est_groups = est_groups
        return q

    def submit(self, q: N.BoltServeQuery) -> int:
        t = C.c_int64()
        N.check(N.lib().bolt_server_submit(self._h, C.byref(q), C.byref(t)), "bolt_server_submit")
        return t.value

    def wait(self, ticket: int, timeout_ms: float = -1.0) -> dict | None:
        r = N.BoltServeResult()
        st = N.lib().bolt_server_wait(self._h, ticket, timeout_ms, C.byref(r))
        if st == N.BOLT_HIP_ERROR_NOT_READY:
            return None
        N.check(st, "bolt_server_wait")
        rows = []
        for i in range(r.n_rows):
            rows.append(tuple(scalar_value(r.cells[i * r.n_cols + c]) for c in range(r.n_cols)))
        out = {"status": r.status, "rows": rows, "n_groups": r.n_groups, "hbm_bytes": r.hbm_bytes,
               "submit_ms": r.submit_ms,
               "admit_ms": r.admit_ms, "done_ms": r.done_ms,
               "error": r.error.decode(errors="replace") if r.error else ""}
        N.lib().bolt_server_result_release(self._h, ticket)
        return out

    def now_ms(self) -> float:
        return N.lib().bolt_server_now_ms(self._h)

    def stats(self) -> dict:
        st, done, pa, pb = C.c_int64(), C.c_int64(), C.c_int32(), C.c_int64()
        sm, fm = C.c_double(), C.c_double()
        N.check(N.lib().bolt_server_stats(self._h, C.byref(st), C.byref(done), C.byref(pa), C.byref(pb),
                                          C.byref(sm), C.byref(fm)))
        return {"steps": st.value, "completed": done.value, "peak_active": pa.value, "peak_bytes": pb.value,
                "step_ms": sm.value, "finish_ms": fm.value}

    def close(self) -> None:
        if self._h:
            N.lib().bolt_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def take(col: DeviceColumn, rows: DeviceBuffer, n: int, stream: Stream | None = None,
         nullable: bool = False) -> DeviceColumn:
    """col[rows[:n]] on the device (bolt_take); a NULL_ROW id gathers a NULL (give `nullable=True`)."""
    out = DeviceColumn.empty(n, col.dtype, with_validity=col.validity is not None or nullable)
    out.scale = col.scale
    oc = out.c()
    N.check(N.lib().bolt_take(_stream(stream), C.byref(col.c()), rows.ptr, n, C.byref(oc)), "bolt_take")
    return out


def execute_star_query(fact: list[DeviceColumn], dims: list[tuple[DeviceColumn, list[DeviceColumn]]],
                       terms: list[Term], joins: list[tuple[int, int]], group: list[tuple[int, int]],
                       aggs: list[tuple[int, int]], order_agg: int = 0, limit: int = 100,
                       stream: Stream | None = None, est_groups: int = 0) -> dict:
    """One star query composed from the operators, one at a time, the way the
    reference's engine runs it (src/exec/engine.rs:2520 sql -> physical plan):
    WHERE as a count plus filter-compaction of the referenced fact columns,
    then per join a hash build on the dimension key, a probe emitting
    (probe row, build row) pairs and a take of every carried column,
    GROUP BY, ORDER BY the chosen aggregate DESC then the group key, LIMIT.
    Aggregate-over-join is not fused (src/plan/physical_plan.rs:3691-3712).
    Same query description and result rows as Server.query / Server.wait."""
    del est_groups  # sizing hint of the served form; the operators size themselves
    L = N.lib()
    s = stream
    fk_cols = [fk for fk, _ in joins]
    agg_cols = [c for op, c in aggs if op != N.COUNT_STAR]
    used = sorted(set(fk_cols + agg_cols + [t.col for t in terms]))
    pos = {c: i for i, c in enumerate(used)}
    cols = [fact[c] for c in used]
    rterms = [Term(pos[t.col], t.op, t.value) for t in terms]
    # WHERE: survivor count, then one compaction pass of the referenced columns
    n_pass = execute_aggregate(cols, [(N.COUNT_STAR, 0)], rterms, stream=s)[0]
    outs = [DeviceColumn.empty(max(n_pass, 1), c.dtype, with_validity=c.validity is not None) for c in cols]
    ca, oa = columns_array(cols), columns_array(outs)
    f, keep = _filter(rterms, cols)
    n_out = C.c_int64()
    N.check(L.bolt_filter_compact(_stream(s), ca, len(cols), C.byref(f) if f else None, None, oa, C.byref(n_out)),
            "bolt_filter_compact")
    n_cur = n_out.value
    cur = {c: outs[pos[c]] for c in set(fk_cols + agg_cols)}
    for c in cur.values():
        c.length = n_cur
    got = {}  # (join index, attribute) -> gathered dimension attribute
    for j, (fk, d) in enumerate(joins):
        dkey, dattrs = dims[d]
        t = JoinTable(dkey, stream=s)
        pr, br = DeviceBuffer(max(n_cur, 1) * 4), DeviceBuffer(max(n_cur, 1) * 4)
        probe = [cur[fk]]
        n_pairs = C.c_int64()
        st = L.bolt_join_probe(_stream(s), t.handle, columns_array(probe), 1, 0, None, pr.ptr, br.ptr,
                               max(n_cur, 1), C.byref(n_pairs))
        if st == N.BOLT_DECLINE:  # duplicate dimension keys: room for every pair
            cap = n_pairs.value
            pr, br = DeviceBuffer(max(cap, 1) * 4), DeviceBuffer(max(cap, 1) * 4)
            st = L.bolt_join_probe(_stream(s), t.handle, columns_array(probe), 1, 0, None, pr.ptr, br.ptr,
                                   max(cap, 1), C.byref(n_pairs))
        N.check(st, "bolt_join_probe")
        t.release()
        m = n_pairs.value
        cur = {c: take(col, pr, m, s) for c, col in cur.items()}
        got = {k: take(col, pr, m, s) for k, col in got.items()}
        for jj, a in group:
            if jj == j:
                got[(j, a)] = take(dattrs[a], br, m, s)
        n_cur = m
    # GROUP BY the dimension attributes
    keys = [got[(j, a)] for j, a in group]
    in_pos = {c: i for i, c in enumerate(sorted(set(agg_cols)))}
    inputs = [cur[c] for c in sorted(set(agg_cols))] or [keys[0]]
    gaggs = [(op, in_pos.get(c, 0)) for op, c in aggs]
    tbl = C.c_void_p()
    N.check(L.bolt_groupby(_stream(s), columns_array(keys), len(keys), columns_array(inputs), len(inputs),
                           _aggs(gaggs), len(gaggs), None, 0, C.byref(tbl)), "bolt_groupby")
    try:
        nr, nc = C.c_int64(), C.c_int32()
        N.check(L.bolt_table_num_rows(tbl, C.byref(nr)))
        N.check(L.bolt_table_num_columns(tbl, C.byref(nc)))
        g = nr.value
        tcols = []
        for i in range(nc.value):
            bc = N.BoltColumn()
            N.check(L.bolt_table_column(tbl, i, C.byref(bc)))
            tcols.append(bc)
        rows = []
        lim = min(limit, g)
        if lim > 0:
            # ORDER BY aggregate DESC NULLS LAST, then the group key ascending
            oc = len(keys) + order_agg
            sk = [tcols[oc]] + tcols[: len(keys)]
            k = len(sk)
            arr = (N.BoltColumn * k)(*sk)
            desc = (C.c_int32 * k)(*([1] + [0] * len(keys)))
            nf = (C.c_int32 * k)(*([0] * k))
            perm = DeviceBuffer(max(g, 1) * 4)
            N.check(L.bolt_sort_indices(_stream(s), arr, k, desc, nf, perm.ptr), "bolt_sort_indices")
            cols_out = []
            for bc in tcols:
                src = DeviceColumn(_Borrowed(bc.data), g, bc.dtype, _Borrowed(bc.validity) if bc.validity else None,
                                   bc.scale)
                cols_out.append(take(src, perm, lim, s))
            vals = [c.to_numpy(lim) for c in cols_out]
            valid = [c.valid_numpy(lim) for c in cols_out]
            for r in range(lim):
                row = []
                for c, bc in enumerate(tcols):
                    if valid[c] is not None and not valid[c][r]:
                        row.append(None)
                    elif bc.dtype == N.DEC128:
                        row.append((int(vals[c][r][1]) << 64) | (int(vals[c][r][0]) & 0xFFFFFFFFFFFFFFFF))
                    elif bc.dtype in (N.F32, N.F64):
                        row.append(float(vals[c][r]))
                    else:
                        row.append(int(vals[c][r]))
                rows.append(tuple(row))
    finally:
        L.bolt_table_release(tbl)
    return {"status": 0, "rows": rows, "n_groups": g}


class _Borrowed:
    """A device pointer owned elsewhere (a result table's column)."""

    def __init__(self, ptr):
        self.ptr = ptr
        self.nbytes = 0

    def free(self) -> None:
        pass


# ---- hash join (src/exec/gpu_join.rs:1583, src/exec/join.rs:169) --------------------

class JoinTable:
    """A built hash table (build side resident on the device)."""

    def __init__(self, key: DeviceColumn, carry: DeviceColumn | None = None, stream: Stream | None = None):
        self._h = C.c_void_p()
        self.stream = stream
        kc = key.c()
        cc = carry.c() if carry is not None else None
        st = N.lib().bolt_join_build(_stream(stream), C.byref(kc), C.byref(cc) if cc is not None else None,
                                     C.byref(self._h))
        N.check(st, "bolt_join_build")

    @property
    def handle(self):
        return self._h

    def unique(self) -> bool:
        u = C.c_int32()
        N.check(N.lib().bolt_join_table_unique(self._h, C.byref(u)), "bolt_join_table_unique")
        return bool(u.value)

    def probe(self, probe_cols: list[DeviceColumn], key_col: int, terms: list[Term] | None = None,
              capacity: int | None = None) -> tuple[np.ndarray, np.ndarray]:
        n = probe_cols[key_col].length
        cap = n if capacity is None else capacity
        op_, ob_ = DeviceBuffer(max(cap, 1) * 4), DeviceBuffer(max(cap, 1) * 4)
        ca = columns_array(probe_cols)
        f, keep = _filter(terms, probe_cols)
        n_out = C.c_int64()
        st = N.lib().bolt_join_probe(_stream(self.stream), self._h, ca, len(probe_cols), key_col,
                                     C.byref(f) if f else None, op_.ptr, ob_.ptr, cap, C.byref(n_out))
        N.check(st, "bolt_join_probe")
        k = n_out.value
        return op_.to_numpy(np.uint32, k), ob_.to_numpy(np.uint32, k)

    def probe_aggregate(self, probe_cols: list[DeviceColumn], key_col: int, aggs: list[tuple[int, int]],
                        terms: list[Term] | None = None) -> list:
        ca = columns_array(probe_cols)
        f, keep = _filter(terms, probe_cols)
        res = (N.BoltScalar * max(len(aggs), 1))()
        st = N.lib().bolt_join_probe_aggregate(_stream(self.stream), self._h, ca, len(probe_cols), key_col,
                                               C.byref(f) if f else None, _aggs(aggs), len(aggs), res)
        N.check(st, "bolt_join_probe_aggregate")
        return [scalar_value(res[i]) for i in range(len(aggs))]

    def probe_aggregate_expr(self, inputs: list[DeviceColumn], program: "Program", n_values: int, key_value: int,
                             aggs: list[tuple[int, int]], filter_value: int = -1,
                             value_scales: list[int] | None = None) -> list:
        """probe_aggregate with the probe key, predicate and aggregate inputs computed by `program` over
        `inputs` (bolt_join_probe_aggregate_expr)."""
        res = (N.BoltScalar * max(len(aggs), 1))()
        st = N.lib().bolt_join_probe_aggregate_expr(_stream(self.stream), self._h, columns_array(inputs), len(inputs),
                                                    program.array(N.BoltExprOp), len(program.ops), n_values,
                                                    _i32_array(value_scales), key_value, filter_value, _aggs(aggs),
                                                    len(aggs), res)
        N.check(st, "bolt_join_probe_aggregate_expr")
        return [scalar_value(res[i]) for i in range(len(aggs))]

    def release(self) -> None:
        if self._h:
            N.lib().bolt_join_table_release(self._h)
            self._h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def execute_inner_join(left_key: DeviceColumn, right_key: DeviceColumn, capacity: int | None = None,
                       stream: Stream | None = None) -> tuple[np.ndarray, np.ndarray]:
    """(left_row, right_row) pairs of the INNER equi-join (unordered)."""
    cap = capacity if capacity is not None else max(left_key.length, right_key.length)
    while True:
        ol, orr = DeviceBuffer(max(cap, 1) * 4), DeviceBuffer(max(cap, 1) * 4)
        lk, rk = left_key.c(), right_key.c()
        n_out = C.c_int64()
        st = N.lib().bolt_inner_join(_stream(stream), C.byref(lk), C.byref(rk), ol.ptr, orr.ptr, cap,
                                     C.byref(n_out))
        if st == N.BOLT_DECLINE and capacity is None and n_out.value > cap:
            cap = n_out.value
            continue
        N.check(st, "bolt_inner_join")
        k = n_out.value
        return ol.to_numpy(np.uint32, k), orr.to_numpy(np.uint32, k)


# ---- per-row expression programs (src/plan/physical_plan.rs:81 Op) ---------------------------

X_LOAD, X_CONST, X_CAST, X_BINARY, X_SELECT, X_NOT, X_IS_NULL, X_STORE, X_EXTRACT, X_DATE_TRUNC = range(1, 11)
# EXTRACT fields / DATE_TRUNC units (bolt_date_field)
DATE_FIELDS = {"year": 1, "month": 2, "day": 3, "hour": 4, "minute": 5, "second": 6}
BIN_OPS = {"+": 1, "-": 2, "*": 3, "/": 4, "%": 5, "==": 6, "!=": 7, "<": 8, "<=": 9, ">": 10, ">=": 11,
           "and": 12, "or": 13, "&": 14, "|": 15}


class Program:
    """Builder of a per-row expression program for bolt_project: each method
    appends one op and returns the register it writes (registers in order)."""

    def __init__(self):
        self.ops: list[dict] = []
        self._next = 0

    def _emit(self, kind: int, **kw) -> int:
        op = dict(kind=kind, dst=0, a=0, b=0, c=0, col=0, op=0, dtype=0, lit_i64=0, lit_f64=0.0, lit_hi=0)
        if kind != X_STORE:
            op["dst"] = self._next
            self._next += 1
        op.update(kw)
        self.ops.append(op)
        return op["dst"]

    def load(self, col: int, dtype: int) -> int:
        return self._emit(X_LOAD, col=col, dtype=dtype)

    def const(self, value, dtype: int, unit: int = 0) -> int:
        """A literal; a Timestamp literal carries its unit (digits of a second)."""
        if dtype == N.TIMESTAMP:
            return self._emit(X_CONST, dtype=dtype, lit_i64=int(value), op=unit)
        if dtype in (N.F32, N.F64):
            return self._emit(X_CONST, dtype=dtype, lit_f64=float(value))
        v = int(value)
        if dtype == N.DEC128:
            m = (1 << 64) - 1
            lo, hi = v & m, (v >> 64) & m
            return self._emit(X_CONST, dtype=dtype, lit_i64=lo - (1 << 64) if lo >> 63 else lo,
                              lit_hi=hi - (1 << 64) if hi >> 63 else hi)
        return self._emit(X_CONST, dtype=dtype, lit_i64=v)

    def cast(self, a: int, dtype: int, scale: int = 0, from_scale: int = 0) -> int:
        """(dtype) a; `scale` is the target scale of a Decimal128 target, `from_scale` the
        scale of a Decimal128 source."""
        return self._emit(X_CAST, a=a, dtype=dtype, op=scale, c=from_scale)

    def extract(self, field: str, a: int) -> int:
        """EXTRACT(field FROM a), Int64; a Date32 or Timestamp."""
        return self._emit(X_EXTRACT, a=a, op=DATE_FIELDS[field])

    def date_trunc(self, unit: str, a: int) -> int:
        """DATE_TRUNC(unit, a), of a's dtype."""
        return self._emit(X_DATE_TRUNC, a=a, op=DATE_FIELDS[unit])

    def binary(self, op: str, a: int, b: int, dtype: int) -> int:
        return self._emit(X_BINARY, a=a, b=b, op=BIN_OPS[op], dtype=dtype)

    def select(self, cond: int, then: int, other: int) -> int:
        return self._emit(X_SELECT, a=cond, b=then, c=other)

    def not_(self, a: int) -> int:
        return self._emit(X_NOT, a=a)

    def is_null(self, col: int, want_null: bool = True) -> int:
        return self._emit(X_IS_NULL, col=col, op=int(want_null))

    def store(self, a: int, col: int) -> None:
        self._emit(X_STORE, a=a, col=col)

    def array(self, struct):
        arr = (struct * len(self.ops))()
        for i, op in enumerate(self.ops):
            for k, v in op.items():
                setattr(arr[i], k, v)
        return arr


def execute_projection(inputs: list[DeviceColumn], program: Program, out_dtypes: list[int],
                       out_validity: list[bool] | None = None, stream: Stream | None = None,
                       out_scales: list[int] | None = None) -> list[DeviceColumn]:
    """Runs `program` over every row of `inputs` (bolt_project); returns the output columns
    (`out_scales`: Decimal128 scales / Timestamp units of the outputs)."""
    n = inputs[0].length if inputs else 0
    outs = [DeviceColumn.empty(n, dt, with_validity=bool(out_validity and out_validity[i]),
                               scale=out_scales[i] if out_scales else 0)
            for i, dt in enumerate(out_dtypes)]
    oc = columns_array(outs)
    st = N.lib().bolt_project(_stream(stream), columns_array(inputs), len(inputs), program.array(N.BoltExprOp),
                              len(program.ops), oc, len(outs))
    N.check(st, "bolt_project")
    return outs


def execute_aggregate_expr(inputs: list[DeviceColumn], program: Program, n_values: int, aggs: list[tuple[int, int]],
                           filter_value: int = -1, stream: Stream | None = None) -> list:
    """Aggregates over the values of `program` for the rows where value `filter_value` is true
    (bolt_aggregate_expr: one fused kernel, nothing materialised)."""
    res = (N.BoltScalar * max(len(aggs), 1))()
    st = N.lib().bolt_aggregate_expr(_stream(stream), columns_array(inputs), len(inputs), program.array(N.BoltExprOp),
                                     len(program.ops), n_values, filter_value, _aggs(aggs), len(aggs), res)
    N.check(st, "bolt_aggregate_expr")
    return [scalar_value(res[i]) for i in range(len(aggs))]


def _i32_array(xs):
    return (C.c_int32 * len(xs))(*xs) if xs else None


def execute_groupby_expr(inputs: list[DeviceColumn], program: Program, n_values: int, key_values: list[int],
                         aggs: list[tuple[int, int]], filter_value: int = -1, value_scales: list[int] | None = None,
                         sorted_output: bool = True, flags: int = 0, stream: Stream | None = None) -> dict:
    """GROUP BY the program values `key_values`, aggregating values (each aggregate's col is a value index),
    over the rows where value `filter_value` is true (bolt_groupby_expr).  Same result as execute_groupby."""
    tbl = C.c_void_p()
    fl = flags | (N.GROUPBY_SORTED if sorted_output else 0)
    st = N.lib().bolt_groupby_expr(_stream(stream), columns_array(inputs), len(inputs), program.array(N.BoltExprOp),
                                   len(program.ops), n_values, _i32_array(value_scales), filter_value,
                                   _i32_array(key_values), len(key_values), _aggs(aggs), len(aggs), fl, C.byref(tbl))
    N.check(st, "bolt_groupby_expr")
    return _table_result(tbl, len(key_values))


JOIN_TYPES = {"inner": 0, "left": 1, "right": 2, "full": 3}
NULL_ROW = 0xFFFFFFFF


def execute_join(left_keys: list[DeviceColumn], right_keys: list[DeviceColumn], how: str = "inner",
                 capacity: int | None = None, stream: Stream | None = None) -> tuple[np.ndarray, np.ndarray]:
    """(left_row, right_row) pairs of an equi-join (execute_join, src/exec/join.rs:132) of kind `how`
    (inner / left / right / full); NULL_ROW marks the padded side of an outer-join pair."""
    if how not in JOIN_TYPES:
        raise ValueError(f"unknown join type {how!r}")
    if len(left_keys) != len(right_keys):
        raise ValueError("left and right need the same number of key columns")
    nl, nr = left_keys[0].length, right_keys[0].length
    cap = capacity if capacity is not None else max(nl + nr, 1)
    la, ra = columns_array(left_keys), columns_array(right_keys)
    while True:
        ol, orr = DeviceBuffer(max(cap, 1) * 4), DeviceBuffer(max(cap, 1) * 4)
        n_out = C.c_int64()
        st = N.lib().bolt_join(_stream(stream), la, ra, len(left_keys), JOIN_TYPES[how], ol.ptr, orr.ptr, cap,
                               C.byref(n_out))
        if st == N.BOLT_DECLINE and capacity is None and n_out.value > cap:
            cap = n_out.value
            continue
        N.check(st, "bolt_join")
        k = n_out.value
        return ol.to_numpy(np.uint32, k), orr.to_numpy(np.uint32, k)


# ---- sort (src/exec/sort.rs:161, src/exec/gpu_sort.rs:944) -----------------------------

def sort_pairs_i64(keys: DeviceBuffer, vals: DeviceBuffer, n: int, descending: bool = False,
                   stream: Stream | None = None) -> tuple[DeviceBuffer, DeviceBuffer]:
    ko, vo = DeviceBuffer(max(n, 1) * 8), DeviceBuffer(max(n, 1) * 4)
    st = N.lib().bolt_sort_pairs_i64(_stream(stream), keys.ptr, vals.ptr, n, 1 if descending else 0, ko.ptr, vo.ptr)
    N.check(st, "bolt_sort_pairs_i64")
    return ko, vo


def execute_sort(keys: list[DeviceColumn], descending: list[bool] | None = None,
                 nulls_first: list[bool] | None = None, stream: Stream | None = None) -> np.ndarray:
    """Stable row permutation for ORDER BY keys (numpy u32)."""
    k = len(keys)
    desc = (C.c_int32 * k)(*[int(bool(x)) for x in (descending or [False] * k)])
    nf = (C.c_int32 * k)(*[int(bool(x)) for x in (nulls_first or [True] * k)])
    n = keys[0].length
    out = DeviceBuffer(max(n, 1) * 4)
    st = N.lib().bolt_sort_indices(_stream(stream), columns_array(keys), k, desc, nf, out.ptr)
    N.check(st, "bolt_sort_indices")
    return out.to_numpy(np.uint32, n)


def execute_sort_limit(keys: list[DeviceColumn], limit: int, offset: int = 0, descending: list[bool] | None = None,
                       nulls_first: list[bool] | None = None, stream: Stream | None = None) -> np.ndarray:
    """Row ids of ORDER BY keys LIMIT limit OFFSET offset (execute_sort + execute_limit,
    src/exec/sort.rs:161, src/exec/limit.rs:18): the slice of the stable permutation."""
    k = len(keys)
    desc = (C.c_int32 * k)(*[int(bool(x)) for x in (descending or [False] * k)])
    nf = (C.c_int32 * k)(*[int(bool(x)) for x in (nulls_first or [True] * k)])
    out = DeviceBuffer(max(limit, 1) * 4)
    n_out = C.c_int64()
    st = N.lib().bolt_sort_limit(_stream(stream), columns_array(keys), k, desc, nf, offset, limit, out.ptr,
                                 C.byref(n_out))
    N.check(st, "bolt_sort_limit")
    return out.to_numpy(np.uint32, n_out.value)
