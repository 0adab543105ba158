"""Helpers shared by the GPU parity tests: upload host arrays as device
columns and as oracle columns from the same numpy data."""
import numpy as np

import oracle as O
from craton_bolt_amd import _native as N
from craton_bolt_amd.device import DeviceColumn

_TO_ORACLE = {N.I32: O.I32, N.I64: O.I64, N.F32: O.F32, N.F64: O.F64, N.DEC128: O.DEC128, N.DATE32: O.DATE32,
              N.BOOL: O.BOOL, N.TIMESTAMP: O.TIMESTAMP}


def both(values: np.ndarray, dtype: int, valid=None, scale: int = 0):
    """(device column, oracle column) over the same data."""
    dev = DeviceColumn.from_numpy(values, dtype, valid, scale)
    host = O.HostCol(values if dtype != N.DEC128 else np.ascontiguousarray(values, np.int64).reshape(-1, 2),
                     _TO_ORACLE[dtype], valid, scale)
    return dev, host


def rel_close(a: float, b: float, rtol: float) -> bool:
    if a is None or b is None:
        return a is b
    if np.isnan(a) or np.isnan(b):
        return bool(np.isnan(a) and np.isnan(b))
    if np.isinf(a) or np.isinf(b):
        return a == b
    return abs(a - b) <= rtol * max(abs(b), 1e-300) or abs(a - b) <= 1e-300


def dec_array(vals):
    m = (1 << 64) - 1
    return np.array([[v & m, (v >> 64) & m] for v in vals], dtype=np.uint64).view(np.int64)
