"""craton_bolt_amd — MI355X-native execution core for the GPU relational
operator path of Craton Bolt (hash GROUP BY, hash join, radix sort,
stream compaction, scalar aggregates).

The compute path is the HIP library `libbolt_hip.so` (C ABI in include/);
this package is its host-side binding.
"""
from ._native import (BoltError, BoltMemoryError, BoltTypeError, DeviceError, GpuCapacity, NoDeviceError,
                      PlanError, UnsupportedError)

__all__ = [
    "BoltError", "BoltMemoryError", "BoltTypeError", "DeviceError", "GpuCapacity", "NoDeviceError",
    "PlanError", "UnsupportedError",
]
__version__ = "0.1.0"
