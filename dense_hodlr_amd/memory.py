"""The library's cache of device allocations (hodlr_memory_* in
include/dense_hodlr.h): blocks released by handles are kept per device and
reused, so repeated handles of one shape stop calling hipMalloc/hipFree."""
from __future__ import annotations

import ctypes as C

from ._lib import DevicePeaks, MemoryStats, check, lib


def memory_stats() -> dict:
    """Counters of the current device: bytes in use and cached, cache hits
    and misses, host seconds spent in hipMalloc and hipFree."""
    s = MemoryStats()
    check(lib().hodlr_memory_get_stats(C.byref(s)))
    return {name: getattr(s, name) for name, _ in MemoryStats._fields_}


def memory_trim() -> None:
    """Free every cached block of the current device."""
    check(lib().hodlr_memory_trim())


def memory_cache(on: bool) -> None:
    """Turn the cache off (emptying it) or back on."""
    check(lib().hodlr_memory_cache(1 if on else 0))


def device_peaks() -> dict:
    """Sustained throughput of the current device, measured by probe kernels
    (hodlr_measure_device_peaks): HBM read and copy GB/s, FP64 matrix-core,
    FP64 vector and FP32 matrix-core TFLOP/s."""
    p = DevicePeaks()
    check(lib().hodlr_measure_device_peaks(C.byref(p)))
    return {name: getattr(p, name) for name, _ in DevicePeaks._fields_}
