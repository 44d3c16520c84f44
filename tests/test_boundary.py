"""CPU tests of the drop-in boundary: the C-ABI library loads, exports every
symbol include/dense_hodlr.h declares, and reports errors as status codes."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from dense_hodlr_amd import _lib
from dense_hodlr_amd import inputs as I

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "dense_hodlr.h")).read()
    return sorted(set(re.findall(r"\b(hodlr_[a-z0-9_]+)\s*\(", txt)))


def test_header_and_binding_agree():
    assert _declared() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    for name in _declared():
        assert hasattr(L, name), name
    assert b"gfx950" in L.hodlr_version()


def _has_gpu():
    """A usable device, asked of the library itself (no torch: its own copy
    of the HIP runtime would then be the one the library binds to)."""
    try:
        return _lib.lib().hodlr_memory_get_stats(C.byref(_lib.MemoryStats())) == 0
    except Exception:
        return False


@pytest.mark.skipif(_has_gpu(), reason="checks the no-device error path")
def test_create_without_device_returns_status():
    L = _lib.lib()
    pts = I.points_1d_sorted(64, 1)
    kd = _lib.KernelDesc(0, 1, pts.ctypes.data_as(C.POINTER(C.c_double)), 0.1, 1e-2, 0)
    h = C.c_void_p()
    rc = L.hodlr_create_kernel(C.byref(h), 64, C.byref(kd), 16)
    assert rc == -3 and h.value is None
    assert b"device" in L.hodlr_last_error(None)


def test_argument_errors_are_status_codes():
    L = _lib.lib()
    h = C.c_void_p()
    assert L.hodlr_create_kernel(C.byref(h), 0, None, 16) == -1
    assert L.hodlr_create_dense(C.byref(h), 4, None, 4, 2, 0) == -1
    assert L.hodlr_set_tolerance(None, 1e-6) == -1
    assert L.hodlr_destroy(None) == -1


def test_batch_argument_errors_are_status_codes():
    L = _lib.lib()
    h = C.c_void_p()
    pts = I.points_1d_sorted(64, 1)
    kd = (_lib.KernelDesc * 2)()
    kd[0] = _lib.KernelDesc(2, 1, pts.ctypes.data_as(C.POINTER(C.c_double)), 0.1, 1e-2, 0)
    kd[1] = _lib.KernelDesc(2, 1, pts.ctypes.data_as(C.POINTER(C.c_double)), -1.0, 1e-2, 0)
    ns = (C.c_int * 2)(64, 64)
    assert L.hodlr_create_batch(C.byref(h), 0, ns, kd, 16) == -1
    rc = L.hodlr_create_batch(C.byref(h), 2, ns, kd, 16)  # instance 1: bad length scale
    assert rc == -1 and h.value is None
    assert b"instance 1" in L.hodlr_last_error(None)
    assert L.hodlr_batch_count(None) == -1
    assert L.hodlr_logabsdet_batch(None, None) == -1


def test_batch_inputs_are_reproducible():
    a = I.batch_inputs(6, 1024, 16384, seed=9)
    b = I.batch_inputs(6, 1024, 16384, seed=9)
    assert a["n"] == b["n"] and a["length_scale"] == b["length_scale"]
    assert all(1024 <= n <= 16384 for n in a["n"])
    assert all(0.01 <= l <= 0.2 for l in a["length_scale"])
    for p, q in zip(a["points"], b["points"]):
        assert np.array_equal(p, q) and np.all(np.diff(p.ravel()) >= 0)
