"""GPU test of hodlr_matvec_device_noalloc, the allocation-free operator
product: through the C ABI on device buffers, against hodlr_matvec (the
host-pointer entry point) on the same operator and input, bit for bit; two
calls in a row on one handle (the second with fewer right-hand sides), then
a smaller operator; the scratch limit reported as HODLR_ERR_ARG and an
uncompressed handle as HODLR_ERR_STATE, before anything is launched."""
import sys

import numpy as np
import pytest

from dense_hodlr_amd import HODLR_Matrix, _lib
from dense_hodlr_amd import inputs as I

from hipmem import DeviceArray

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def noalloc(g, X):
    """Y = A_H X by the allocation-free entry point on device buffers."""
    n, nrhs = X.shape
    dX = DeviceArray(n, nrhs)
    dX.upload(X)
    dY = DeviceArray(n, nrhs, fill=np.nan)  # every entry must be written
    rc = _lib.lib().hodlr_matvec_device_noalloc(g._h, dX.ptr, n, nrhs, dY.ptr, n)
    Y = dY.download()
    dX.free()
    dY.free()
    return rc, Y


def test_noalloc_matvec_matches_matvec():
    _case("noalloc matvec: gauss2d n=16384, 8 then 3 right-hand sides")
    n = 16384
    pts = I.points_2d_kd(n, 81)
    g = HODLR_Matrix.from_kernel(pts, "gaussian", 0.05, 1e-4, sizeThreshold=256)
    g.set_LRTolerance(1e-10)
    X = I.rhs_normal(n, 8, 82)
    rc, _ = noalloc(g, X)
    assert rc == -2, rc  # HODLR_ERR_STATE: not compressed yet
    g.storeLRinTree()
    for nrhs in (8, 3):
        Xk = np.asfortranarray(X[:, :nrhs])
        rc, Y = noalloc(g, Xk)
        assert rc == 0, _lib.lib().hodlr_last_error(g._h)
        Yref = g.matvec(Xk)
        assert np.array_equal(Y, Yref), np.max(np.abs(Y - Yref))
    _case("noalloc matvec: smaller operator, 1D n=3000")
    h = HODLR_Matrix.from_kernel(I.points_1d_sorted(3000, 83), "gaussian", 0.1, 1e-2,
                                 sizeThreshold=64)
    h.set_LRTolerance(1e-10)
    h.storeLRinTree()
    Xs = I.rhs_uniform(3000, 2, 84)
    rc, Y = noalloc(h, Xs)
    assert rc == 0, _lib.lib().hodlr_last_error(h._h)
    assert np.array_equal(Y, h.matvec(Xs))
    _case("noalloc matvec: beyond the scratch limit")
    big = 8192
    dX = DeviceArray(n, big)
    dY = DeviceArray(n, big)
    rc = _lib.lib().hodlr_matvec_device_noalloc(g._h, dX.ptr, n, big, dY.ptr, n)
    dX.free()
    dY.free()
    msg = _lib.lib().hodlr_last_error(g._h).decode()
    print(msg, flush=True)
    assert rc == -1 and "maximum" in msg
    # the handle still works after the rejected call
    rc, Y = noalloc(g, X)
    assert rc == 0 and np.array_equal(Y, g.matvec(X))
