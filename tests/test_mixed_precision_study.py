"""Why the FP32 factorization cannot precondition the N=262144 Gaussian
system (condition number ~2e7): a CPU study on the same kernel and leaf size
at N=16384 (tests/mixed_study.py restates the mixed recSM in numpy on the
oracle's compressed operator).

With the diagonal shift 1e-4 (condition number ~1e6) the FP32 node products
give a preconditioner whose refinement step contracts by ~0.1 (the device's
mixed path converges there).  With the shift 1e-5 (condition number ~1e7,
the headline system's order) the same preconditioner amplifies the error by
~1e2-1e4, and so does FP32 at any single level below the root: no split
between FP64 upper levels and FP32 lower levels (or the reverse) keeps the
step a contraction, while the FP64 factorization contracts by 1e-8."""
import sys

import numpy as np
import pytest

from dense_hodlr_amd import inputs as I

sys.path.insert(0, __import__("os").path.dirname(__file__))
import mixed_study as M  # noqa: E402


def _operator(shift, n=16384):
    from oracle import pyoracle
    pts = I.points_2d_kd(n, 12345)
    o = pyoracle.OracleHODLR(n, leaf=256, points=pts, family="gaussian", length_scale=0.05,
                             diag=shift)
    o.set_LRTolerance(1e-10)
    return M.Hodlr(o, pts, "gaussian", 0.05, shift)


@pytest.fixture(scope="module")
def ops():
    return {s: _operator(s) for s in (1e-4, 1e-5)}


def test_fp32_products_contract_at_condition_1e6(ops):
    H = ops[1e-4]
    B = I.rhs_normal(H.n, 2, 7)
    r64 = M.contraction(H, M.factorize(H, lambda d: "f64"), B)
    r32 = M.contraction(H, M.factorize(H, lambda d: "f32", store32=True), B)
    print(f"shift 1e-4: FP64 {r64:.2e}, FP32 products + storage {r32:.2e}", flush=True)
    assert r64 < 1e-8
    assert r32 < 0.5


def test_fp32_products_at_any_level_below_the_root_diverge_at_condition_1e7(ops):
    H = ops[1e-5]
    B = I.rhs_normal(H.n, 2, 7)
    r64 = M.contraction(H, M.factorize(H, lambda d: "f64"), B)
    F32 = M.factorize(H, lambda d: "f32", store32=True)
    r32 = M.contraction(H, F32, B)
    single = [M.contraction(H, M.factorize(H, lambda d, l=l: "f32" if d == l else "f64"), B)
              for l in range(H.nlev)]
    res = M.fgmres(H, F32, B[:, 0], 8)
    print(f"shift 1e-5: FP64 {r64:.2e}, FP32 everywhere {r32:.2e}; FP32 at one level: "
          + ", ".join(f"{l}: {r:.2e}" for l, r in enumerate(single))
          + f"; FGMRES residuals with the FP32 preconditioner {['%.2e' % v for v in res]}",
          flush=True)
    assert r64 < 1e-6
    assert r32 > 10.0
    # every split that puts FP32 on some level below the root diverges
    assert all(r > 1.0 for r in single[1:])
    # and a Krylov method does not rescue it within a cycle of 8
    assert min(res) > 1e-2
