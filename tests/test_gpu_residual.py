"""GPU tests of the double-double residual R = B - A_H X (hodlr_residual).

The exact residual is computed with rational arithmetic (fractions.Fraction)
from the FP64 values the operator is made of (dense leaf blocks and the U, V
factors read back from the device), then rounded once; the GPU result must
agree to a few units in the last place even where FP64 evaluation loses
many digits to cancellation.
"""
import sys

import numpy as np
import pytest

from dense_hodlr_amd import HODLR_Matrix
from dense_hodlr_amd import inputs as I

from exact_residual import exact_residual

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def test_dd_residual_matches_exact():
    _case("dd residual exact n=192")
    n, leaf = 192, 24
    pts = np.sort(I.points_1d_sorted(n, 11).ravel())
    d = pts[:, None] - pts[None, :]
    A = np.exp(-(d / 0.1) ** 2) + 1e-7 * np.eye(n)
    g = HODLR_Matrix(A, sizeThreshold=leaf)
    g.set_LRTolerance(1e-12)
    B = I.rhs_normal(n, 2, 12)
    X = g.recSM_Solve(B)  # |X| >> |B|: the terms of each row cancel
    R = g.residual(B, X)
    Rx = exact_residual(g, A, B, X)
    R64 = B - g.matvec(X)
    err_dd = np.max(np.abs(R - Rx))
    err_64 = np.max(np.abs(R64 - Rx))
    scale = np.max(np.abs(Rx))
    print(f"|X|/|B| = {np.linalg.norm(X) / np.linalg.norm(B):.1e}; max |R| {scale:.2e}; "
          f"dd error {err_dd:.2e}, fp64 error {err_64:.2e}", flush=True)
    assert err_dd <= 4 * np.finfo(float).eps * scale + 1e-300
    assert err_64 > err_dd


def test_dd_residual_zero_and_linear():
    _case("dd residual identities n=4096")
    cfg = I.config_inputs("gauss1d_4096")
    g = HODLR_Matrix.from_kernel(cfg["points"], "gaussian", 0.1, 1e-2, sizeThreshold=64)
    g.set_LRTolerance(1e-10)
    X = I.rhs_normal(4096, 3, 5)
    Z = np.zeros_like(X)
    # B - A_H 0 = B exactly
    assert np.array_equal(g.residual(X, Z), X)
    # 0 - A_H X agrees with -matvec to FP64 rounding of the product
    Y = g.matvec(X)
    R = g.residual(Z, X)
    assert np.max(np.abs(R + Y)) <= 1e-12 * np.max(np.abs(Y))
