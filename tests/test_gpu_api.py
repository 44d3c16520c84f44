"""GPU tests of reference API behaviour that the parity suite does not reach:
determinant() without the permutation sign, set_minPivot's truncation to an
int, the status of a refinement that ends above its target, and the
factorization time reported with the roots' LU on the side stream."""
import sys

import numpy as np
import pytest

from dense_hodlr_amd import HODLR_Matrix, HodlrNotConverged
from dense_hodlr_amd import inputs as I

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_determinant_matches_oracle(oracle):
    """determinant() (HODLR_Matrix.cpp:1681-1690) multiplies the diagonals
    of every leaf LU and node capacitance LU, without the sign of the row
    interchanges.  A well-scaled exact-HODLR matrix (diagonal shifted by 3)
    keeps the product finite.  The GPU product equals the oracle's, and its
    magnitude is exp(logAbsDeterminant())."""
    _case("determinant vs oracle")
    n, leaf = 400, 40
    A = 0.2 * I.exact_hodlr(n, leaf, 4, seed=11) + 3.0 * np.eye(n)
    g = HODLR_Matrix(A, leaf, "partialPiv_ACA")
    g.set_LRTolerance(1e-12)
    o = oracle.OracleHODLR(n, leaf=leaf, dense=A)
    o.set_LRTolerance(1e-12)
    dg, do = g.determinant(), o.determinant()
    lg = g.logAbsDeterminant()
    print(f"det GPU {dg:.15e} oracle {do:.15e}; log|det| {lg:.15f}", flush=True)
    assert np.isfinite(dg) and dg != 0.0
    assert abs(dg - do) <= 1e-12 * abs(do)
    assert abs(np.log(abs(dg)) - lg) <= 1e-12 * abs(lg)
    # the permutation sign is not applied: |det| agrees with LAPACK's
    # determinant of the assembled operator, the sign need not
    sign, ld = np.linalg.slogdet(A)
    assert abs(ld - lg) <= 1e-9 * abs(ld)


@pytest.mark.parametrize("pivot", [0.75, 2.5])
def test_min_pivot_truncates_to_int(oracle, pivot):
    """set_minPivot passes an int to the ACA (include/lowRank.hpp:25): 0.75
    acts as 0 (ranks and solution identical to set_minPivot(0)), 2.5 as 2.
    The matrix is 3 x a Gaussian kernel plus the identity, so entries exceed
    2 near the diagonal only and the threshold 2 changes the pivots; ranks
    and the solution equal the oracle's at the same setting."""
    _case(f"minPivot {pivot}")
    n, leaf = 1024, 64
    pts = I.points_1d_sorted(n, 21)
    K = oracle.kernel_block(pts, "gaussian", 0.1, 0.0, 0, 0, n, n)
    A = 3.0 * K + np.eye(n)
    b = I.rhs_uniform(n, 2, 22)
    g = HODLR_Matrix(A, leaf, "partialPiv_ACA")
    g.set_LRTolerance(1e-10)
    g.set_minPivot(0.0)
    x0 = g.recSM_Solve(b)
    r0 = g.nodes().copy()
    g.set_minPivot(pivot)
    x1 = g.recSM_Solve(b)
    r1 = g.nodes().copy()
    o = oracle.OracleHODLR(n, leaf=leaf, dense=A)
    o.set_LRTolerance(1e-10)
    o.set_minPivot(pivot)
    xo = o.recSM_Solve(b)
    assert np.array_equal(r1, o.nodes())
    assert rel(x1, xo) < 1e-10
    if int(pivot) == 0:
        assert np.array_equal(r0, r1) and np.array_equal(x0, x1)
    else:
        assert not np.array_equal(r0[:, 4:], r1[:, 4:])


@pytest.mark.parametrize("prec,steps,target", [("mixed", 1, 1e-10), ("fp64_refined", 2, 1e-15)])
def test_refinement_not_converged_status(prec, steps, target):
    """A refinement that ends above its target returns HODLR_ERR_NOT_CONVERGED
    (HodlrNotConverged in Python) with the best iterate in the output: its
    independently evaluated residual is the one the stats report, and it is
    no worse than the unrefined FP64 solution."""
    _case(f"not converged {prec} steps={steps} target={target:g}")
    n = 16384
    pts = I.points_2d_kd(n, 5)
    b = I.rhs_normal(n, 4, 6)
    g = HODLR_Matrix.from_kernel(pts, "gaussian", 0.05, 1e-4, sizeThreshold=256)
    g.set_LRTolerance(1e-10)
    x64 = g.recSM_Solve(b)
    r64 = np.linalg.norm(g.residual(b, x64), axis=0) / np.linalg.norm(b, axis=0)
    g.set_precision(prec)
    g.set_refinement(target, steps)
    with pytest.raises(HodlrNotConverged) as ei:
        g.recSM_Solve(b)
    x = ei.value.x
    st = g.stats()
    rc = np.linalg.norm(g.residual(b, x), axis=0) / np.linalg.norm(b, axis=0)
    print(f"{prec}: reported {st['last_residual']:.3e}, checked {rc.max():.3e}, "
          f"fp64 {r64.max():.3e}, steps {st['refine_iterations']}", flush=True)
    assert st["last_residual"] > target
    assert abs(rc.max() - st["last_residual"]) <= 1e-3 * st["last_residual"] + 1e-15
    assert st["refine_iterations"] <= steps
    if prec == "fp64_refined":
        assert rc.max() <= r64.max()


def test_factor_time_includes_root_lu():
    """factor_time spans the factorization's kernels on both streams (the
    roots' capacitance LU runs on the side stream while the next solve
    starts); factor_stream_time ends with the engine stream's part."""
    _case("factor time with the roots' LU")
    n = 16384
    pts = I.points_2d_kd(n, 7)
    b = I.rhs_normal(n, 2, 8)
    g = HODLR_Matrix.from_kernel(pts, "gaussian", 0.05, 1e-4, sizeThreshold=256)
    g.set_LRTolerance(1e-10)
    g.recSM_Solve(b)
    st = g.stats()
    print(f"factor {st['factor_time']:.4e} s, engine stream {st['factor_stream_time']:.4e} s",
          flush=True)
    assert st["factor_time"] > 0 and st["factor_stream_time"] > 0
    assert st["factor_time"] >= st["factor_stream_time"]


def test_factorization_kernel_groups_profiled():
    """With profiling on, the factorization's kernel groups carry their event
    times and work: the LU flops are (2/3) n^3 over every leaf and node
    matrix (n = kt + kb), the solves 2 n^2 per column over the leaves'
    ancestor columns, and the grouped GEMMs' flops the node products'; the
    groups' times fit inside the factorization's."""
    _case("factorization kernel groups")
    n = 16384
    pts = I.points_2d_kd(n, 7)
    g = HODLR_Matrix.from_kernel(pts, "gaussian", 0.05, 1e-4, sizeThreshold=256)
    g.set_LRTolerance(1e-10)
    g.storeLRinTree()
    g.set_profiling(True)
    g.recSM_Factorize()
    st = g.stats()
    fg = st["factor_groups"]
    nodes = g.nodes()
    leaf = n // (2 ** (int(nodes[:, 0].max()) + 1))
    K = nodes[:, 4] + nodes[:, 5]
    lu = 2.0 / 3.0 * (float(np.sum(K.astype(float) ** 3)) + (n // leaf) * float(leaf) ** 3)
    print({k: (v["time"], v["flops"], v["launches"]) for k, v in fg.items()}, flush=True)
    assert abs(fg["lu"]["flops"] - lu) <= 1e-9 * lu
    for k in ("gemm", "lu", "leaf_solve", "node_solve"):
        assert fg[k]["time"] > 0 and fg[k]["launches"] > 0 and fg[k]["flops"] > 0, k
        assert fg[k]["bound_time"] > 0, k
    assert sum(v["time"] for v in fg.values()) <= 1.05 * st["factor_time"]


def test_dense_constructor_from_binary_file(tmp_path, oracle):
    """HODLR_Matrix(readBinaryIntoMatrixXd(path), ...) as the reference's
    drivers build it (helperFunctions.cpp:370-371): the handle made from the
    file equals the one made from the array it was written from."""
    from dense_hodlr_amd import saveMatrixXdToBinary
    _case("dense constructor from a binary file")
    n, leaf = 1200, 50
    A = I.exact_hodlr(n, leaf, 6, seed=31)
    path = tmp_path / "A.bin"
    saveMatrixXdToBinary(A, path)
    b = I.rhs_uniform(n, 2, 32)
    g1 = HODLR_Matrix(A, leaf, "partialPiv_ACA")
    g2 = HODLR_Matrix.from_binary_file(path, leaf, "partialPiv_ACA")
    for g in (g1, g2):
        g.set_LRTolerance(1e-10)
    x1, x2 = g1.recSM_Solve(b), g2.recSM_Solve(b)
    assert np.array_equal(g1.nodes(), g2.nodes()) and np.array_equal(x1, x2)
    o = oracle.OracleHODLR(n, leaf=leaf, dense=A)
    o.set_LRTolerance(1e-10)
    assert rel(x2, o.recSM_Solve(b)) < 1e-10


def test_device_peaks_are_sustained_ceilings():
    """hodlr_measure_device_peaks: every probe returns a positive figure no
    larger than the MI355X datasheet peak it is a sustained share of (HBM
    8 TB/s, FP64 78.6 TFLOP/s, FP32 157.3 TFLOP/s), and the streaming read
    reaches at least a quarter of the HBM peak."""
    from dense_hodlr_amd import device_peaks
    p = device_peaks()
    print(p, flush=True)
    assert 2000.0 < p["hbm_read_gbs"] <= 8000.0 * 1.02
    assert 0.0 < p["hbm_copy_gbs"] <= 8000.0 * 1.02
    assert 0.0 < p["fp64_mfma_tflops"] <= 78.6 * 1.02
    assert 0.0 < p["fp64_valu_tflops"] <= 78.6 * 1.02
    assert 0.0 < p["fp32_mfma_tflops"] <= 157.3 * 1.02
