"""GPU parity tests: the HIP path through the C ABI against the oracle on the
same seeded inputs, plus the reference's own solution-level tests
(test/test.cpp) restated at sizes the oracle finishes in seconds."""
import sys

import numpy as np
import pytest

from dense_hodlr_amd import HODLR_Matrix, HodlrNotConverged
from dense_hodlr_amd import inputs as I

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def _pair(oracle, n, pts, family, ell, diag, leaf, tol, diag_mode="add"):
    g = HODLR_Matrix.from_kernel(pts, family, ell, diag, diag_mode, sizeThreshold=leaf)
    g.set_LRTolerance(tol)
    o = oracle.OracleHODLR(n, leaf=leaf, points=pts, family=family, length_scale=ell, diag=diag,
                           diag_mode={"add": 0, "replace": 1}[diag_mode])
    o.set_LRTolerance(tol)
    return g, o


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def gauss1d(oracle):
    """1D Gaussian workload: N=4096, l=0.1, +1e-2 on the diagonal, leaf 64, tol 1e-10."""
    cfg = I.config_inputs("gauss1d_4096")
    _case("gauss1d_4096")
    g, o = _pair(oracle, cfg["n"], cfg["points"], "gaussian", 0.1, 1e-2, 64, 1e-10)
    g.storeLRinTree()
    o.compress()
    return cfg, g, o


def test_ranks_identical(gauss1d):
    cfg, g, o = gauss1d
    _case("ranks gauss1d_4096")
    gn, on = g.nodes(), o.nodes()
    assert np.array_equal(gn[:, :4], on[:, :4])  # same tree
    assert np.array_equal(gn[:, 4:], on[:, 4:])  # same ACA ranks, block by block


def test_factors_match_oracle(gauss1d):
    """Kernel entries and ACA residuals are computed with the same IEEE
    operations in the same order as the oracle, so with identical pivots the
    factors are bit-identical."""
    cfg, g, o = gauss1d
    _case("factors gauss1d_4096")
    bad = []
    for idx in range(len(o.nodes())):
        for which in range(4):
            a, b = g.factor(idx, which), o.factor(idx, which)
            assert a.shape == b.shape
            if not np.array_equal(a, b):
                cols = np.nonzero(np.any(a != b, axis=0))[0]
                bad.append((idx, which, int(cols[0]), float(np.abs(a - b).max())))
    assert not bad, bad[:10]


def test_factors_match_oracle_many_blocks(oracle):
    """The same bit-for-bit equality when the ACA launches are large enough
    (more than 768 workgroups: 2046 blocks of a 1024-leaf tree) to take the
    half-width-piece variant of the panel passes in the early iterations and
    the full-width one later (the variant is chosen by workgroup count), over
    many chunks of stored columns: 2D Gaussian N=8192 in kd order, l=0.1,
    leaf 8 (ranks up to ~300)."""
    n = 8192
    pts = I.points_2d_kd(n, 4242)
    _case("factors gauss2d_8192 leaf 8")
    g, o = _pair(oracle, n, pts, "gaussian", 0.1, 1e-2, 8, 1e-10)
    g.storeLRinTree()
    o.compress()
    assert np.array_equal(g.nodes(), o.nodes())
    bad = []
    for idx in range(len(o.nodes())):
        for which in range(4):
            a, b = g.factor(idx, which), o.factor(idx, which)
            if a.shape != b.shape or not np.array_equal(a, b):
                bad.append((idx, which))
    assert not bad, bad[:10]


def test_factors_match_oracle_multi_workgroup_finalize(oracle):
    """Bit-for-bit factors and ranks when the top-level blocks finalize over
    several workgroups: 2D Gaussian N=16384 in kd order, l=0.05, leaf 256.
    Once only the two 8192-row top blocks are left, their panel rows spread
    to one per thread (32 row and 32 column workgroups each), so a block has
    64 partial slots (more than the 32 of a one-workgroup finalize) and,
    past rank 65, several finalize workgroups whose partials the last one
    combines."""
    n = 16384
    pts = I.points_2d_kd(n, 515)
    _case("factors gauss2d_16384 leaf 256")
    g, o = _pair(oracle, n, pts, "gaussian", 0.05, 1e-4, 256, 1e-10)
    g.storeLRinTree()
    o.compress()
    assert np.array_equal(g.nodes(), o.nodes())
    assert g.nodes()[0, 4] > 65 and g.nodes()[0, 5] > 65  # several finalize chunks
    bad = []
    for idx in range(len(o.nodes())):
        for which in range(4):
            a, b = g.factor(idx, which), o.factor(idx, which)
            if a.shape != b.shape or not np.array_equal(a, b):
                bad.append((idx, which))
    assert not bad, bad[:10]


def test_recsm_matches_oracle(gauss1d, oracle):
    cfg, g, o = gauss1d
    _case("recSM gauss1d_4096")
    xg = g.recSM_Solve(cfg["rhs"])
    xo = o.recSM_Solve(cfg["rhs"])
    assert rel(xg, xo) < 1e-10


def test_reclu_matches_oracle(gauss1d):
    cfg, g, o = gauss1d
    _case("recLU gauss1d_4096")
    xg = g.recLU_Solve(cfg["rhs"])
    xo = o.recLU_Solve(cfg["rhs"])
    assert rel(xg, xo) < 1e-10
    assert rel(xg, g.recSM_Solve(cfg["rhs"])) < 1e-10


def test_logdet_matches_oracle(gauss1d):
    cfg, g, o = gauss1d
    _case("logdet gauss1d_4096")
    lg, lo = g.logAbsDeterminant(), o.logAbsDeterminant()
    assert abs(lg - lo) <= 1e-12 * abs(lo)


def test_matvec_matches_oracle(gauss1d):
    cfg, g, o = gauss1d
    _case("matvec gauss1d_4096")
    X = I.rhs_normal(cfg["n"], 3, 77)
    assert rel(g.matvec(X), o.matvec(X)) < 1e-13


def test_solution_against_dense_lu(gauss1d, oracle):
    cfg, g, o = gauss1d
    _case("dense LU gauss1d_4096")
    n = cfg["n"]
    A = oracle.kernel_block(cfg["points"], "gaussian", 0.1, 1e-2, 0, 0, n, n)
    xd, _ = oracle.dense_solve(A, cfg["rhs"])
    xg = g.recSM_Solve(cfg["rhs"])
    assert rel(xg, xd) < 1e-7
    assert np.linalg.norm(A @ xg - cfg["rhs"]) / (np.linalg.norm(A, 2) * np.linalg.norm(xg)) < 1e-10


def test_exact_hodlr_dense_constructor(oracle):
    # test/test.cpp:57-69, :76-88 at n = 1500: rel. solution error < 1e-8
    _case("exact HODLR 1500")
    n, leaf = 1500, 50
    A = I.exact_hodlr(n, leaf, 10, seed=4)
    x = I.linspace_m2_2(n)
    b = A @ x
    g = HODLR_Matrix(A, leaf, "partialPiv_ACA")
    g.set_LRTolerance(1e-10)
    assert rel(g.recSM_Solve(b)[:, 0], x) < 1e-8
    assert rel(g.recLU_Solve(b)[:, 0], x) < 1e-8
    o = oracle.OracleHODLR(n, leaf=leaf, dense=A)
    o.set_LRTolerance(1e-10)
    o.compress()
    assert np.array_equal(g.nodes(), o.nodes())


@pytest.mark.parametrize("n", [900, 1500])
def test_blocked_one_launch_solve_matches_step_kernel(monkeypatch, n):
    """Systems of 513-2048 rows are substituted in blocks of 64 rows (the
    owning wave solves the diagonal block, the workgroup updates the rest);
    each row receives its updates in the same order as in the step-by-step
    kernel, so the results are bit-identical with the blocks switched off
    (HODLR_NO_LU_SOLVE_BLOCKS).  One dense leaf of n rows (the leaf inverse
    solves n columns through it), random (pivoting at most columns), and a
    kernel matrix whose leaves have ~750 rows."""
    _case(f"blocked one-launch solve n={n}")
    rng = np.random.default_rng(7 + n)
    A = rng.standard_normal((n, n))
    b = rng.standard_normal((n, 3))
    pts = I.points_2d_kd(6000, 71)
    bk = I.rhs_normal(6000, 2, 72)
    out = []
    for flag in (None, "1"):
        if flag:
            monkeypatch.setenv("HODLR_NO_LU_SOLVE_BLOCKS", flag)
        else:
            monkeypatch.delenv("HODLR_NO_LU_SOLVE_BLOCKS", raising=False)
        g = HODLR_Matrix(A, n + 1, "partialPiv_ACA")
        xs = g.recSM_Solve(b)
        ld = g.logAbsDeterminant()
        del g
        h = HODLR_Matrix.from_kernel(pts, "gaussian", 0.05, 1e-3, sizeThreshold=800)
        h.set_LRTolerance(1e-10)
        xk = h.recSM_Solve(bk)
        xl = h.recLU_Solve(bk)
        del h
        out.append((xs, ld, xk, xl))
    for u, v in zip(out[0], out[1]):
        assert np.array_equal(np.asarray(u), np.asarray(v))
    assert rel(out[0][0], np.linalg.solve(A, b)) < 1e-13 * np.linalg.cond(A)


@pytest.mark.parametrize("n", [200, 500, 1000, 2000, 3000])
def test_dense_lu_pivots_match_lapack(n):
    """The batched LU's panels pick the first maximum of each column, as
    LAPACK's getrf (idamax): random dense matrices of one leaf each, sized
    for each register-panel variant (256 or 512 threads, 1-8 rows per
    thread), pivot at almost every column.  determinant() is the product of
    the leaf LU's diagonal without the interchanges' sign, so it equals
    LAPACK's product of U's diagonal, sign included, only if every pivot row
    is the same.  Scaled to |det| = 1."""
    import scipy.linalg as sla
    _case(f"dense LU pivots n={n}")
    for seed in range(3):
        G = np.random.default_rng(90 + 7 * n + seed).standard_normal((n, n))
        A = G / np.exp(np.linalg.slogdet(G)[1] / n)
        dl = float(np.prod(np.diag(sla.lu_factor(A)[0])))
        g = HODLR_Matrix(A, n + 1, "partialPiv_ACA")
        dg = g.determinant()
        del g
        print(f"n={n} seed={seed} det GPU {dg:.15e} LAPACK {dl:.15e}", flush=True)
        assert np.sign(dg) == np.sign(dl) and abs(dg - dl) <= 1e-9 * abs(dl)


def test_root_lu_side_stream_matches_in_order(monkeypatch):
    """The roots' capacitance LU runs on a second stream while the following
    solve works through the lower levels (the solve waits for it at the
    root level); switched off (HODLR_NO_ROOT_OVERLAP) everything runs in
    order on one stream.  Same kernels, same data: recSM and recLU
    solutions, the refined solve and the log-det are bit-identical, also for
    a batch of instances (one root each) and across repeated solves."""
    from dense_hodlr_amd import HODLR_Batch
    _case("root LU side stream vs in order")
    n = 6000
    pts = I.points_2d_kd(n, 51)
    b = I.rhs_normal(n, 4, 52)
    bpts = [I.points_2d_kd(m, 60 + m) for m in (1500, 2600, 900)]
    out = []
    for flag in (None, "1"):
        if flag:
            monkeypatch.setenv("HODLR_NO_ROOT_OVERLAP", flag)
        else:
            monkeypatch.delenv("HODLR_NO_ROOT_OVERLAP", raising=False)
        g = HODLR_Matrix.from_kernel(pts, "gaussian", 0.05, 1e-4, sizeThreshold=128)
        g.set_LRTolerance(1e-10)
        xs1 = g.recSM_Solve(b)
        xs2 = g.recSM_Solve(b)
        xl = g.recLU_Solve(b)
        ld = g.logAbsDeterminant()
        g.set_precision("fp64_refined")
        xr = g.recSM_Solve(b)
        del g
        bt = HODLR_Batch(bpts, "se", 0.1, 1e-2, sizeThreshold=64)
        bb = I.rhs_normal(sum(len(p) for p in bpts), 1, 53)
        xb = bt.recSM_Solve(bb)
        lb = bt.logAbsDeterminants()
        del bt
        out.append((xs1, xs2, xl, ld, xr, xb, lb))
    for u, v in zip(out[0], out[1]):
        assert np.array_equal(np.asarray(u), np.asarray(v))
    assert np.array_equal(out[0][0], out[0][1])


def test_streamed_updates_match_tiled(monkeypatch):
    """The factorization's small-rank updates W -= W' Y (rank <= 64, at least
    1024 workgroups) go through the streaming matrix-core kernel, which keeps
    its rows of W' in registers and prefetches the next column tile of W;
    switched off (HODLR_NO_STREAM_UPDATE) they go through the tiled kernel.
    The accumulation order is the same, so recSM and recLU solutions and the
    log-det are bit-identical (65536 points in 2D, leaf 128: the deepest
    levels' updates have 1024 workgroups and more)."""
    _case("streamed vs tiled small-rank updates")
    n = 65536
    pts = I.points_2d_kd(n, 71)
    b = I.rhs_normal(n, 2, 72)
    out = []
    for flag in (None, "1"):
        if flag:
            monkeypatch.setenv("HODLR_NO_STREAM_UPDATE", flag)
        else:
            monkeypatch.delenv("HODLR_NO_STREAM_UPDATE", raising=False)
        g = HODLR_Matrix.from_kernel(pts, "gaussian", 0.05, 1e-3, sizeThreshold=128)
        g.set_LRTolerance(1e-8)
        out.append((g.recSM_Solve(b), g.recLU_Solve(b), g.logAbsDeterminant(), g.nodes()))
        del g
    assert np.array_equal(out[0][3], out[1][3])
    assert max(np.max(out[0][3][out[0][3][:, 0] >= 6][:, 4:]), 1) <= 64  # deep ranks stream
    for u, v in zip(out[0][:3], out[1][:3]):
        assert np.array_equal(np.asarray(u), np.asarray(v))


@pytest.mark.parametrize("n", [3000, 4500])
def test_single_large_leaf_lu_paths(n):
    """One dense leaf of n rows (sizeThreshold above n): the batched LU at
    n = 3000 runs 4-column register panels whose interchanges the step
    kernel applies to the columns left and right of each panel; at n = 4500
    the global-memory panel kernel swaps whole rows itself.  A random dense
    matrix pivots at almost every column.  Solutions against LAPACK's
    (numpy.linalg.solve) within cond(A) x 1e-13, the normwise residual
    within 1e-12 (the leaf solve multiplies by the explicit inverse, whose
    residual is not LAPACK's few ulps; measured 1.3e-14 at n = 3000),
    log|det| against slogdet."""
    _case(f"single leaf n={n}")
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    b = rng.standard_normal((n, 2))
    g = HODLR_Matrix(A, n + 1, "partialPiv_ACA")
    xs = g.recSM_Solve(b)
    xl = g.recLU_Solve(b)
    xr = np.linalg.solve(A, b)
    sign, ld = np.linalg.slogdet(A)
    cond = np.linalg.cond(A)
    print(f"cond {cond:.2e} rel diff recSM {rel(xs, xr):.2e} recLU {rel(xl, xr):.2e}", flush=True)
    assert rel(xs, xr) < 1e-13 * cond and rel(xl, xr) < 1e-13 * cond
    assert np.linalg.norm(A @ xs - b) / (np.linalg.norm(A, 2) * np.linalg.norm(xs)) < 1e-12
    assert abs(g.logAbsDeterminant() - ld) < 1e-10 * abs(ld)


def test_radial_imq_replace_diagonal(oracle):
    # test/test.cpp:112-148 at n = 3000: IMQ, diagonal 0, leaf 30, tol 1e-8, < 1e-6
    _case("IMQ radial 3000")
    n = 3000
    pts = np.linspace(-1, 1, n).reshape(-1, 1)
    A = oracle.kernel_block(pts, "imq", 1.0, 0.0, 0, 0, n, n, diag_mode=1)
    x = I.linspace_m2_2(n)
    b = A @ x
    g, o = _pair(oracle, n, pts, "imq", 1.0, 0.0, 30, 1e-8, diag_mode="replace")
    xs, xl = g.recSM_Solve(b)[:, 0], g.recLU_Solve(b)[:, 0]
    assert rel(xs, x) < 1e-6 and rel(xl, x) < 1e-6
    assert rel(xs, o.recSM_Solve(b)[:, 0]) < 1e-9


@pytest.mark.parametrize("n,leaf", [(1000, 30), (777, 64), (50, 64), (129, 64)])
def test_uneven_trees(oracle, n, leaf):
    _case(f"uneven n={n} leaf={leaf}")
    pts = I.points_1d_sorted(n, n)
    g, o = _pair(oracle, n, pts, "gaussian", 0.2, 0.1, leaf, 1e-10)
    b = I.rhs_uniform(n, 2, 3)
    xg, xo = g.recSM_Solve(b), o.recSM_Solve(b)
    assert rel(xg, xo) < 1e-11
    if n >= leaf:
        assert np.array_equal(g.nodes(), o.nodes())
        assert rel(g.recLU_Solve(b), o.recLU_Solve(b)) < 1e-11
        assert abs(g.logAbsDeterminant() - o.logAbsDeterminant()) < 1e-11 * abs(o.logAbsDeterminant())


def test_matern2d_small(oracle):
    _case("matern2d 8192")
    n = 8192
    pts = I.points_2d_kd(n, 31)
    g, o = _pair(oracle, n, pts, "matern32", 0.2, 1e-3, 128, 1e-10)
    b = I.rhs_normal(n, 4, 32)
    g.storeLRinTree()
    o.compress()
    gn, on = g.nodes(), o.nodes()
    assert np.array_equal(gn, on)
    xg, xo = g.recSM_Solve(b), o.recSM_Solve(b)
    assert rel(xg, xo) < 1e-10
    r = b - o.matvec(xg)
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-8


def test_callback_constructor(oracle):
    _case("callback constructor")
    n = 300
    xs = np.linspace(-2, 2, n)

    def quad(i, j):  # test/test.cpp:12-17
        if i == j:
            return 0.0
        r = xs[i] - xs[j]
        return r * r + 1

    g = HODLR_Matrix.from_callback(n, quad, sizeThreshold=30)
    A = np.array([[quad(i, j) for j in range(n)] for i in range(n)])
    o = oracle.OracleHODLR(n, leaf=30, dense=A)
    b = A @ xs
    assert rel(g.recLU_Solve(b), o.recLU_Solve(b)) < 1e-9


def test_settings_invalidate_cache(oracle):
    _case("settings")
    n = 2048
    pts = I.points_1d_sorted(n, 9)
    g = HODLR_Matrix.from_kernel(pts, "gaussian", 0.1, 1e-2, sizeThreshold=64)
    g.set_LRTolerance(1e-4)
    b = I.rhs_uniform(n, 1, 1)
    x1 = g.recSM_Solve(b)
    r1 = g.nodes()[:, 4:].sum()
    g.set_LRTolerance(1e-10)
    x2 = g.recSM_Solve(b)
    assert g.nodes()[:, 4:].sum() > r1
    assert rel(x1, x2) > 1e-12


def test_underflowing_kernel_matches_oracle(oracle):
    """Squared exponential with l = 0.0103 on 1321 sorted points: the far
    entries of the top blocks are subnormal or zero.  Pivot candidates below
    the floor 2^-511 count as zero on both sides, so the factors stay finite
    and bit-identical to the oracle's."""
    _case("se l=0.0103 n=1321 subnormal far field")
    cfg = I.batch_inputs(24, 1024, 8192, seed=5)
    q = 19
    n, ell = cfg["n"][q], cfg["length_scale"][q]
    assert n == 1321 and abs(ell - 0.0103) < 1e-4
    g, o = _pair(oracle, n, cfg["points"][q], "se", ell, 1e-2, 64, 1e-10)
    g.storeLRinTree()
    o.compress()
    assert np.array_equal(g.nodes(), o.nodes())
    for idx in range(len(o.nodes())):
        for which in range(4):
            a, b = g.factor(idx, which), o.factor(idx, which)
            assert np.all(np.isfinite(a)) and np.array_equal(a, b), (idx, which)
    b = cfg["rhs"][q]
    xg, xo = g.recSM_Solve(b), o.recSM_Solve(b)
    assert np.all(np.isfinite(xg)) and rel(xg, xo) < 1e-10


@pytest.mark.parametrize("n,leaf,ell", [(512, 32, 0.1), (777, 64, 0.2)])
def test_svd_method_matches_oracle(oracle, n, leaf, ell):
    """set_LRMethod("SVD"): batched one-sided Jacobi on the GPU.  Same ranks
    as the oracle's Jacobi SVD (absolute threshold), the same low-rank blocks
    U Sigma V^T to rounding, the same recSM solution and log-determinant."""
    _case(f"svd method n={n} leaf={leaf}")
    pts = I.points_1d_sorted(n, 31)
    b = I.rhs_uniform(n, 2, 32)
    g, o = _pair(oracle, n, pts, "gaussian", ell, 1e-2, leaf, 1e-10)
    g.set_LRMethod("SVD")
    o.set_LRMethod("SVD")
    g.storeLRinTree()
    o.compress()
    gn, on = g.nodes(), o.nodes()
    assert np.array_equal(gn, on)
    for idx in range(len(on)):
        for wu, wv in ((0, 1), (2, 3)):
            Sg = g.factor(idx, wu) @ g.factor(idx, wv).T
            So = o.factor(idx, wu) @ o.factor(idx, wv).T
            assert np.linalg.norm(Sg - So) <= 1e-12 * max(np.linalg.norm(So), 1e-300), (idx, wu)
    xg, xo = g.recSM_Solve(b), o.recSM_Solve(b)
    assert rel(xg, xo) < 1e-10
    assert abs(g.logAbsDeterminant() - o.logAbsDeterminant()) < 1e-10 * abs(o.logAbsDeterminant())
    # switching back to ACA recompresses
    g.set_LRMethod("partialPiv_ACA")
    o.set_LRMethod("partialPiv_ACA")
    assert rel(g.recSM_Solve(b), o.recSM_Solve(b)) < 1e-10


def test_svd_method_gauss1d_4096(gauss1d):
    """SVD compression at the N=4096 workload (2048 x 2048 root blocks): the
    solution agrees with the ACA one to the compression tolerance and with a
    dense LU of the same kernel matrix."""
    cfg, g, o = gauss1d
    _case("svd gauss1d_4096")
    import time
    h = HODLR_Matrix.from_kernel(cfg["points"], "gaussian", 0.1, 1e-2, sizeThreshold=64)
    h.set_LRTolerance(1e-10)
    h.set_LRMethod("SVD")
    t = time.perf_counter()
    h.storeLRinTree()
    t = time.perf_counter() - t
    xs = h.recSM_Solve(cfg["rhs"])
    xa = g.recSM_Solve(cfg["rhs"])
    st = h.stats()
    print(f"SVD compression {t:.3f} s ({st['aca_iterations']} Jacobi sweeps); ranks {st['max_rank']}; "
          f"rel diff to ACA solution {rel(xs, xa):.2e}", flush=True)
    assert rel(xs, xa) < 1e-7
    assert max(st["max_rank"]) <= max(g.stats()["max_rank"]) + 2


@pytest.mark.parametrize("shapes", [
    dict(aca_rows_per_thread=1, gemm_tile_m=64, gemm_split_target=1024, lu_panel=16,
         gemm_mfma=False),
    dict(aca_rows_per_thread=4, gemm_tile_m=128, gemm_split_target=4096, lu_panel=32,
         gemm_mfma=False),
    dict(aca_rows_per_thread=2, gemm_tile_m=128, gemm_split_target=2048, lu_panel=64,
         gemm_mfma=True),
    dict(aca_rows_per_thread=2, gemm_tile_m=128, gemm_split_target=64, lu_panel=64,
         gemm_mfma=True),
    dict(aca_rows_per_thread=4, gemm_tile_m=64, gemm_split_target=4096, lu_panel=32,
         gemm_mfma=True),
    dict(aca_rows_per_thread=4, gemm_tile_m=0, gemm_split_target=1024, lu_panel=32,
         gemm_mfma=True),
    dict(aca_rows_per_thread=4, gemm_tile_m=0, gemm_split_target=4096, lu_panel=32,
         gemm_mfma=True, leaf_solve=1),
    dict(aca_rows_per_thread=4, gemm_tile_m=0, gemm_split_target=4096, lu_panel=32,
         gemm_mfma=True, leaf_solve=2),
    dict(aca_rows_per_thread=4, gemm_tile_m=0, gemm_split_target=4096, lu_panel=32,
         gemm_mfma=True, leaf_solve=3),
])
def test_launch_shapes_match_oracle(oracle, shapes):
    """Every launch-shape variant (ACA rows per thread, vector or matrix-core
    FP64 GEMM tiles of either shape or chosen per launch, split-K, LU panel
    width) gives the oracle's ranks and its
    recSM / recLU solutions and log-determinant."""
    _case(f"launch shapes {shapes}")
    cfg = I.config_inputs("gauss1d_4096")
    g, o = _pair(oracle, cfg["n"], cfg["points"], "gaussian", 0.1, 1e-2, 64, 1e-10)
    g.set_launch_shapes(**shapes)
    b = I.rhs_uniform(cfg["n"], 3, 5)
    xs = g.recSM_Solve(b)
    xl = g.recLU_Solve(b)
    xo = o.recSM_Solve(b)
    assert np.array_equal(g.nodes(), o.nodes())
    assert rel(xs, xo) < 1e-10 and rel(xl, xo) < 1e-9
    assert abs(g.logAbsDeterminant() - o.logAbsDeterminant()) < 1e-11 * abs(o.logAbsDeterminant())
    assert g.stats()["tuning"]["gemm_mfma"] == bool(shapes.get("gemm_mfma", True))


@pytest.mark.parametrize("limit", [0, 40])
def test_blocked_lu_solves_match_oracle(oracle, monkeypatch, limit):
    """Systems above the one-launch solve's size limit (2048 rows; the root
    capacitance of large problems) take the blocked path: permutation, panel
    TRSM and GEMM updates.  Lowering the limit for a handle (0: every system,
    40: leaves and deep nodes one-launch, the rest blocked) routes the
    workload through that path; ranks, solutions and log-det stay the
    oracle's."""
    _case(f"blocked LU solves below limit {limit}")
    monkeypatch.setenv("HODLR_LU_SOLVE_SMALL_MAX", str(limit))
    cfg = I.config_inputs("gauss1d_4096")
    g, o = _pair(oracle, cfg["n"], cfg["points"], "gaussian", 0.1, 1e-2, 64, 1e-10)
    b = I.rhs_uniform(cfg["n"], 3, 7)
    xs = g.recSM_Solve(b)
    xl = g.recLU_Solve(b)
    xo = o.recSM_Solve(b)
    assert np.array_equal(g.nodes(), o.nodes())
    assert rel(xs, xo) < 1e-10 and rel(xl, xo) < 1e-9
    assert abs(g.logAbsDeterminant() - o.logAbsDeterminant()) < 1e-11 * abs(o.logAbsDeterminant())


@pytest.mark.parametrize("leaf", [40, 100, 200, 400])
def test_wave_lu_solves_match_workgroup_kernel(monkeypatch, leaf):
    """Systems of at most 512 rows are solved one wave per system and column
    chunk (no barriers, broadcasts by readlane); the same operations as the
    workgroup kernel, so recSM, recLU and the log-det are bit-identical with
    the wave kernel switched off.  Leaf sizes 40/100/200/400 (ragged trees
    of 3000 points) put the leaves and node systems on one, two, four and
    eight rows per lane; the oracle tests above run the wave kernel (the
    default)."""
    _case(f"wave LU solves vs workgroup kernel, leaf {leaf}")
    n = 3000
    pts = I.points_2d_kd(n, 41)
    b = I.rhs_normal(n, 5, 42)
    out = []
    for wave_max in ("256", "0"):
        monkeypatch.setenv("HODLR_LU_SOLVE_WAVE_MAX", wave_max)
        g = HODLR_Matrix.from_kernel(pts, "gaussian", 0.08, 1e-3, sizeThreshold=leaf)
        g.set_LRTolerance(1e-10)
        out.append((g.recSM_Solve(b), g.recLU_Solve(b), g.logAbsDeterminant(), g.nodes()))
        del g
    (xs0, xl0, d0, n0), (xs1, xl1, d1, n1) = out
    assert np.array_equal(n0, n1)
    assert np.array_equal(xs0, xs1) and np.array_equal(xl0, xl1)
    assert d0 == d1


def test_matern2d_65536_against_oracle_summary():
    """The N=65536 Matern-3/2 workload at full size (16 right-hand sides, the
    log-determinant) against the oracle's recSM on the same seeded inputs,
    through the committed summary (tests/golden/matern2d_65536_oracle.npz,
    made by make_matern2d_summary.py): identical ACA ranks node by node; the
    plain recSM solution within 1e-10 of the oracle's (512 sampled rows,
    column norms); the solution refined with double-double residuals within
    1e-10 of the oracle's refined one (both converge to the solution of
    A_H x = b rounded, whatever order the solvers sum in); log-det within
    1e-12."""
    import os

    from dense_hodlr_amd import HodlrNotConverged
    _case("matern2d_65536 vs oracle summary")
    f = np.load(os.path.join(os.path.dirname(__file__), "golden", "matern2d_65536_oracle.npz"))
    cfg = I.config_inputs("matern2d_65536")
    g = HODLR_Matrix.from_kernel(cfg["points"], "matern32", 0.2, 1e-3, sizeThreshold=cfg["leaf"])
    g.set_LRTolerance(cfg["tol"])
    b = cfg["rhs"]
    x = g.recSM_Solve(b)
    assert np.array_equal(g.nodes(), f["nodes"])

    def cmp(x, xs_ref, cn_ref):
        cn = np.linalg.norm(x, axis=0)
        dn = np.max(np.abs(cn - cn_ref) / cn_ref)
        xs = x[f["idx"], :]
        return dn, np.max(np.linalg.norm(xs - xs_ref, axis=0) / np.linalg.norm(xs_ref, axis=0))

    dn, rel = cmp(x, f["xs"], f["colnorm"])
    res = np.max(np.linalg.norm(g.residual(b, x), axis=0) / np.linalg.norm(b, axis=0))
    g.set_precision("fp64_refined")
    g.set_refinement(1e-13, 4)
    try:
        xr = g.recSM_Solve(b)
    except HodlrNotConverged as e:  # the floor of this system's residual: the best iterate
        xr = e.x
    dnr, relr = cmp(xr, f["xs_refined"], f["colnorm_refined"])
    print(f"plain: column norms {dn:.2e}, sampled relative error {rel:.2e}, residual {res:.2e}; "
          f"refined: column norms {dnr:.2e}, sampled relative error {relr:.2e} (oracle residual "
          f"{float(f['refine_residual']):.1e})", flush=True)
    assert dn < 1e-10 and rel < 1e-10
    assert dnr < 1e-12 and relr < 1e-10
    ld = g.logAbsDeterminant()
    assert abs(ld - float(f["logdet"])) <= 1e-12 * abs(float(f["logdet"]))


@pytest.mark.parametrize("leaf", [40, 100, 200, 400])
def test_shared_factor_leaf_solves_match_wave_kernel(leaf):
    """The factorization's leaf steps (hundreds of columns) solve with one
    wave per 8-column chunk streaming the whole factor (leaf_solve 0), with
    the factor shared by a workgroup's four waves (2), or with one wave per
    16-column chunk (4, the default for leaves of at most 256 rows): the same
    operations in the same order per element, so recSM, recLU and the log-det
    are bit-identical.  Leaf sizes 40/100/200/400 (ragged trees
    of 6000 points) cover one, two, four and eight rows per lane."""
    _case(f"shared-factor leaf solves vs one wave, leaf {leaf}")
    n = 6000
    pts = I.points_2d_kd(n, 43)
    b = I.rhs_normal(n, 3, 44)
    out = []
    for ls in (0, 2, 4):
        g = HODLR_Matrix.from_kernel(pts, "gaussian", 0.06, 1e-3, sizeThreshold=leaf)
        g.set_launch_shapes(leaf_solve=ls)
        g.set_LRTolerance(1e-10)
        out.append((g.recSM_Solve(b), g.recLU_Solve(b), g.logAbsDeterminant(), g.nodes()))
        del g
    (xs0, xl0, d0, n0) = out[0]
    assert np.sum(n0[:, 4:]) > 128  # the leaf steps have many columns
    for xs1, xl1, d1, n1 in out[1:]:
        assert np.array_equal(n0, n1)
        assert np.array_equal(xs0, xs1) and np.array_equal(xl0, xl1)
        assert d0 == d1


@pytest.mark.parametrize("n,nrhs", [(37, 130), (64, 200), (130, 128), (250, 200), (256, 333)])
def test_mfma_leaf_solve_matches_lapack(n, nrhs):
    """The matrix-core LU solve (leaf_solve 3; solves of 128 or more columns
    with leaves of at most 256 rows) on one dense leaf: a random matrix
    (interchanges at almost every column), row counts off and on the 16-row
    tiles, column counts off and on the 16-column chunks, against LAPACK's
    solve of the same system."""
    _case(f"matrix-core leaf solve n={n} nrhs={nrhs}")
    rng = np.random.default_rng(1000 + n)
    A = rng.standard_normal((n, n))
    B = rng.standard_normal((n, nrhs))
    g = HODLR_Matrix(A, 256)
    g.set_launch_shapes(leaf_solve=3)
    X = g.recSM_Solve(B)
    Xr = np.linalg.solve(A, B)
    err = np.max(np.abs(X - Xr)) / np.max(np.abs(Xr))
    print(f"max error {err:.2e}, cond {np.linalg.cond(A):.1e}", flush=True)
    assert err < 1e-11


@pytest.mark.parametrize("leaf", [40, 100, 200, 256])
def test_mfma_leaf_solves_match_oracle(oracle, leaf):
    """The factorization's leaf steps on the matrix cores (leaf_solve 3): the
    off-diagonal tiles are summed four terms at a time, so the results are
    not bit-identical to the substitution kernels.  The ranks (decided before
    the factorization) are the oracle's; the unrefined recSM and recLU
    solutions agree with the oracle's to the rounding-order difference times
    the conditioning (1e-9) with a backward error no larger than the
    oracle's (double-double residuals against A_H); the refined solutions
    (both sides refined with double-double residuals, which removes the
    rounding-order dependence) agree to 1e-10; the log-det to 1e-11.
    Ragged trees of 6000 points with leaves of up to 40, 100, 200 and 256
    rows."""
    _case(f"matrix-core leaf solves vs oracle, leaf {leaf}")
    n = 6000
    pts = I.points_2d_kd(n, 43)
    b = I.rhs_normal(n, 3, 44)
    g = HODLR_Matrix.from_kernel(pts, "gaussian", 0.06, 1e-3, sizeThreshold=leaf)
    g.set_launch_shapes(leaf_solve=3)
    g.set_LRTolerance(1e-10)
    xs, xl, ld = g.recSM_Solve(b), g.recLU_Solve(b), g.logAbsDeterminant()
    o = oracle.OracleHODLR(n, leaf=leaf, points=pts, family="gaussian", length_scale=0.06,
                           diag=1e-3)
    o.set_LRTolerance(1e-10)
    xo = o.recSM_Solve(b)
    bn = np.linalg.norm(b, axis=0)
    rg = np.max(np.linalg.norm(g.residual(b, xs), axis=0) / bn)
    ro = np.max(np.linalg.norm(o.residual_dd(b, xo), axis=0) / bn)
    xor, _, _ = o.refined_solve(b, 1e-14, 3)
    g.set_precision("fp64_refined")
    g.set_refinement(1e-14, 4)
    try:
        xr = g.recSM_Solve(b)
    except HodlrNotConverged as e:  # the best iterate near the residual floor
        xr = e.x
    print(f"recSM {rel(xs, xo):.2e}, recLU {rel(xl, xo):.2e}, residual {rg:.2e} (oracle {ro:.2e}), "
          f"refined {rel(xr, xor):.2e}", flush=True)
    assert np.array_equal(g.nodes(), o.nodes())
    assert np.sum(g.nodes()[:, 4:]) > 128
    assert rel(xs, xo) < 1e-9 and rel(xl, xo) < 1e-9
    assert rg <= 4.0 * ro + 1e-15
    assert rel(xr, xor) < 1e-10
    assert abs(ld - o.logAbsDeterminant()) <= 1e-11 * abs(o.logAbsDeterminant())
