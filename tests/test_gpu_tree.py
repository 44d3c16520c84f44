"""GPU tests of user partitions (kd trees, per-node rank bounds and methods),
entry access (block / row / col) and tree surgery (splitAtTop, topDiag,
bottDiag), each against the oracle on the same inputs and tree."""
import sys

import numpy as np
import pytest

from dense_hodlr_amd import (HODLR_Batch, HODLR_Matrix, HodlrError, UserTreeNode,
                             get_KDTree_Sorted, user_IndexTree)
from dense_hodlr_amd import inputs as I
from dense_hodlr_amd.user_tree import tree_of

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _clusters(ncent, per, seed):
    """ncent cluster centres in the unit square, per points around each, in
    kd order of the centres (each cluster's points contiguous)."""
    r = I.SplitMix64(seed)
    cent = r.uniform(2 * ncent).reshape(ncent, 2)
    off = 0.01 * r.normal(2 * ncent * per).reshape(ncent, per, 2)
    return cent, off


def test_kd_partition_matches_oracle(oracle):
    """get_KDTree_Sorted over 64 cluster centres with 32 unknowns each (leaf
    32, rank bound 40 on every block): the GPU builds the oracle's partition,
    its ranks (all within the bound) and its recSM / recLU solutions and
    log-determinant."""
    _case("kd partition 64 x 32")
    cent, off = _clusters(64, 32, 5)
    perm, tree = get_KDTree_Sorted(cent, 32, 40, "partialPiv_ACA")
    pts = np.ascontiguousarray((cent[perm][:, None, :] + off[perm]).reshape(-1, 2))
    n = len(pts)
    b = I.rhs_normal(n, 2, 6)
    g = HODLR_Matrix.from_kernel(pts, "gaussian", 0.2, 1e-2, sizeThreshold=32, indexTree=tree)
    g.set_LRTolerance(1e-10)
    o = oracle.OracleHODLR(n, leaf=32, points=pts, family="gaussian", length_scale=0.2, diag=1e-2,
                           tree=tree.to_array())
    o.set_LRTolerance(1e-10)
    xs, xo = g.recSM_Solve(b), o.recSM_Solve(b)
    gn = g.nodes()
    assert np.array_equal(gn, o.nodes())
    assert gn[:, 4:].max() <= 40
    assert np.any(gn[:, 4:] == 40)  # the bound is active somewhere
    assert rel(xs, xo) < 1e-10
    assert rel(g.recLU_Solve(b), o.recLU_Solve(b)) < 1e-10
    assert abs(g.logAbsDeterminant() - o.logAbsDeterminant()) < 1e-12 * abs(o.logAbsDeterminant())


def test_rank_bounds_and_methods_match_oracle(oracle):
    """A user root with a maximum rank 5 (top) and 0 (bottom: a zero block),
    another with a minimum rank 40, and one whose method is SVD while the
    default nodes below use ACA: ranks and factors as the oracle's (ACA bit
    for bit, SVD products to rounding), solutions to 1e-10."""
    n, leaf = 512, 32
    pts = I.points_1d_sorted(n, 10)
    K = oracle.kernel_block(pts, "gaussian", 0.1, 1e-2, 0, 0, n, n)
    b = I.rhs_uniform(n, 2, 11)
    for t in ([[255, -1, -1, 5, 0, 0, -1, -1]], [[255, 40, -1, -1, -1, 0, -1, -1]],
              [[255, -1, -1, -1, -1, 1, -1, -1]]):
        _case(f"rank bounds / methods {t}")
        g = HODLR_Matrix(K, leaf, indexTree=t)
        g.set_LRTolerance(1e-10)
        o = oracle.OracleHODLR(n, leaf=leaf, dense=K, tree=t)
        o.set_LRTolerance(1e-10)
        xg, xo = g.recSM_Solve(b), o.recSM_Solve(b)
        assert np.array_equal(g.nodes(), o.nodes())
        for idx in range(len(o.nodes())):
            for wu, wv in ((0, 1), (2, 3)):
                if idx == 0 and t[0][5] == 1:  # SVD: the products
                    Pg = g.factor(idx, wu) @ g.factor(idx, wv).T
                    Po = o.factor(idx, wu) @ o.factor(idx, wv).T
                    assert np.max(np.abs(Pg - Po)) <= 1e-12 * np.max(np.abs(Po))
                else:
                    assert np.array_equal(g.factor(idx, wu), o.factor(idx, wu)), (idx, wu)
                    assert np.array_equal(g.factor(idx, wv), o.factor(idx, wv)), (idx, wv)
        assert rel(xg, xo) < 1e-10


def test_set_lr_method_overrides_user_methods(oracle):
    """set_LRMethod("SVD") sets every node's method, the user partition's
    included (HODLR_Tree.cpp:377-390); set_LRMethod back to the default
    changes nothing for nodes that carry their own method."""
    _case("set_LRMethod on a user partition")
    n, leaf = 400, 40
    pts = I.points_1d_sorted(n, 12)
    t = [[199, -1, -1, -1, -1, 0, -1, -1]]
    b = I.rhs_uniform(n, 1, 13)
    g = HODLR_Matrix.from_kernel(pts, "gaussian", 0.2, 1e-2, sizeThreshold=leaf, indexTree=t)
    o = oracle.OracleHODLR(n, leaf=leaf, points=pts, family="gaussian", length_scale=0.2,
                           diag=1e-2, tree=t)
    for h in (g, o):
        h.set_LRTolerance(1e-10)
        h.set_LRMethod("SVD")
    assert rel(g.recSM_Solve(b), o.recSM_Solve(b)) < 1e-10
    assert np.array_equal(g.nodes(), o.nodes())


def test_user_tree_errors():
    _case("user tree errors")
    A = np.eye(100) + 0.01
    with pytest.raises(HodlrError) as e:
        HODLR_Matrix(A, 10, indexTree=[[120, -1, -1, -1, -1, 0, -1, -1]])  # split outside
    assert e.value.code == -1
    with pytest.raises(HodlrError) as e:  # a node of 100 rows needs a split
        HODLR_Matrix(A, 10, indexTree=[[-1, -1, -1, -1, -1, 0, -1, -1]])
    assert e.value.code == -1 and "split" in str(e.value)
    with pytest.raises(ValueError):
        user_IndexTree(UserTreeNode(49, UserTreeNode(20), None)).to_array()


@pytest.mark.parametrize("kind", ["dense", "kernel"])
def test_split_at_top_matches_oracle_of_the_blocks(oracle, kind):
    """splitAtTop / topDiag / bottDiag keep each diagonal block's tree part
    and factors: the halves have the oracle's partition and ranks for that
    block (built from the block's data with the same partition), equal
    solutions (1e-10) and, entry by entry, the parent's compressed
    operator restricted to the block.  n = 3001 makes the halves 1500 and
    1501 rows: the bottom half's splits follow the right-child rule, not a
    default tree's."""
    n, leaf = 3001, 64
    pts = I.points_1d_sorted(n, 14)
    if kind == "dense":
        A = oracle.kernel_block(pts, "imq", 0.3, 1.0, 0, 0, n, n)
        g = HODLR_Matrix(A, leaf)
    else:
        g = HODLR_Matrix.from_kernel(pts, "gaussian", 0.1, 1e-2, sizeThreshold=leaf)
    g.set_LRTolerance(1e-10)
    top, bot = g.splitAtTop()
    sp = int(g.nodes()[0, 2])
    for side, h in ((0, top), (1, bot)):
        lo, hi = (0, sp) if side == 0 else (sp + 1, n - 1)
        s = hi - lo + 1
        _case(f"split {kind} side {side}: {s} rows")
        assert h.rows() == s
        t = tree_of(h.nodes(), s)
        if kind == "dense":
            o = oracle.OracleHODLR(s, leaf=leaf, dense=A[lo:hi + 1, lo:hi + 1], tree=t)
        else:
            o = oracle.OracleHODLR(s, leaf=leaf, points=pts[lo:hi + 1], family="gaussian",
                                   length_scale=0.1, diag=1e-2, tree=t)
        o.set_LRTolerance(1e-10)
        b = I.rhs_uniform(s, 2, 15 + side)
        xh, xo = h.recSM_Solve(b), o.recSM_Solve(b)
        assert np.array_equal(h.nodes(), o.nodes())
        assert rel(xh, xo) < 1e-10
        Bh = h.block_compressed(0, 0, s, s)
        Bg = g.block_compressed(lo, lo, s, s)
        assert np.array_equal(Bh, Bg)
        # topDiag / bottDiag give the same halves
        d = g.topDiag() if side == 0 else g.bottDiag()
        assert np.array_equal(d.nodes(), h.nodes()) and np.array_equal(d.recSM_Solve(b), xh)
    # the parent is unchanged
    assert g.rows() == n
    _case("split of a leaf root")
    small = HODLR_Matrix.from_kernel(pts[:50], "gaussian", 0.1, 1e-2, sizeThreshold=64)
    with pytest.raises(HodlrError) as e:
        small.splitAtTop()
    assert e.value.code == -2


def test_block_row_col(oracle):
    """block() returns the source entries (the kernel on the device: the
    oracle's bit for bit, numpy's closed form to a few ulp; the dense copy
    exactly), block_compressed the operator A_H (the oracle's assembled
    operator to rounding); row and col are 1-row and 1-column blocks; a
    batch is block-diagonal; blocks outside the matrix are rejected."""
    _case("block / row / col")
    n = 1500
    pts = I.points_1d_sorted(n, 16)
    g = HODLR_Matrix.from_kernel(pts, "gaussian", 0.1, 1e-2, sizeThreshold=64)
    g.set_LRTolerance(1e-10)
    B = g.block(100, 700, 300, 500)
    Bo = oracle.kernel_block(pts, "gaussian", 0.1, 1e-2, 100, 700, 300, 500)
    x = pts[:, 0]
    arg = ((x[100:400, None] - x[None, 700:1200]) / 0.1) ** 2
    Bn = np.exp(-arg)
    assert np.array_equal(B, Bo)
    # the exponent's own rounding (a few ulp of |arg|) carries into exp
    eps = np.finfo(float).eps
    assert np.all(np.abs(B - Bn) <= 8 * eps * (1 + arg) * Bn + 1e-300)
    assert np.array_equal(g.row(123), oracle.kernel_block(pts, "gaussian", 0.1, 1e-2, 123, 0, 1, n))
    assert np.array_equal(g.col(77), oracle.kernel_block(pts, "gaussian", 0.1, 1e-2, 0, 77, n, 1))
    o = oracle.OracleHODLR(n, leaf=64, points=pts, family="gaussian", length_scale=0.1, diag=1e-2)
    o.set_LRTolerance(1e-10)
    AH = o.dense_operator()
    C = g.block_compressed(0, 0, n, n)
    assert np.max(np.abs(C - AH)) <= 1e-14 * np.max(np.abs(AH))
    assert np.max(np.abs(C - oracle.kernel_block(pts, "gaussian", 0.1, 1e-2, 0, 0, n, n))) < 1e-8
    # test/test.cpp:400-413 on a dense exact-HODLR matrix: full, part, row, col
    m = 2000
    E = I.exact_hodlr(m, 30, 10, seed=17)
    h = HODLR_Matrix(E, 30)
    assert np.array_equal(h.block(0, 0, m, m), E)
    assert np.array_equal(h.block(m // 4, m // 4, m // 2, m // 2),
                          E[m // 4:m // 4 + m // 2, m // 4:m // 4 + m // 2])
    assert np.array_equal(h.row(50), E[50:51, :]) and np.array_equal(h.col(1532), E[:, 1532:1533])
    h.set_LRTolerance(1e-12)
    assert np.max(np.abs(h.block_compressed(0, 0, m, m) - E)) < 1e-12
    # a batch: zero between instances
    bp = [I.points_1d_sorted(k, 18 + k) for k in (300, 500)]
    bt = HODLR_Batch(bp, "se", 0.1, 1e-2, sizeThreshold=64)
    F = bt.block(0, 0, 800, 800)
    assert not np.any(F[:300, 300:]) and not np.any(F[300:, :300])
    assert np.array_equal(F[300:, 300:], oracle.kernel_block(bp[1], "se", 0.1, 1e-2, 0, 0, 500, 500))
    with pytest.raises(HodlrError) as e:
        g.block(1400, 0, 200, 10)
    assert e.value.code == -1
