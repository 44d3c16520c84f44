"""CPU tests of the user-partition and kd-ordering host code (through the C
ABI, no GPU needed) and of the oracle's restatement of userTree_To_HODLRTree
(HODLR_Tree.cpp:207-321), per-node methods and rank bounds."""
import numpy as np
import pytest

from dense_hodlr_amd import UserTreeNode, get_KDTree_Sorted, kd_sort, user_IndexTree
from dense_hodlr_amd import inputs as I
from dense_hodlr_amd.user_tree import tree_of


def test_kd_sort_matches_an_independent_kd_order():
    """Without equal coordinates the merge sort's tie rule never applies, so
    the order equals inputs.kd_order's (numpy's stable sort per level)."""
    for n, seed in ((1000, 3), (777, 4), (4096, 5)):
        p = I.SplitMix64(seed).uniform(2 * n).reshape(n, 2)
        assert np.array_equal(kd_sort(p), I.kd_order(p))


def test_kd_sort_tie_rule_hand_traced():
    """Equal keys take the right list's point first (MobFunctions.h:274-281),
    and each half is sorted again one depth further.  x = [1, 1, 0]: the top
    sort gives [2, 1, 0] (the equal points 0 and 1 swap: right first), the
    right part [1, 0] is sorted again and swaps back: [2, 0, 1]."""
    assert kd_sort(np.array([[1.0], [1.0], [0.0]])).tolist() == [2, 0, 1]
    # a lattice: every level has ties; still a permutation that is a kd tree
    g = np.array([(i, j) for i in range(8) for j in range(8)], dtype=float)
    perm = kd_sort(g)
    assert sorted(perm.tolist()) == list(range(64))
    q = g[perm]
    assert q[:32, 0].max() <= q[32:, 0].min()


def test_kd_tree_splits_hand_traced():
    """Four points, 3 unknowns each (the reference's NDIM * pointsPerSphere):
    root split 3 * 2 - 1 = 5, left child 3 * 1 - 1 = 2, right child
    3 * (2 + 1) - 1 = 8; single points are leaves; every node carries the rank
    bound and the method."""
    perm, t = get_KDTree_Sorted(np.array([[0.0], [1.0], [2.0], [3.0]]), 3, 10)
    a = t.to_array()
    assert perm.tolist() == [0, 1, 2, 3]
    assert len(a) == 7
    assert [r[0] for r in a] == [5, 2, -1, -1, 8, -1, -1]
    assert all(r[3] == 10 and r[4] == 10 and r[5] == 0 for r in a)
    assert t.rootNode.left.splitIndex == 2 and t.rootNode.right.splitIndex == 8


def test_tree_array_round_trip():
    t = user_IndexTree(UserTreeNode(99, UserTreeNode(49), UserTreeNode(-1), "SVD", 2, 3, 40, 50))
    a = t.to_array()
    b = user_IndexTree.from_array(a).to_array()
    assert np.array_equal(a, b)
    assert a[0].tolist() == [99, 2, 3, 40, 50, 1, 1, 2]
    with pytest.raises(ValueError):
        user_IndexTree(UserTreeNode(5, UserTreeNode(2), None)).to_array()


def test_oracle_user_tree_equal_to_default(oracle):
    """The default tree written out as a user tree builds the same partition,
    ranks and solution."""
    n, leaf = 1000, 30
    pts = I.points_1d_sorted(n, 7)
    b = I.rhs_uniform(n, 1, 8)
    o = oracle.OracleHODLR(n, leaf=leaf, points=pts, family="gaussian", length_scale=0.2, diag=0.1)
    o.set_LRTolerance(1e-10)
    x = o.recSM_Solve(b)
    t = tree_of(o.nodes(), n)
    u = oracle.OracleHODLR(n, leaf=leaf, points=pts, family="gaussian", length_scale=0.2, diag=0.1,
                           tree=t)
    u.set_LRTolerance(1e-10)
    xu = u.recSM_Solve(b)
    assert np.array_equal(u.nodes(), o.nodes())
    assert np.array_equal(xu, x)


def test_oracle_user_node_default_children_hand_traced(oracle):
    """A user root split at 5 with no children over 13 rows, leaf 2: default
    nodes below it, the left one [0, 5] split at 0 + 6/2 - 1 = 2, the right
    one [6, 12] split at 5 + 7/2 = 8 (the reference's rule for a default
    right node under a user node; the default tree would split it at
    12 - 3 = 9)."""
    n = 13
    pts = I.points_1d_sorted(n, 9)
    u = oracle.OracleHODLR(n, leaf=2, points=pts, family="gaussian", length_scale=0.3, diag=1.0,
                           tree=[[5, -1, -1, -1, -1, -1, -1, -1]])
    nd = u.nodes()
    assert nd[0, :4].tolist() == [0, 0, 5, 12]
    lvl1 = [r[:4].tolist() for r in nd if r[0] == 1]
    assert lvl1 == [[1, 0, 2, 5], [1, 6, 8, 12]]


def test_oracle_rank_bounds_and_methods(oracle):
    """Per-node rank bounds and methods: a maximum rank caps the ACA rank
    (lowRank.cpp:246-247), 0 stores one zero column (:120-125), a minimum
    rank forces more terms (:149); an SVD node's blocks are the SVD of those
    blocks while the rest use ACA."""
    n, leaf = 512, 32
    pts = I.points_1d_sorted(n, 10)
    K = oracle.kernel_block(pts, "gaussian", 0.1, 1e-2, 0, 0, n, n)
    free = oracle.OracleHODLR(n, leaf=leaf, dense=K)
    free.set_LRTolerance(1e-10)
    free.compress()
    r0 = free.nodes()[0]
    # root: top block at most 5 terms, bottom block exactly zero; no children -> default below
    t = [[255, -1, -1, 5, 0, 0, -1, -1]]
    o = oracle.OracleHODLR(n, leaf=leaf, dense=K, tree=t)
    o.set_LRTolerance(1e-10)
    o.compress()
    nd = o.nodes()
    assert nd[0, 4] == min(5, r0[4]) and nd[0, 5] == 1
    assert not np.any(o.factor(0, 2)) and not np.any(o.factor(0, 3))
    assert np.array_equal(nd[1:, 4:], free.nodes()[1:, 4:])
    # minimum rank 40 on the top block
    o2 = oracle.OracleHODLR(n, leaf=leaf, dense=K, tree=[[255, 40, -1, -1, -1, 0, -1, -1]])
    o2.set_LRTolerance(1e-10)
    o2.compress()
    assert o2.nodes()[0, 4] == max(40, r0[4])
    # SVD at the root, ACA below (method -1 = the matrix's)
    o3 = oracle.OracleHODLR(n, leaf=leaf, dense=K, tree=[[255, -1, -1, -1, -1, 1, -1, -1]])
    o3.set_LRTolerance(1e-10)
    o3.compress()
    US, V, r, _ = oracle.svd_lowrank(K[:256, 256:], 1e-10)
    assert o3.nodes()[0, 4] == r
    assert np.allclose(o3.factor(0, 0) @ o3.factor(0, 1).T, US @ V.T, rtol=0, atol=1e-14)
    USb, Vb, rb, _ = oracle.svd_lowrank(K[256:, :256], 1e-10)  # the node's method: both blocks
    assert o3.nodes()[0, 5] == rb
    assert np.allclose(o3.factor(0, 2) @ o3.factor(0, 3).T, USb @ Vb.T, rtol=0, atol=1e-14)
    # the children keep ACA: their factors are the standalone ACA's, bit for bit
    assert np.array_equal(o3.nodes()[1:, 4:], free.nodes()[1:, 4:])
    W, V, r1, _, _ = oracle.aca(K[:128, 128:256], 1e-10)
    assert o3.nodes()[1, 4] == r1 and np.array_equal(o3.factor(1, 0), W)
