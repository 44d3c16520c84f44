"""User-defined partitions (the reference's user_IndexTree,
include/user_IndexTree.hpp:7-38) and the kd-tree ordering that builds them
(rigidBody/MobFunctions.h:334-430).

A tree crosses the C ABI as an array of hodlr_tree_node (include/
dense_hodlr.h): one row of 8 int32 per node, root first -- split, top and
bottom minimum rank, top and bottom maximum rank, method, left, right.
user_IndexTree / UserTreeNode give the reference's pointer form in Python;
to_array() / from_array() convert.  kd_sort and kd_tree run on the host in
the library (no GPU needed).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib

FIELDS = ("split", "top_min_rank", "bottom_min_rank", "top_max_rank", "bottom_max_rank",
          "lr_method", "left", "right")
_METHODS = {None: -1, "partialPiv_ACA": 0, "SVD": 1}
_NAMES = {-1: None, 0: "partialPiv_ACA", 1: "SVD"}


class UserTreeNode:
    """user_IndexTree::node: splitIndex (last row of the left part, -1 for a
    leaf), rank bounds of the node's two off-diagonal blocks (-1: none),
    LR_Method (None: the matrix's method) and the children (both or none;
    none gives default subtrees below the split)."""

    def __init__(self, splitIndex=-1, left=None, right=None, LR_Method=None,
                 topOffDiag_minRank=-1, bottOffDiag_minRank=-1, topOffDiag_maxRank=-1,
                 bottOffDiag_maxRank=-1):
        if LR_Method not in _METHODS:
            raise ValueError(f"unsupported low-rank method {LR_Method!r}")
        self.splitIndex = int(splitIndex)
        self.left, self.right = left, right
        self.LR_Method = LR_Method
        self.topOffDiag_minRank = int(topOffDiag_minRank)
        self.bottOffDiag_minRank = int(bottOffDiag_minRank)
        self.topOffDiag_maxRank = int(topOffDiag_maxRank)
        self.bottOffDiag_maxRank = int(bottOffDiag_maxRank)


class user_IndexTree:
    """A user partition (include/user_IndexTree.hpp:7-38); rootNode None
    means the default tree."""

    def __init__(self, rootNode: UserTreeNode | None = None):
        self.rootNode = rootNode

    def to_array(self) -> np.ndarray:
        rows = []

        def visit(nd):
            me = len(rows)
            rows.append([nd.splitIndex, nd.topOffDiag_minRank, nd.bottOffDiag_minRank,
                         nd.topOffDiag_maxRank, nd.bottOffDiag_maxRank,
                         _METHODS[nd.LR_Method], -1, -1])
            if (nd.left is None) != (nd.right is None):
                raise ValueError("a user tree node has both children or none")
            if nd.left is not None:
                rows[me][6] = visit(nd.left)
                rows[me][7] = visit(nd.right)
            return me

        if self.rootNode is not None:
            visit(self.rootNode)
        return np.ascontiguousarray(np.array(rows, dtype=np.int32).reshape(-1, 8))

    @classmethod
    def from_array(cls, arr) -> "user_IndexTree":
        arr = np.asarray(arr, dtype=np.int64).reshape(-1, 8)
        if len(arr) == 0:
            return cls(None)

        def make(i):
            r = arr[i]
            nd = UserTreeNode(r[0], None, None, _NAMES[int(r[5])], r[1], r[2], r[3], r[4])
            if r[6] >= 0:
                nd.left, nd.right = make(int(r[6])), make(int(r[7]))
            return nd

        return cls(make(0))


def as_tree_array(tree) -> np.ndarray:
    """A user_IndexTree, or an array of rows of 8 ints, as the C node array."""
    if isinstance(tree, user_IndexTree):
        return tree.to_array()
    return np.ascontiguousarray(np.asarray(tree, dtype=np.int32).reshape(-1, 8))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def kd_sort(points) -> np.ndarray:
    """kd order of the points (get_KDTree_Sorted(locations, 0),
    MobFunctions.h:334-369): the locations in order are points[perm]."""
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64))
    if pts.ndim == 1:
        pts = pts.reshape(-1, 1)
    perm = np.zeros(len(pts), dtype=np.int32)
    check(lib().hodlr_kd_sort(pts.ctypes.data_as(C.POINTER(C.c_double)), len(pts), pts.shape[1],
                              _ip(perm)))
    return perm


def get_KDTree_Sorted(points, unknowns_per_point: int, rankUpperBound: int = -1,
                      LR_Method: str | None = "partialPiv_ACA"):
    """get_KDTree_Sorted(locations, pointsPerSphere, rankUpperBound,
    LR_Method) (MobFunctions.h:371-430) with unknowns_per_point = the
    reference's NDIM * pointsPerSphere: returns (perm, user_IndexTree); the
    matrix rows are ordered point by point in perm order, unknowns_per_point
    rows per point."""
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64))
    if pts.ndim == 1:
        pts = pts.reshape(-1, 1)
    n = len(pts)
    perm = np.zeros(n, dtype=np.int32)
    nodes = np.zeros((max(1, 2 * n - 1), 8), dtype=np.int32)
    count = C.c_int()
    check(lib().hodlr_kd_tree(pts.ctypes.data_as(C.POINTER(C.c_double)), n, pts.shape[1],
                              int(unknowns_per_point), int(rankUpperBound), _METHODS[LR_Method],
                              _ip(perm), _ip(nodes), C.byref(count)))
    return perm, user_IndexTree.from_array(nodes[:count.value])


def tree_of(nodes, n) -> np.ndarray:
    """The partition of a matrix as a user tree array (every node explicit)
    from its nodes() table (level, min_i, split, max_i, ...): the tree that
    rebuilds the same partition."""
    at = {(int(r[0]), int(r[1])): r for r in nodes}
    rows = []

    def visit(level, lo, hi):
        me = len(rows)
        rows.append([-1, -1, -1, -1, -1, -1, -1, -1])
        r = at.get((level, lo))
        if r is None or int(r[3]) != hi:
            return me
        sp = int(r[2])
        rows[me][0] = sp
        rows[me][6] = visit(level + 1, lo, sp)
        rows[me][7] = visit(level + 1, sp + 1, hi)
        return me

    visit(0, 0, n - 1)
    return np.ascontiguousarray(np.array(rows, dtype=np.int32))


__all__ = ["UserTreeNode", "user_IndexTree", "kd_sort", "get_KDTree_Sorted", "tree_of",
           "as_tree_array", "FIELDS"]
