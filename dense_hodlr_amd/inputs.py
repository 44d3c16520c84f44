"""Portable synthetic inputs for the HODLR engine.

Every random quantity comes from one splitmix64 stream (Steele, Lea, Flood,
"Fast splittable pseudorandom number generators", OOPSLA 2014): uniforms take
the top 53 bits of each output, normals use the Box-Muller transform.  The
reference's tests draw from unseeded ``std::rand`` through
``Eigen::MatrixXd::Random`` (src/HODLR_Matrix.cpp:456, :471-474), which is not
portable, so the engine's tests and benchmarks use this generator instead.

Point orderings follow the reference:
  * 1D: sorted uniform points (the 1D builders, src/helperFunctions.cpp:73-126);
  * 2D: kd order of rigidBody/MobFunctions.h:334-369 -- sort the segment by
    coordinate ``depth % dim``, give the left child ``n // 2`` points, recurse.
"""
from __future__ import annotations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    """Sequential splitmix64 stream; ``take(n)`` returns the next n outputs."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def take(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            k = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + k * _GAMMA
            self.state = self.state + np.uint64(n) * _GAMMA
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            return z ^ (z >> np.uint64(31))

    def uniform(self, n: int) -> np.ndarray:
        """n doubles in [0, 1): top 53 bits times 2^-53."""
        return (self.take(n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)

    def uniform_pm1(self, n: int) -> np.ndarray:
        """n doubles in [-1, 1) (the range of Eigen's Random)."""
        return 2.0 * self.uniform(n) - 1.0

    def normal(self, n: int) -> np.ndarray:
        """n standard normals by Box-Muller on pairs of uniforms."""
        m = (n + 1) // 2
        u = self.uniform(2 * m)
        u1 = 1.0 - u[0::2]  # (0, 1]
        u2 = u[1::2]
        rad = np.sqrt(-2.0 * np.log(u1))
        z = np.empty(2 * m)
        z[0::2] = rad * np.cos(2.0 * np.pi * u2)
        z[1::2] = rad * np.sin(2.0 * np.pi * u2)
        return z[:n]


def points_1d_sorted(n: int, seed: int) -> np.ndarray:
    """n uniform points in [0,1), sorted ascending; shape (n, 1)."""
    return np.sort(SplitMix64(seed).uniform(n)).reshape(n, 1)


def kd_order(pts: np.ndarray) -> np.ndarray:
    """Permutation putting points in kd order (rigidBody/MobFunctions.h:334-369).

    Each segment is sorted by coordinate ``depth % dim`` and split with the
    left part holding ``len // 2`` points.  The reference's merge
    (MobFunctions.h:263-289) takes the right list first on equal keys; ties
    have probability zero for continuous random points, and a stable sort is
    used here.
    """
    n, dim = pts.shape
    perm = np.arange(n)
    segs = [(0, n)]
    depth = 0
    while segs:
        coord = pts[perm, depth % dim]
        seg_id = np.full(n, -1, dtype=np.int64)
        nxt = []
        for s, e in segs:
            seg_id[s:e] = s
            if e - s > 1:
                h = (e - s) // 2
                nxt.append((s, s + h))
                nxt.append((s + h, e))
        active = seg_id >= 0
        idx = np.nonzero(active)[0]
        order = np.lexsort((coord[idx], seg_id[idx]))
        perm[idx] = perm[idx][order]
        segs = [(s, e) for (s, e) in nxt if e - s > 1]
        depth += 1
    return perm


def points_2d_kd(n: int, seed: int) -> np.ndarray:
    """n uniform points in the unit square (point-major), in kd order."""
    p = SplitMix64(seed).uniform(2 * n).reshape(n, 2)
    return np.ascontiguousarray(p[kd_order(p)])


def rhs_uniform(n: int, nrhs: int, seed: int) -> np.ndarray:
    """n x nrhs right-hand side, entries uniform in [-1,1), column-major."""
    return np.asfortranarray(SplitMix64(seed).uniform_pm1(n * nrhs).reshape(nrhs, n).T)


def rhs_normal(n: int, nrhs: int, seed: int) -> np.ndarray:
    """n x nrhs right-hand side, entries N(0,1), column-major."""
    return np.asfortranarray(SplitMix64(seed).normal(n * nrhs).reshape(nrhs, n).T)


def linspace_m2_2(n: int) -> np.ndarray:
    """The reference tests' exact solution LinSpaced(n, -2, 2) (test/test.cpp:62)."""
    if n == 1:
        return np.array([-2.0])
    return -2.0 + 4.0 * np.arange(n) / (n - 1)


def default_tree(n: int, leaf: int):
    """Default HODLR tree (src/HODLR_Tree.cpp:33-132).

    Returns (nodes, leaves): nodes is a list of (level, min_i, split, max_i)
    for non-leaf nodes, level by level and left to right; leaves is a list of
    (level, min_i, max_i) left to right.
    """
    nodes, leaves = [], []
    if n < leaf:
        return nodes, [(0, 0, n - 1)]
    stack = [(0, 0, n // 2 - 1, n - 1)]
    allnodes = []
    while stack:
        lvl, lo, sp, hi = stack.pop()
        allnodes.append((lvl, lo, sp, hi))
        ts, bs = sp - lo + 1, hi - sp
        # right child first on the stack so the left is expanded first
        if bs > leaf:
            stack.append((lvl + 1, sp + 1, hi - bs // 2, hi))
        else:
            leaves.append((lvl + 1, sp + 1, hi))
        if ts > leaf:
            stack.append((lvl + 1, lo, lo + ts // 2 - 1, sp))
        else:
            leaves.append((lvl + 1, lo, sp))
    nodes = sorted(allnodes, key=lambda t: (t[0], t[1]))
    leaves = sorted(leaves, key=lambda t: t[1])
    return nodes, leaves


def exact_hodlr(n: int, leaf: int, rank: int, seed: int) -> np.ndarray:
    """Dense exact-HODLR matrix (createExactHODLR, src/HODLR_Matrix.cpp:432-502).

    Leaves and the U, V factors of every off-diagonal block are uniform in
    [-1, 1); values are drawn in preorder (leaf block, or top U, top V, bottom
    U, bottom V, then left and right subtrees), each matrix column-major.
    """
    rng = SplitMix64(seed)
    A = np.zeros((n, n), order="F")

    def build(lvl, lo, sp, hi):
        ts, bs = sp - lo + 1, hi - sp
        if rank > 0:
            tU = rng.uniform_pm1(ts * rank).reshape(rank, ts).T
            tV = rng.uniform_pm1(bs * rank).reshape(rank, bs).T
            bU = rng.uniform_pm1(bs * rank).reshape(rank, bs).T
            bV = rng.uniform_pm1(ts * rank).reshape(rank, ts).T
            A[lo:sp + 1, sp + 1:hi + 1] = tU @ tV.T
            A[sp + 1:hi + 1, lo:sp + 1] = bU @ bV.T
        if ts > leaf:
            build(lvl + 1, lo, lo + ts // 2 - 1, sp)
        else:
            A[lo:sp + 1, lo:sp + 1] = rng.uniform_pm1(ts * ts).reshape(ts, ts).T
        if bs > leaf:
            build(lvl + 1, sp + 1, hi - bs // 2, hi)
        else:
            A[sp + 1:hi + 1, sp + 1:hi + 1] = rng.uniform_pm1(bs * bs).reshape(bs, bs).T

    if n < leaf:
        A[:, :] = rng.uniform_pm1(n * n).reshape(n, n).T
    else:
        build(0, 0, n // 2 - 1, n - 1)
    return A


def config_inputs(name: str, seed: int = 12345):
    """Inputs of the named workloads (points, kernel, leaf, tol, rhs)."""
    if name == "gauss1d_4096":
        n = 4096
        return dict(n=n, points=points_1d_sorted(n, seed), family="gaussian", length_scale=0.1,
                    diag=1e-2, leaf=64, tol=1e-10, rhs=rhs_uniform(n, 1, seed + 1))
    if name == "matern2d_65536":
        n = 65536
        return dict(n=n, points=points_2d_kd(n, seed), family="matern32", length_scale=0.2,
                    diag=1e-3, leaf=128, tol=1e-10, rhs=rhs_normal(n, 16, seed + 1))
    if name == "gauss2d_262144":
        n = 262144
        return dict(n=n, points=points_2d_kd(n, seed), family="gaussian", length_scale=0.05,
                    diag=1e-4, leaf=256, tol=1e-10, rhs=rhs_normal(n, 64, seed + 1))
    if name == "gp_batch_512":
        return batch_inputs(512, 1024, 16384, seed)
    raise KeyError(name)


def batch_inputs(count: int, nmin: int, nmax: int, seed: int = 12345):
    """Independent Gaussian-process covariance systems: sizes log-uniform in
    [nmin, nmax], 1D sorted uniform points, squared-exponential kernel
    exp(-r^2 / (2 l^2)) with l uniform in [0.01, 0.2], noise variance 1e-2 on
    the diagonal, one N(0,1) right-hand side each, leaf 64, ACA tolerance
    1e-10.  One generator draws, per instance in order: the size, the length
    scale, then (from seeds derived per instance) points and right-hand side."""
    g = SplitMix64(seed)
    ns, ls, pts, rhs = [], [], [], []
    lo, hi = np.log(nmin), np.log(nmax)
    for q in range(count):
        n = int(np.floor(np.exp(lo + (hi - lo) * g.uniform(1)[0])))
        n = min(max(n, nmin), nmax)
        ell = float(0.01 + 0.19 * g.uniform(1)[0])
        ns.append(n)
        ls.append(ell)
        pts.append(points_1d_sorted(n, seed * 7919 + 2 * q + 1))
        rhs.append(rhs_normal(n, 1, seed * 7919 + 2 * q + 2))
    return dict(count=count, n=ns, length_scale=ls, points=pts, family="se", diag=1e-2, leaf=64,
                tol=1e-10, rhs=rhs)
