"""ORACLE -- test infrastructure only.

ctypes wrapper around oracle/build/libhodlr_oracle.so, the CPU FP64
restatement of the reference HODLR path (see hodlr_oracle.cpp for the
file:line map).  Imported only by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker; the product never loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libhodlr_oracle.so")
_lib = None

FAMILIES = {"gaussian": 0, "matern32": 1, "se": 2, "imq": 3, "quadratic": 4, "exponential": 5}

_d = C.c_double
_i = C.c_int
_p = C.c_void_p
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_exp.restype = _d
        L.orc_exp.argtypes = [_d]
        L.orc_kernel_block.argtypes = [_i, _i, _i, _dp, _d, _d, _i, _i, _i, _i, _i, _dp]
        L.orc_create_kernel.restype = _p
        L.orc_create_kernel.argtypes = [_i, _i, _i, _dp, _d, _d, _i, _i]
        L.orc_create_dense.restype = _p
        L.orc_create_dense.argtypes = [_i, _dp, _i, _i]
        L.orc_create_kernel_tree.restype = _p
        L.orc_create_kernel_tree.argtypes = [_i, _i, _i, _dp, _d, _d, _i, _i, _ip, _i]
        L.orc_create_dense_tree.restype = _p
        L.orc_create_dense_tree.argtypes = [_i, _dp, _i, _i, _ip, _i]
        L.orc_destroy.argtypes = [_p]
        L.orc_set_tolerance.argtypes = [_p, _d]
        L.orc_set_min_pivot.argtypes = [_p, _d]
        L.orc_set_lr_method.argtypes = [_p, _i]
        for f in ("orc_num_levels", "orc_num_nodes", "orc_num_leaves"):
            getattr(L, f).restype = _i
            getattr(L, f).argtypes = [_p]
        L.orc_compress.argtypes = [_p]
        L.orc_get_nodes.argtypes = [_p, _ip, _ip, _ip, _ip, _ip, _ip]
        L.orc_get_leaves.argtypes = [_p, _ip, _ip, _ip]
        L.orc_get_factor.restype = _i
        L.orc_get_factor.argtypes = [_p, _i, _i, _dp]
        L.orc_recsm_solve.argtypes = [_p, _dp, _i, _dp]
        L.orc_reclu_compute.argtypes = [_p]
        L.orc_reclu_solve.argtypes = [_p, _dp, _i, _dp]
        L.orc_logabsdet.restype = _d
        L.orc_logabsdet.argtypes = [_p, _dp]
        L.orc_matvec.argtypes = [_p, _dp, _i, _dp]
        L.orc_residual_dd.argtypes = [_p, _dp, _dp, _i, _dp]
        L.orc_set_leaf_inverse.argtypes = [_p, _i]
        L.orc_iterative_solve.restype = _i
        L.orc_iterative_solve.argtypes = [_p, _dp, _i, _dp, _i, _d, _d, _i, _i, _dp]
        L.orc_refined_solve.restype = _i
        L.orc_refined_solve.argtypes = [_p, _dp, _i, _dp, _d, _i, _dp]
        L.orc_time.restype = _d
        L.orc_time.argtypes = [_p, _i]
        L.orc_aca.restype = _i
        L.orc_aca.argtypes = [_dp, _i, _i, _i, _d, _i, _i, _i, _dp, _dp, _i, _ip, _ip, _ip]
        L.orc_svd_lowrank.restype = _i
        L.orc_svd_lowrank.argtypes = [_dp, _i, _i, _i, _d, _dp, _dp, _dp]
        L.orc_lu.argtypes = [_dp, _i, _ip]
        L.orc_dense_solve.restype = _d
        L.orc_dense_solve.argtypes = [_dp, _i, _dp, _i, _dp]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _iptr(a):
    return a.ctypes.data_as(_ip)


def hexp(x: float) -> float:
    return lib().orc_exp(float(x))


def kernel_block(points, family, length_scale, diag, i0, j0, m, n, diag_mode=0):
    pts = np.ascontiguousarray(points, dtype=np.float64)
    out = np.zeros((m, n), order="F")
    lib().orc_kernel_block(pts.shape[0], FAMILIES[family], pts.shape[1], _ptr(pts),
                           length_scale, diag, diag_mode, i0, j0, m, n, _ptr(out))
    return out


def aca(A, tol, min_rank=-1, max_rank=-1, min_pivot=0):
    """Partial-pivot ACA of a dense block; returns (W, V, rank, pivot rows, pivot cols)."""
    A = np.asfortranarray(A, dtype=np.float64)
    m, n = A.shape
    cap = min(m, n) + 2
    W = np.zeros((m, cap), order="F")
    V = np.zeros((n, cap), order="F")
    pr = np.zeros(cap, dtype=np.int32)
    pc = np.zeros(cap, dtype=np.int32)
    npv = np.zeros(1, dtype=np.int32)
    r = lib().orc_aca(_ptr(A), m, m, n, tol, min_rank, max_rank, int(min_pivot), _ptr(W), _ptr(V),
                      cap, _iptr(pr), _iptr(pc), _iptr(npv))
    assert r >= 0
    k = int(npv[0])
    return W[:, :r].copy(), V[:, :r].copy(), r, pr[:k].copy(), pc[:k].copy()


def svd_lowrank(A, tol):
    A = np.asfortranarray(A, dtype=np.float64)
    m, n = A.shape
    mn = min(m, n)
    US = np.zeros((m, mn), order="F")
    V = np.zeros((n, mn), order="F")
    s = np.zeros(mn)
    r = lib().orc_svd_lowrank(_ptr(A), m, m, n, tol, _ptr(US), _ptr(V), _ptr(s))
    return US[:, :r].copy(), V[:, :r].copy(), r, s


def lu(A):
    A = np.array(A, dtype=np.float64, order="F", copy=True)
    piv = np.zeros(A.shape[0], dtype=np.int32)
    lib().orc_lu(_ptr(A), A.shape[0], _iptr(piv))
    return A, piv


def dense_solve(A, B):
    A = np.asfortranarray(A, dtype=np.float64)
    B = np.asfortranarray(B.reshape(A.shape[0], -1), dtype=np.float64)
    X = np.zeros_like(B, order="F")
    lad = lib().orc_dense_solve(_ptr(A), A.shape[0], _ptr(B), B.shape[1], _ptr(X))
    return X, lad


class OracleHODLR:
    """CPU restatement of the reference HODLR_Matrix for the kernel and dense
    constructors (include/HODLR_Matrix.hpp:76, :162)."""

    def __init__(self, n, leaf=30, points=None, family=None, length_scale=1.0, diag=0.0,
                 diag_mode=0, dense=None, tree=None):
        """tree: user partition, rows of 8 ints (split, top/bottom min rank,
        top/bottom max rank, method, left, right), root first."""
        L = lib()
        self.n = n
        T = None if tree is None else np.ascontiguousarray(np.asarray(tree, dtype=np.int32).reshape(-1, 8))
        if dense is not None:
            A = np.asfortranarray(dense, dtype=np.float64)
            if T is None:
                self.h = L.orc_create_dense(n, _ptr(A), n, leaf)
            else:
                self.h = L.orc_create_dense_tree(n, _ptr(A), n, leaf, _iptr(T), len(T))
        else:
            self._pts = np.ascontiguousarray(points, dtype=np.float64)
            if T is None:
                self.h = L.orc_create_kernel(n, FAMILIES[family], self._pts.shape[1],
                                             _ptr(self._pts), length_scale, diag, diag_mode, leaf)
            else:
                self.h = L.orc_create_kernel_tree(n, FAMILIES[family], self._pts.shape[1],
                                                  _ptr(self._pts), length_scale, diag, diag_mode,
                                                  leaf, _iptr(T), len(T))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def set_LRTolerance(self, tol):
        lib().orc_set_tolerance(self.h, tol)

    def set_minPivot(self, mp):
        lib().orc_set_min_pivot(self.h, mp)

    def set_LRMethod(self, method):
        lib().orc_set_lr_method(self.h, {"partialPiv_ACA": 0, "SVD": 1}[method])

    def num_levels(self):
        return lib().orc_num_levels(self.h)

    def compress(self):
        lib().orc_compress(self.h)

    def nodes(self):
        """(level, min_i, split, max_i, kt, kb) per non-leaf node, level order."""
        L = lib()
        c = L.orc_num_nodes(self.h)
        arrs = [np.zeros(c, dtype=np.int32) for _ in range(6)]
        L.orc_get_nodes(self.h, *[_iptr(a) for a in arrs])
        return np.stack(arrs, axis=1)

    def leaves(self):
        L = lib()
        c = L.orc_num_leaves(self.h)
        arrs = [np.zeros(c, dtype=np.int32) for _ in range(3)]
        L.orc_get_leaves(self.h, *[_iptr(a) for a in arrs])
        return np.stack(arrs, axis=1)

    def factor(self, idx, which):
        nd = self.nodes()[idx]
        lvl, lo, sp, hi, kt, kb = nd
        rows = {0: sp - lo + 1, 1: hi - sp, 2: hi - sp, 3: sp - lo + 1}[which]
        k = kt if which < 2 else kb
        out = np.zeros((rows, max(k, 1)), order="F")
        lib().orc_get_factor(self.h, idx, which, _ptr(out))
        return out

    def recSM_Solve(self, B):
        B = np.asfortranarray(B.reshape(self.n, -1), dtype=np.float64)
        X = np.zeros_like(B, order="F")
        lib().orc_recsm_solve(self.h, _ptr(B), B.shape[1], _ptr(X))
        return X

    def recLU_Compute(self):
        lib().orc_reclu_compute(self.h)

    def recLU_Solve(self, B):
        B = np.asfortranarray(B.reshape(self.n, -1), dtype=np.float64)
        X = np.zeros_like(B, order="F")
        lib().orc_reclu_solve(self.h, _ptr(B), B.shape[1], _ptr(X))
        return X

    def logAbsDeterminant(self):
        return lib().orc_logabsdet(self.h, None)

    def determinant(self):
        d = np.zeros(1)
        lib().orc_logabsdet(self.h, _ptr(d))
        return float(d[0])

    def matvec(self, X):
        X = np.asfortranarray(X.reshape(self.n, -1), dtype=np.float64)
        Y = np.zeros_like(X, order="F")
        lib().orc_matvec(self.h, _ptr(X), X.shape[1], _ptr(Y))
        return Y

    def iterative_Solve(self, B, maxIterations=1000, stop_tolerance=1e-10, init_LRTolerance=1e-2,
                        method="partialPiv_ACA", direct="recLU"):
        """(X, iterations, final preconditioned residual); iterations -1: not converged."""
        B = np.asfortranarray(B.reshape(self.n, -1), dtype=np.float64)
        X = np.zeros_like(B, order="F")
        res = np.zeros(1)
        it = lib().orc_iterative_solve(self.h, _ptr(B), B.shape[1], _ptr(X), int(maxIterations),
                                       float(stop_tolerance), float(init_LRTolerance),
                                       {"partialPiv_ACA": 0, "SVD": 1}[method],
                                       {"recSM": 0, "recLU": 1}[direct], _ptr(res))
        return X, int(it), float(res[0])

    def set_leaf_inverse(self, on=True):
        """Measurement variant, not the reference: recSM's leaf solves multiply
        by explicitly formed inverses (as the GPU engine's leaf steps do)."""
        lib().orc_set_leaf_inverse(self.h, int(bool(on)))

    def residual_dd(self, B, X):
        """B - A_H X in double-double arithmetic, rounded once."""
        B = np.asfortranarray(B.reshape(self.n, -1), dtype=np.float64)
        X = np.asfortranarray(X.reshape(self.n, -1), dtype=np.float64)
        R = np.zeros_like(B, order="F")
        lib().orc_residual_dd(self.h, _ptr(B), _ptr(X), B.shape[1], _ptr(R))
        return R

    def refined_solve(self, B, target=1e-14, max_steps=3):
        """recSM solution refined with double-double residuals: returns
        (X, corrections made, final relative residual)."""
        B = np.asfortranarray(B.reshape(self.n, -1), dtype=np.float64)
        X = np.zeros_like(B, order="F")
        res = np.zeros(1)
        k = lib().orc_refined_solve(self.h, _ptr(B), B.shape[1], _ptr(X), float(target),
                                    int(max_steps), _ptr(res))
        return X, int(k), float(res[0])

    def dense_operator(self):
        """A_H assembled as a dense matrix (A_H e_j column by column)."""
        return self.matvec(np.eye(self.n, order="F"))

    def times(self):
        L = lib()
        return {k: L.orc_time(self.h, i) for i, k in enumerate(
            ["compress", "sm_factor", "sm_solve", "lu_factor", "lu_solve"])}
