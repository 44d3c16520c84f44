"""Host-side mirror of the reference's HODLR_Matrix (include/HODLR_Matrix.hpp:38-413).

Same method names and argument meaning as the reference class for the
kernel/dense constructors, set_LRMethod / set_LRTolerance / set_minPivot,
recSM_Solve, recLU_Solve, recLU_Compute, logAbsDeterminant and determinant.
Where the reference prints "Error! ..." and exits, or asserts, this class
raises HodlrError (or ValueError for argument checks made on the host).
Everything runs through the C ABI of libdense_hodlr_amd.so on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import HodlrError, HodlrNotConverged, check, lib
from .user_tree import as_tree_array, tree_of


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _as_f64_2d(B, n):
    B = np.asarray(B, dtype=np.float64)
    if B.ndim == 1:
        B = B.reshape(-1, 1)
    if B.shape[0] != n:
        # the reference asserts input_RHS.rows() == matrixSize (HODLR_Matrix.cpp:806-807)
        raise ValueError(f"right-hand side has {B.shape[0]} rows, matrix has {n}")
    return np.asfortranarray(B)


class HODLR_Matrix:
    """HODLR matrix on one MI355X.

    HODLR_Matrix(matrix, sizeThreshold=30, LR_Method="partialPiv_ACA") is the
    dense constructor (HODLR_Matrix.hpp:76); use from_kernel() for the typed
    kernel constructor and from_callback() for the host-callback one
    (HODLR_Matrix.hpp:162).
    """

    def __init__(self, matrix=None, sizeThreshold: int = 30, LR_Method: str = "partialPiv_ACA",
                 *, indexTree=None, _handle=None, _n=None):
        """indexTree: a user partition (user_IndexTree, or its node array), the
        constructor HODLR_Matrix(MatrixXd&, int, user_IndexTree&)
        (HODLR_Matrix.hpp:85)."""
        self._h = None
        if _handle is not None:
            self._h = _handle
            self._n = _n
            return
        A = np.asfortranarray(np.asarray(matrix, dtype=np.float64))
        if A.ndim != 2 or A.shape[0] != A.shape[1]:
            # HODLR_Matrix.cpp:92: assert(isSquareMatrix == true)
            raise ValueError("the HODLR matrix must be square")
        if LR_Method not in _lib.LR_METHODS:
            raise ValueError(f"Invalid low-rank approximation scheme ({LR_Method})")
        h = C.c_void_p()
        if indexTree is None:
            check(lib().hodlr_create_dense(C.byref(h), A.shape[0], _dptr(A), A.shape[0],
                                           int(sizeThreshold), _lib.LR_METHODS[LR_Method]))
        else:
            T = as_tree_array(indexTree)
            check(lib().hodlr_create_dense_tree(C.byref(h), A.shape[0], _dptr(A), A.shape[0],
                                                int(sizeThreshold), _iptr(T), len(T)))
            self._h = h
            self._n = A.shape[0]
            if LR_Method != "partialPiv_ACA":
                self.set_LRMethod(LR_Method)
            return
        self._h = h
        self._n = A.shape[0]

    @classmethod
    def from_kernel(cls, points, kernel: str = "gaussian", length_scale: float = 1.0,
                    diag: float = 0.0, diag_mode: str = "add", sizeThreshold: int = 30,
                    indexTree=None):
        """indexTree: a user partition, the constructor with a
        user_IndexTree& (HODLR_Matrix.hpp:184)."""
        pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64))
        if pts.ndim == 1:
            pts = pts.reshape(-1, 1)
        n, dim = pts.shape
        kd = _lib.KernelDesc(_lib.KERNELS[kernel], dim, _dptr(pts), float(length_scale),
                             float(diag), {"add": 0, "replace": 1}[diag_mode])
        h = C.c_void_p()
        if indexTree is None:
            check(lib().hodlr_create_kernel(C.byref(h), n, C.byref(kd), int(sizeThreshold)))
        else:
            T = as_tree_array(indexTree)
            check(lib().hodlr_create_kernel_tree(C.byref(h), n, C.byref(kd), int(sizeThreshold),
                                                 _iptr(T), len(T)))
        return cls(_handle=h, _n=n)

    @classmethod
    def from_binary_file(cls, path, sizeThreshold: int = 30, LR_Method: str = "partialPiv_ACA"):
        """HODLR_Matrix(readBinaryIntoMatrixXd(path), sizeThreshold, LR_Method), the
        reference drivers' way to build from a file (helperFunctions.cpp:370-371),
        with the file streamed straight into the handle's copy."""
        import os
        if LR_Method not in _lib.LR_METHODS:
            raise ValueError(f"Invalid low-rank approximation scheme ({LR_Method})")
        h = C.c_void_p()
        check(lib().hodlr_create_dense_from_file(C.byref(h), os.fsencode(path), int(sizeThreshold),
                                                 _lib.LR_METHODS[LR_Method]))
        return cls(_handle=h, _n=lib().hodlr_rows(h))

    @classmethod
    def from_callback(cls, n: int, fn, data=None, sizeThreshold: int = 30):
        """fn(i, j) -> float, evaluated on the host for every entry."""
        cb = _lib.CALLBACK(lambda i, j, _p: float(fn(i, j)))
        h = C.c_void_p()
        check(lib().hodlr_create_kernel_callback(C.byref(h), int(n), cb, None, int(sizeThreshold)))
        return cls(_handle=h, _n=n)

    def __del__(self):
        if getattr(self, "_h", None) is not None and _lib._lib is not None:
            _lib._lib.hodlr_destroy(self._h)
            self._h = None

    # ---- attribute modification (HODLR_Matrix.hpp:222-230) -------------
    def set_LRTolerance(self, tolerance: float):
        check(lib().hodlr_set_tolerance(self._h, float(tolerance)), self._h)

    def set_minPivot(self, minPivot: float):
        check(lib().hodlr_set_min_pivot(self._h, float(minPivot)), self._h)

    def set_LRMethod(self, LR_Method: str):
        if LR_Method not in _lib.LR_METHODS:
            raise ValueError(f"Invalid low-rank approximation scheme ({LR_Method})")
        check(lib().hodlr_set_lr_method(self._h, _lib.LR_METHODS[LR_Method]), self._h)

    def set_precision(self, precision: str):
        check(lib().hodlr_set_precision(self._h, _lib.PRECISIONS[precision]), self._h)

    def set_profiling(self, on: bool = True):
        check(lib().hodlr_set_profiling(self._h, int(bool(on))), self._h)

    def set_rank_predictor(self, model=None, margin: float = 0.25):
        """Size compression from a RankModel's predictions (None detaches)."""
        check(lib().hodlr_set_rank_predictor(self._h, model._h if model is not None else None,
                                             float(margin)), self._h)

    def autotune(self):
        """Pick launch shapes by timing candidates on this problem."""
        check(lib().hodlr_autotune(self._h), self._h)

    def set_launch_shapes(self, aca_rows_per_thread=2, gemm_tile_m=0, gemm_split_target=2048,
                          lu_panel=64, gemm_mfma=True, leaf_solve=3):
        """Fix the launch shapes (see hodlr_set_launch_shapes)."""
        check(lib().hodlr_set_launch_shapes(self._h, int(aca_rows_per_thread), int(gemm_tile_m),
                                            int(gemm_split_target), int(lu_panel),
                                            int(bool(gemm_mfma)), int(leaf_solve)),
              self._h)

    def set_graphs(self, on: bool = True):
        """Replay FP64 solves from captured hipGraphs."""
        check(lib().hodlr_set_graphs(self._h, int(bool(on))), self._h)

    def set_refinement(self, target: float = 1e-10, max_iters: int = 30):
        check(lib().hodlr_set_refinement(self._h, float(target), int(max_iters)), self._h)

    # ---- compute ---------------------------------------------------------
    def storeLRinTree(self):
        check(lib().hodlr_compress(self._h), self._h)

    def recSM_Factorize(self):
        check(lib().hodlr_factorize(self._h, _lib.FACTOR_SM), self._h)

    def recLU_Compute(self):
        check(lib().hodlr_factorize(self._h, _lib.FACTOR_LU), self._h)

    def _solve(self, fn, B):
        B = _as_f64_2d(B, self._n)
        X = np.zeros_like(B, order="F")
        try:
            check(fn(self._h, _dptr(B), self._n, B.shape[1], _dptr(X), self._n), self._h)
        except HodlrNotConverged as e:
            e.x = X  # the best iterate
            raise
        return X

    def recSM_Solve(self, input_RHS):
        return self._solve(lib().hodlr_solve_sm, input_RHS)

    def recLU_Solve(self, input_RHS):
        return self._solve(lib().hodlr_solve_lu, input_RHS)

    def iterative_Solve(self, input_RHS, maxIterations: int = 1000, stop_tolerance: float = 1e-10,
                        init_LRTolerance: float = 1e-2, input_LRMethod: str = "partialPiv_ACA",
                        directSolve_Method: str = "recLU"):
        """iterative_Solve (HODLR_Matrix.hpp:219): GMRES(32) on the exact
        operator, left-preconditioned by the HODLR solve compressed at
        init_LRTolerance.  Unlike the reference, maxIterations and
        stop_tolerance are honoured (its fixed values are the defaults)."""
        if input_LRMethod not in _lib.LR_METHODS:
            raise ValueError(f"Invalid low-rank approximation scheme ({input_LRMethod})")
        which = {"recLU": _lib.FACTOR_LU, "recSM": _lib.FACTOR_SM}.get(directSolve_Method)
        if which is None:
            raise ValueError(f"Unknown direct solver type ({directSolve_Method})")
        B = _as_f64_2d(input_RHS, self._n)
        X = np.zeros_like(B, order="F")
        try:
            check(lib().hodlr_iterative_solve(self._h, _dptr(B), self._n, B.shape[1], _dptr(X),
                                              self._n, int(maxIterations), float(stop_tolerance),
                                              float(init_LRTolerance),
                                              _lib.LR_METHODS[input_LRMethod], which), self._h)
        except HodlrNotConverged as e:
            e.x = X
            raise
        return X

    def logAbsDeterminant(self) -> float:
        out = C.c_double()
        check(lib().hodlr_logabsdet(self._h, C.byref(out)), self._h)
        return out.value

    def determinant(self) -> float:
        out = C.c_double()
        check(lib().hodlr_determinant(self._h, C.byref(out)), self._h)
        return out.value

    def matvec(self, X):
        """A_H X with the compressed FP64 operator."""
        X = _as_f64_2d(X, self._n)
        Y = np.zeros_like(X, order="F")
        check(lib().hodlr_matvec(self._h, _dptr(X), self._n, X.shape[1], _dptr(Y), self._n), self._h)
        return Y

    def residual(self, B, X):
        """B - A_H X in double-double arithmetic, rounded once."""
        B = _as_f64_2d(B, self._n)
        X = _as_f64_2d(X, self._n)
        if B.shape[1] != X.shape[1]:
            raise ValueError("B and X have different numbers of columns")
        R = np.zeros_like(B, order="F")
        check(lib().hodlr_residual(self._h, _dptr(B), self._n, B.shape[1], _dptr(X), self._n,
                                   _dptr(R), self._n), self._h)
        return R

    # ---- tree surgery (HODLR_Matrix.hpp:256-261) ---------------------------
    def _diag(self, side):
        h = C.c_void_p()
        check(lib().hodlr_diag_block(self._h, int(side), C.byref(h)), self._h)
        return HODLR_Matrix(_handle=h, _n=lib().hodlr_rows(h))

    def topDiag(self):
        """The root's top-left diagonal block as a HODLR matrix (tree part and
        factors kept; this matrix is unchanged)."""
        return self._diag(0)

    def bottDiag(self):
        return self._diag(1)

    def splitAtTop(self):
        """splitAtTop(self, topHODLR, bottHODLR): (top, bottom); unlike the
        reference, this matrix keeps its tree."""
        t, b = C.c_void_p(), C.c_void_p()
        check(lib().hodlr_split_at_top(self._h, C.byref(t), C.byref(b)), self._h)
        return (HODLR_Matrix(_handle=t, _n=lib().hodlr_rows(t)),
                HODLR_Matrix(_handle=b, _n=lib().hodlr_rows(b)))

    def user_tree(self):
        """This matrix's partition as a user tree node array (every node
        explicit, no rank bounds)."""
        return tree_of(self.nodes(), self._n)

    # ---- entries (HODLR_Matrix.hpp:252-254) -------------------------------
    def _block(self, which, min_i, min_j, numRows, numCols):
        out = np.zeros((numRows, numCols), order="F")
        check(lib().hodlr_block(self._h, which, int(min_i), int(min_j), int(numRows),
                                int(numCols), _dptr(out), max(1, int(numRows))), self._h)
        return out

    def block(self, min_i: int, min_j: int, numRows: int, numCols: int):
        """Entries of the matrix the handle was built from (the reference's
        block() when it holds the matrix data)."""
        return self._block(0, min_i, min_j, numRows, numCols)

    def row(self, row: int):
        return self.block(row, 0, 1, self._n)

    def col(self, col: int):
        return self.block(0, col, self._n, 1)

    def block_compressed(self, min_i: int, min_j: int, numRows: int, numCols: int):
        """Entries of the compressed operator A_H (leaves plus U V^T: the
        reference's fill_Block, HODLR_Matrix.cpp:1424-1480)."""
        return self._block(1, min_i, min_j, numRows, numCols)

    # ---- accessors -------------------------------------------------------
    def rows(self) -> int:
        return self._n

    def cols(self) -> int:
        return self._n

    def stats(self) -> dict:
        s = _lib.Stats()
        check(lib().hodlr_get_stats(self._h, C.byref(s)), self._h)
        L = s.num_levels
        return dict(num_levels=L, num_nodes=s.num_nodes, num_leaves=s.num_leaves,
                    aca_iterations=s.aca_iterations, compress_time=s.compress_time,
                    factor_time=s.factor_time, factor_stream_time=s.factor_stream_time,
                    solve_time=s.solve_time,
                    refine_time=s.refine_time, refine_iterations=s.refine_iterations,
                    last_residual=s.last_residual, max_rank=list(s.max_rank[:L]),
                    mean_rank=list(s.mean_rank[:L]), device_bytes=s.device_bytes,
                    aca_row_time=s.aca_row_time, aca_col_time=s.aca_col_time,
                    aca_launches=s.aca_launches, aca_bytes=s.aca_bytes,
                    pred_blocks=s.pred_blocks, pred_misses=s.pred_misses,
                    pred_log_err=s.pred_log_err, graph_captures=s.graph_captures,
                    graph_launches=s.graph_launches, iter_iterations=s.iter_iterations,
                    iter_precond_residual=s.iter_precond_residual,
                    iter_residual=s.iter_residual,
                    aca_panel_grows=s.aca_panel_grows, aca_panel_bytes=s.aca_panel_bytes,
                    factor_groups={
                        name: dict(time=s.fac_time[i], flops=s.fac_flops[i],
                                   bytes=s.fac_bytes[i], bound_time=s.fac_bound_time[i],
                                   launches=s.fac_launches[i])
                        for i, name in enumerate(("gemm", "lu", "leaf_solve", "node_solve"))},
                    tuning=dict(aca_rows_per_thread=s.tune_aca_rows_per_thread,
                                gemm_tile_m=s.tune_gemm_tile_m,
                                gemm_split_target=s.tune_gemm_split_target,
                                lu_panel=s.tune_lu_panel, gemm_mfma=bool(s.tune_gemm_mfma),
                                leaf_solve=s.tune_leaf_solve,
                                tuned=bool(s.tuned)))

    def get_LR_ComputationTime(self) -> float:
        return self.stats()["compress_time"]

    def nodes(self):
        """(level, min_i, split, max_i, rank_top, rank_bottom) per non-leaf node."""
        L = lib()
        c = L.hodlr_num_nodes(self._h)
        arrs = [np.zeros(c, dtype=np.int32) for _ in range(6)]
        check(L.hodlr_get_nodes(self._h, *[a.ctypes.data_as(C.POINTER(C.c_int)) for a in arrs]), self._h)
        return np.stack(arrs, axis=1)

    def factor(self, idx: int, which: int):
        """Low-rank factor of node idx: 0 top U, 1 top V, 2 bottom U, 3 bottom V."""
        nd = self.nodes()[idx]
        lvl, lo, sp, hi, kt, kb = (int(v) for v in nd)
        rows = {0: sp - lo + 1, 1: hi - sp, 2: hi - sp, 3: sp - lo + 1}[which]
        k = kt if which < 2 else kb
        out = np.zeros((rows, k), order="F")
        check(lib().hodlr_get_factor(self._h, idx, which, _dptr(out)), self._h)
        return out


class HODLR_Batch(HODLR_Matrix):
    """Many independent kernel systems in one handle (stacked rows).

    points: list of arrays (n_q x dim); kernels, length_scales, diags: one
    value per instance or a scalar for all.  Solves take and return the
    stacked right-hand sides (sum n_q rows); split() cuts a stacked array into
    the instances' parts.  No reference counterpart: a batch of the
    reference's HODLR_Matrix objects processed together.
    """

    def __init__(self, points, kernel="se", length_scales=1.0, diags=0.0, sizeThreshold=30,
                 diag_mode="add"):
        count = len(points)
        pts = [np.ascontiguousarray(np.asarray(p, dtype=np.float64)) for p in points]
        pts = [p.reshape(-1, 1) if p.ndim == 1 else p for p in pts]

        def per(v):
            return list(v) if np.ndim(v) else [v] * count

        fam, ls, dg = per(kernel), per(length_scales), per(diags)
        descs = (_lib.KernelDesc * count)()
        for q, p in enumerate(pts):
            descs[q] = _lib.KernelDesc(_lib.KERNELS[fam[q]], p.shape[1], _dptr(p), float(ls[q]),
                                       float(dg[q]), {"add": 0, "replace": 1}[diag_mode])
        ns = np.array([p.shape[0] for p in pts], dtype=np.int32)
        h = C.c_void_p()
        check(lib().hodlr_create_batch(C.byref(h), count, ns.ctypes.data_as(C.POINTER(C.c_int)),
                                       descs, int(sizeThreshold)))
        super().__init__(_handle=h, _n=int(ns.sum()))
        self._count = count

    def count(self) -> int:
        return lib().hodlr_batch_count(self._h)

    def offsets(self):
        out = np.zeros(self._count + 1, dtype=np.int32)
        check(lib().hodlr_batch_offsets(self._h, out.ctypes.data_as(C.POINTER(C.c_int))), self._h)
        return out

    def split(self, X):
        o = self.offsets()
        return [X[o[q]:o[q + 1]] for q in range(self._count)]

    def logAbsDeterminants(self):
        out = np.zeros(self._count)
        check(lib().hodlr_logabsdet_batch(self._h, _dptr(out)), self._h)
        return out


__all__ = ["HODLR_Matrix", "HODLR_Batch", "HodlrError", "HodlrNotConverged"]
