"""Dense matrix files in the reference's matrixIO formats (src/matrixIO.cpp),
through the C ABI (hodlr_read_binary_matrix and friends): the binary format
of saveMatrixXdToBinary / readBinaryIntoMatrixXd and the text format of
readTxtIntoMatrix, with the reference's function names.  File access is
host code; no GPU is needed."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._lib import check, lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _read(fn, path):
    p = os.fsencode(path)
    r, c = C.c_int(), C.c_int()
    check(fn(p, C.byref(r), C.byref(c), None, 0))
    A = np.zeros((r.value, c.value), order="F")
    check(fn(p, C.byref(r), C.byref(c), _dp(A), max(1, r.value)))
    return A


def readBinaryIntoMatrixXd(path) -> np.ndarray:
    """readBinaryIntoMatrixXd (include/matrixIO.hpp:51)."""
    return _read(lib().hodlr_read_binary_matrix, path)


def saveMatrixXdToBinary(A, path) -> None:
    """saveMatrixXdToBinary (include/matrixIO.hpp:42)."""
    A = np.asfortranarray(np.asarray(A, dtype=np.float64))
    if A.ndim != 2:
        raise ValueError("a matrix is two-dimensional")
    check(lib().hodlr_write_binary_matrix(os.fsencode(path), A.shape[0], A.shape[1], _dp(A),
                                          max(1, A.shape[0])))


def readTxtIntoMatrix(path) -> np.ndarray:
    """readTxtIntoMatrix (include/matrixIO.hpp:20)."""
    return _read(lib().hodlr_read_text_matrix, path)


__all__ = ["readBinaryIntoMatrixXd", "saveMatrixXdToBinary", "readTxtIntoMatrix"]
