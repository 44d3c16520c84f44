"""MI355X-native HODLR engine: drop-in for the HODLR_Matrix path of
amiraa127/Dense_HODLR (kernel/dense constructors, partial-pivot ACA or SVD
compression, recSM/recLU solves, log-determinant), built as hand-written HIP
kernels for gfx950 behind the C ABI in include/dense_hodlr.h."""
from ._lib import HodlrError, HodlrNotConverged, lib
from .hodlr_matrix import HODLR_Batch, HODLR_Matrix
from .matrix_io import readBinaryIntoMatrixXd, readTxtIntoMatrix, saveMatrixXdToBinary
from .memory import device_peaks, memory_cache, memory_stats, memory_trim
from .rank_model import RankModel
from .user_tree import UserTreeNode, get_KDTree_Sorted, kd_sort, user_IndexTree

__all__ = ["HODLR_Matrix", "HODLR_Batch", "RankModel", "HodlrError", "HodlrNotConverged", "lib", "memory_stats", "device_peaks",
           "memory_trim", "memory_cache", "readBinaryIntoMatrixXd", "saveMatrixXdToBinary",
           "readTxtIntoMatrix", "UserTreeNode", "user_IndexTree", "kd_sort", "get_KDTree_Sorted"]
