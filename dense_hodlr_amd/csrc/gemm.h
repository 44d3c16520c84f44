/*
 * Grouped FP64 GEMM (see gemm.hip).
 */
#ifndef DENSE_HODLR_GEMM_H
#define DENSE_HODLR_GEMM_H

#include <vector>

#include "common.h"

namespace hodlr {

/* Tile height (128 or 64) that suits a group of products.  The tiles are
 * 128 x 64 or 64 x 64 on the vector units; FP64 on the matrix cores
 * (mfma) uses 128 x 64 or 64 x 128, whichever pads the group less. */
template <class T>
int gemm_tile_m(const std::vector<GemmDescT<T>>& descs, bool mfma = false);

/* Tile width that goes with tile height bm. */
template <class T>
int gemm_tile_n(int bm, bool mfma);

/* Assigns tile offsets and K splits for tiles bm x gemm_tile_n(bm, mfma);
 * returns the workspace size in elements. */
template <class T>
size_t gemm_plan(std::vector<GemmDescT<T>>& descs, std::vector<int>& split_ids, int& ntiles,
                 int target_tiles, int bm, bool mfma = false);

/* mfma (FP64): run the tiles on the matrix cores; bm and mfma as given to
 * gemm_plan */
template <class T>
void launch_gemm(const GemmDescT<T>* d_descs, int count, int ntiles, T* d_ws,
                 const int* d_split_ids, int nsplit, int bm, hipStream_t s, bool mfma = false);

/* The factorization's updates C -= A B (op(A) = A, alpha = +-1, beta = 1)
 * of rank K <= GEMM_UPDATE_KMAX and at least 128 columns, streamed on the
 * matrix cores (bit-identical to launch_gemm with mfma).  A workgroup takes
 * gemm_update_rows(maxk) rows of C; gemm_update_plan numbers the problems'
 * workgroups (tile0) for that row count and returns their total. */
constexpr int GEMM_UPDATE_KMAX = 64;
int gemm_update_variant(int maxk);
int gemm_update_rows(int maxk);
bool gemm_update_eligible(const GemmDescT<double>& g);
int gemm_update_plan(std::vector<GemmDescT<double>>& descs, int rows);
void launch_gemm_update(const GemmDescT<double>* d_descs, int count, int nwg, int maxk,
                        hipStream_t s);

}  // namespace hodlr

#endif
