/*
 * Grouped FP64 GEMM (see gemm.hip).
 */
#ifndef DENSE_HODLR_GEMM_H
#define DENSE_HODLR_GEMM_H

#include <vector>

#include "common.h"

namespace hodlr {

/* Tile height (128 or 64) that suits a group of products. */
template <class T>
int gemm_tile_m(const std::vector<GemmDescT<T>>& descs);

/* Assigns tile offsets and K splits for tiles bm x 64; returns the workspace
 * size in elements. */
template <class T>
size_t gemm_plan(std::vector<GemmDescT<T>>& descs, std::vector<int>& split_ids, int& ntiles,
                 int target_tiles, int bm);

/* mfma (FP64, bm 128 only): run the tiles on the matrix cores */
template <class T>
void launch_gemm(const GemmDescT<T>* d_descs, int count, int ntiles, T* d_ws,
                 const int* d_split_ids, int nsplit, int bm, hipStream_t s, bool mfma = false);

}  // namespace hodlr

#endif
