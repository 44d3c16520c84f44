/*
 * Grouped FP64 GEMM (see gemm.hip).
 */
#ifndef DENSE_HODLR_GEMM_H
#define DENSE_HODLR_GEMM_H

#include <vector>

#include "common.h"

namespace hodlr {

/* Assigns tile offsets and K splits; returns the workspace size in doubles. */
size_t gemm_plan(std::vector<GemmDesc>& descs, std::vector<int>& split_ids, int& ntiles,
                 int target_tiles);

void launch_gemm(const GemmDesc* d_descs, int count, int ntiles, double* d_ws,
                 const int* d_split_ids, int nsplit, hipStream_t s);

}  // namespace hodlr

#endif
