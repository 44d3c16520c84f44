/*
 * Level-batched partial-pivot ACA (see aca.hip).
 */
#ifndef DENSE_HODLR_ACA_H
#define DENSE_HODLR_ACA_H

#include "common.h"

namespace hodlr {

constexpr int ACA_MAX_LEVELS = 32;
constexpr int MAX_FIN_CHUNKS = 256;  // ranks up to 256 * 64 = 16384

/* A block whose row and column phases have at most FIN_PACK_SLOTS
 * workgroups together is finalized four 64-value chunks per workgroup, a
 * wave per chunk (aca_fin_kernel); the others one chunk per workgroup. */
constexpr int FIN_PACK_SLOTS = 32;
__host__ __device__ inline bool aca_fin_packed(const AcaBlock& b) {
  return b.row_nwg + b.col_nwg <= FIN_PACK_SLOTS;
}

/* Device panels of the levels being compressed.  W[l], V[l]: ld x cap[l]
 * column-major; P[l]: per-workgroup partial slots of (cap + 4) doubles. */
struct AcaLevels {
  double* W[ACA_MAX_LEVELS];
  double* V[ACA_MAX_LEVELS];
  double* P[ACA_MAX_LEVELS];
  int cap[ACA_MAX_LEVELS];
  int ld;
};

struct AcaLaunch {
  const AcaBlock* blocks;
  AcaState* state;
  uint32_t* bits;
  const int* wg2blk_row;
  const int* wg2blk_col;
  const int* wg2blk_fin;
  const int* wg2chunk_fin;
  const int* nchunk;
  double* fin_part;
  int* counters;
  int n_row_wg, n_col_wg, n_fin_wg;
  int jpt, ipt;
  double tol;
  int min_pivot;
};

void aca_launch_init(const AcaBlock* d_blocks, AcaState* d_state, uint32_t* d_bits, int nblocks,
                     hipStream_t s);
/* One ACA iteration (row, column, finalize launches).  With ev non-null,
 * events ev[0..2] are recorded before the row phase, between the row and
 * column phases and after the column phase. */
void aca_launch_iteration(const AcaLevels& L, const AcaLaunch& a, const EntrySource* d_srcs,
                          hipStream_t s, hipEvent_t* ev = nullptr);
void aca_launch_post(const AcaLevels& L, const AcaBlock* d_blocks, AcaState* d_state,
                     const EntrySource* d_srcs, int* d_ranks, int nblocks, hipStream_t s);

}  // namespace hodlr

#endif
