/*
 * The exact kernel operator Y = K X with entries evaluated on the fly (see
 * exact.hip).
 */
#ifndef DENSE_HODLR_EXACT_H
#define DENSE_HODLR_EXACT_H

#include "common.h"

namespace hodlr {

/* One instance's product: Y (n x nrhs, ld ldy) = K_inst X (n x nrhs, ld ldx),
 * n the instance size; tile0 its first workgroup. */
struct KopDesc {
  const double* X;
  double* Y;
  int inst;
  int ldx, ldy;
  int tile0;
};

/* workgroups of one instance of n rows */
int kop_tiles(int n, int nrhs);
/* every instance of the launch, ntiles the sum of kop_tiles (kernel sources
 * only: dense sources are products of stored matrices) */
void launch_kernel_operator(const KopDesc* d, int count, int ntiles, const EntrySource* srcs,
                            int nrhs, hipStream_t s);

}  // namespace hodlr

#endif
