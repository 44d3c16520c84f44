/*
 * Batched one-sided Jacobi SVD of the off-diagonal blocks (the reference's
 * SVD_LowRankApprox, lowRank.cpp:520-569): every block of every level is
 * orthogonalized in the same launches.
 */
#ifndef DENSE_HODLR_SVD_H
#define DENSE_HODLR_SVD_H

#include "common.h"

namespace hodlr {

/* One working matrix A (m x n, m >= n, column-major, ld m) and its
 * accumulated right rotations V (n x n, ld n).  Pairs of columns are rotated
 * in round-robin order: step s of a sweep pairs positions p and n' - 1 - p,
 * n' = n rounded up to even, position 0 fixed and the others cycling. */
struct SvdDesc {
  double* A;
  double* V;
  double floor2; /* couplings |a_p . a_q| <= floor2 are not rotated:
                    (eps * largest initial column norm)^2 */
  int m, n;
  int wg0;  /* first workgroup of this block in the step launch */
  int pad_;
};

/* Copy of one column, scaled: dst[i] = src[i] * f (mode 0), src[i] / f
 * (mode 1; f = 0 gives zeros). */
struct ColCopyDesc {
  const double* src;
  double* dst;
  double f;
  int rows;
  int mode;
};

int svd_step_wgs(const SvdDesc& d);  /* workgroups of a block per step (n'/2) */

/* one rotation step s for every block; rotations performed are counted in
 * *rotated */
void launch_svd_step(const SvdDesc* d, int count, int nwg, int step, unsigned* rotated,
                     hipStream_t s);
/* out[c + off[b]] = ||A_b(:, c)|| for every block b and column c < n_b */
void launch_svd_norms(const SvdDesc* d, int count, const int* off, int maxn, double* out,
                      hipStream_t s);
void launch_col_copy(const ColCopyDesc* d, int count, hipStream_t s);

}  // namespace hodlr

#endif
