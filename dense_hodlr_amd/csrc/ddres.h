/*
 * Compensated (double-double) residual of the HODLR operator:
 * R = B - A_H X with A_H = leaves + U V^T, every product and sum carried as an
 * unevaluated pair (hi, lo) by error-free transformations (TwoProd with FMA,
 * TwoSum), i.e. Ogita-Rump-Oishi Dot2 accumulation.  The result is as
 * accurate as if computed in twice the working precision and rounded once.
 *
 * Why: on smooth kernels with small shifts each row of A_H X sums thousands
 * of terms of size |x| >> |b| that cancel to O(|b|); in FP64 the rounding of
 * that sum alone reaches a relative residual near 1e-10 at N = 262144, so a
 * refinement target of 1e-10 cannot be verified (nor reached) with an FP64
 * residual.
 */
#ifndef DENSE_HODLR_DDRES_H
#define DENSE_HODLR_DDRES_H

#include "common.h"

namespace hodlr {

/* T(c, j) = sum_r V(r, c) X(r, j) over rows r < rows, c < k, j < nrhs, in
 * row chunks of DD_CHUNK; chunk partials (hi, lo) go to part_hi/part_lo at
 * part_off + (chunk * nrhs + j) * k + c, the reduced result to t_hi/t_lo at
 * t_off + j * k + c. */
struct DdProjDesc {
  const double* V;
  const double* X;
  int ldv, ldx;
  int rows, k, nrhs;
  int nchunks;
  int wg0;     /* first workgroup of this problem (projection launch) */
  int rwg0;    /* first workgroup of this problem (reduction launch) */
  long long part_off;
  long long t_off;
};

/* Y(rows x nrhs) (+)= A(rows x k) T(k x nrhs) in double-double; T is (t_hi,
 * t_lo) with leading dimension ldt (t_lo may be null); accumulate = 0 writes
 * Y instead of adding to it. */
struct DdApplyDesc {
  const double* A;
  const double* t_hi;
  const double* t_lo;
  double* y_hi;
  double* y_lo;
  int lda, ldt, ldy;
  int rows, k, nrhs;
  int accumulate;
  int wg0;
};

constexpr int DD_CHUNK = 4096;  /* rows per projection partial */

/* number of workgroups of a projection / reduction / apply problem */
int dd_proj_wgs(const DdProjDesc& d);
int dd_reduce_wgs(const DdProjDesc& d);
int dd_apply_wgs(const DdApplyDesc& d);

/* nrhs: the right-hand side count of every problem of the launch (selects
 * the tile shape, as dd_proj_wgs / dd_apply_wgs do) */
void launch_dd_proj(const DdProjDesc* d, int count, int nwg, int nrhs, double* part_hi,
                    double* part_lo, hipStream_t s);
void launch_dd_reduce(const DdProjDesc* d, int count, int nwg, const double* part_hi,
                      const double* part_lo, double* t_hi, double* t_lo, hipStream_t s);
void launch_dd_apply(const DdApplyDesc* d, int count, int nwg, int nrhs, hipStream_t s);
/* R = B - (Y_hi + Y_lo), rounded once (n x ncols) */
void launch_dd_residual(const double* B, int ldb, const double* y_hi, const double* y_lo, int ldy,
                        double* R, int ldr, int n, int ncols, hipStream_t s);

}  // namespace hodlr

#endif
