/*
 * Batched dense kernels of the HODLR engine: leaf assembly, partial-pivot LU,
 * LU solves, determinant reductions, and the gather/scatter kernels that
 * stitch level panels together.
 */
#ifndef DENSE_HODLR_DENSE_H
#define DENSE_HODLR_DENSE_H

#include "common.h"

namespace hodlr {

/* One square matrix of a batch (column-major, in place). */
struct LuDesc {
  double* A;
  int* piv;  /* 0-based row interchanges, LAPACK style */
  int n;
  int ld;
};

/* In-place solve A X = B with A given by an LuDesc; B is n x ncols. */
struct SolveDesc {
  const double* LU;
  const int* piv;
  double* B;
  int n;
  int ld_lu;
  int ldb;
  int ncols;
  int chunk0; /* first workgroup of this problem */
  int pad_;
};

/* Dense block of entries: out[i + j*ldo] = K(i0 + i, j0 + j) of instance inst. */
struct AssembleDesc {
  double* out;
  int inst;
  int i0, j0;
  int m, n;
  int ldo;
};

/* Copy of an m x n column-major block. */
struct CopyDesc {
  const double* src;
  double* dst;
  int m, n;
  int lds, ldd;
};

/* Assemble the small node matrix and its right-hand side from the Gram
 * products of one node (see engine.cpp for the layout). */
struct CapDesc {
  const double* Zt; /* kb x ncols, ld kb: V_b^T [X_top | U_t^SM] */
  const double* Zb; /* kt x ncols, ld kt: V_t^T [X_bot | U_b^SM] */
  double* C;        /* K x K, K = kt + kb */
  double* Y;        /* K x nanc, ld K */
  int kt, kb;
  int nanc;         /* ancestor (right-hand side) columns */
  int own0;         /* column of U^SM own block inside Z */
  int lu_order;     /* 0: recSM capacitance; 1: recLU Schur ordering */
  int pad_;
};

/* Triangular solve with a pw x pw diagonal block (pw <= 64) of an LU
 * factor: unit lower (upper = 0) or upper (upper = 1); B is pw x ncols. */
struct TrsmDesc {
  const double* T;
  double* B;
  int ldt, ldb;
  int pw;
  int upper;
  int ncols;
  int chunk0; /* first workgroup (64 columns each) of this problem */
};

/* Row interchanges piv[0..n) applied in order to B (n x ncols). */
struct PermDesc {
  double* B;
  const int* piv;
  int ldb;
  int n;
  int ncols;
  int chunk0; /* first workgroup (256 columns each) */
};

void launch_lu_panel(const LuDesc* d, int count, int p0, int nb, hipStream_t s);
void launch_trsm(const TrsmDesc* d, int count, int nwg, hipStream_t s);
void launch_permute(const PermDesc* d, int count, int nwg, hipStream_t s);
void launch_assemble(const AssembleDesc* d, int count, int max_elems, const EntrySource* srcs,
                     hipStream_t s);
void launch_set_identity(const CopyDesc* d, int count, hipStream_t s);
void launch_copy(const CopyDesc* d, int count, hipStream_t s);
void launch_logdet(const LuDesc* d, int count, double* out_log, double* out_prod, hipStream_t s);
void launch_capacitance(const CapDesc* d, int count, hipStream_t s);

}  // namespace hodlr

#endif
