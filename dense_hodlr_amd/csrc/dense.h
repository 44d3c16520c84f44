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
  /* LU_SWP_INTS ints of scratch: the rows the latest panel's interchanges
   * move in the columns outside it (destinations at [0, 2 pw), their
   * sources at [LU_SWP_SRC, LU_SWP_SRC + 2 pw)), for lu_step_kernel */
  int* swp = nullptr;
};
constexpr int LU_SWP_SRC = 128;
constexpr int LU_SWP_INTS = 256;

/* One blocked-LU step outside the panel [p0, p0 + pw) of a matrix: the
 * panel's row interchanges for the columns left of it (chunks [0, nleft))
 * and right of it, and for the right ones the triangular solve
 * U12 = L11^-1 A12.  64 columns per workgroup. */
struct LuStepDesc {
  double* A;
  const int* swp;
  int n, ld, p0, pw;
  int chunk0; /* first workgroup of this matrix */
  int nleft;  /* workgroups of the columns left of the panel */
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
  /* the interchanges as one permutation (row i of P B is row perm[i] of B);
   * read by the matrix-core kernel only */
  const int* perm = nullptr;
};

/* Build job of one permutation: perm[0..n) from the interchanges piv[0..n). */
struct PermBuild {
  const int* piv;
  int* perm;
  int n;
  int pad_;
};

/* Dense block of entries: out[i + j*ldo] = K(i0 + i, j0 + j) of instance
 * inst (i < m, j < n), or its transpose out[j + i*ldo] when trans = 1. */
struct AssembleDesc {
  double* out;
  int inst;
  int i0, j0;
  int m, n;
  int ldo;
  int trans;
  int pad_;
};

/* out(i, j) = sum_l U(i, l) V(j, l) over an m x n piece of a low-rank block
 * (U m x k with ld ldu, V n x k with ld ldv; out ld ldo). */
struct LrFillDesc {
  const double* U;
  const double* V;
  double* out;
  int m, n, k;
  int ldu, ldv, ldo;
};

/* Copy of an m x n column-major block. */
struct CopyDesc {
  const double* src;
  double* dst;
  int m, n;
  int lds, ldd;
};

/* Assemble the small node matrix and its right-hand side from the Gram
 * products of one node (see engine.cpp for the layout).  The products are
 * FP64 (plain path) or FP32 (mixed path); C and Y are always FP64. */
template <class TZ>
struct CapDescT {
  const TZ* Zt; /* kb x ncols, ld kb: V_b^T [X_top | U_t^SM] */
  const TZ* Zb; /* kt x ncols, ld kt: V_t^T [X_bot | U_b^SM] */
  double* C;        /* K x K, K = kt + kb */
  double* Y;        /* K x nanc, ld K */
  int kt, kb;
  int nanc;         /* ancestor (right-hand side) columns */
  int own0;         /* column of U^SM own block inside Z */
  int lu_order;     /* 0: recSM capacitance; 1: recLU Schur ordering */
  int pad_;
};
using CapDesc = CapDescT<double>;
using CapDescF = CapDescT<float>;

/* m x n block copy with conversion between FP64 and FP32. */
template <class S, class D>
struct ConvDesc {
  const S* src;
  D* dst;
  int m, n;
  int lds, ldd;
};

/* FP64 -> FP32 conversion of one low-rank block U V^T (U: mu x k rows of a
 * level panel, V: mv x k) with every term rebalanced: u_c s_c and v_c / s_c,
 * s_c = sqrt(|v_c| / |u_c|).  The product is unchanged; partial-pivot ACA
 * leaves 1/pivot in u_c, which for the small late terms exceeds the FP32
 * range while v_c is tiny. */
struct BalanceDesc {
  const double* U;
  const double* V;
  float* U32;
  float* V32;
  int mu, mv, k;
  int ldu, ldv, ldu32, ldv32;
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

/* Panel step of the batched blocked LU at columns [p0, p0 + nb); matrices of
 * the batch have at most maxn rows.  lu_panel_width(maxn, nb) is the widest
 * panel the register-resident kernel takes for that size (it falls back to
 * the global-memory kernel beyond 4096 rows). */
int lu_panel_width(int maxn, int nb);
void launch_lu_panel(const LuDesc* d, int count, int p0, int nb, int maxn, hipStream_t s);
/* Whether the panel kernel for (maxn, nb) leaves the interchanges of the
 * columns outside the panel to launch_lu_step (register panels; the
 * global-memory panel kernel swaps whole rows itself). */
bool lu_panel_defers_swaps(int maxn, int nb);
/* swap: apply the interchanges recorded in LuDesc::swp (see above); nb: the
 * batch's panel width (at most 32 when swap). */
void launch_lu_step(const LuStepDesc* d, int count, int nwg, bool swap, int nb, hipStream_t s);
void launch_trsm(const TrsmDesc* d, int count, int nwg, hipStream_t s);
void launch_permute(const PermDesc* d, int count, int nwg, hipStream_t s);
/* Whole solves of systems of at most LU_SOLVE_SMALL_MAX rows in one launch
 * (interchanges, both substitutions); SolveDesc::chunk0 numbers the work
 * units, lu_solve_small_wgs(ncols, maxn, wave) per problem (maxn: the
 * largest system of the launch; nwg: the units of all problems).  A unit is
 * a workgroup, or with wave set and maxn <= LU_SOLVE_WAVE_MAX a single wave
 * (the barrier-free kernel; same results).  With blocked set, systems above
 * 512 rows are substituted in blocks of 64 rows (same results).  Returns
 * false, and launches nothing, when maxn exceeds the limit. */
constexpr int LU_SOLVE_SMALL_MAX = 2048;
constexpr int LU_SOLVE_WAVE_MAX = 512;
int lu_solve_small_wgs(int ncols, int maxn, bool wave);
/* The same solves for many right-hand sides of systems of at most
 * LU_SOLVE_WAVE_MAX rows, the factor shared by the four waves of a workgroup
 * (bit-identical to the one-wave kernel); SolveDesc::chunk0 numbers
 * workgroups, lu_solve_shared_wgs(ncols, maxn) per problem.  Returns false,
 * and launches nothing, when maxn exceeds the limit. */
int lu_solve_shared_wgs(int ncols, int maxn);
/* The same solves on the matrix cores for systems of at most
 * LU_SOLVE_MFMA_MAX rows (SolveDesc::perm required): one workgroup per
 * system and 64 columns, lu_solve_mfma_wgs(ncols) per problem numbered by
 * chunk0.  Not bit-identical to the substitution kernels (the off-diagonal
 * tiles are summed four terms at a time by the matrix cores).  Returns
 * false, and launches nothing, when maxn exceeds the limit. */
constexpr int LU_SOLVE_MFMA_MAX = 256;
int lu_solve_mfma_wgs(int ncols);
bool launch_lu_solve_mfma(const SolveDesc* d, int count, int nwg, int maxn, hipStream_t s);
/* The one-wave solves with 16 columns per wave (twice the columns per factor
 * load; bit-identical to launch_lu_solve_small's one-wave kernel) for
 * systems of at most LU_SOLVE_WIDE_MAX rows; lu_solve_wide_units(ncols) per
 * problem numbered by chunk0.  Returns false, and launches nothing, when
 * maxn exceeds the limit. */
constexpr int LU_SOLVE_WIDE_MAX = 256;
int lu_solve_wide_units(int ncols);
bool launch_lu_solve_wide(const SolveDesc* d, int count, int nunits, int maxn, hipStream_t s);
/* One wave per job; n <= LU_PERM_MAX. */
constexpr int LU_PERM_MAX = 4096;
void launch_lu_perm(const PermBuild* d, int count, hipStream_t s);
bool launch_lu_solve_shared(const SolveDesc* d, int count, int nwg, int maxn, hipStream_t s);
bool launch_lu_solve_small(const SolveDesc* d, int count, int nwg, int maxn, bool wave,
                           bool blocked, hipStream_t s);
void launch_assemble(const AssembleDesc* d, int count, int max_elems, const EntrySource* srcs,
                     hipStream_t s);
void launch_set_identity(const CopyDesc* d, int count, hipStream_t s);
void launch_copy(const CopyDesc* d, int count, hipStream_t s);
/* every piece of the launch (max_elems: the largest m * n) */
void launch_lr_fill(const LrFillDesc* d, int count, size_t max_elems, hipStream_t s);
void launch_logdet(const LuDesc* d, int count, double* out_log, double* out_prod, hipStream_t s);
template <class TZ>
void launch_capacitance(const CapDescT<TZ>* d, int count, hipStream_t s);
template <class S, class D>
void launch_convert(const ConvDesc<S, D>* d, int count, hipStream_t s);
/* one workgroup per (term c < kmax, block); terms c >= k of a block are left
 * untouched */
void launch_balance(const BalanceDesc* d, int count, int kmax, hipStream_t s);
/* out[c] = sum_i X[i + c*ld]^2 for c < ncols (one workgroup per column, fixed
 * reduction order) */
void launch_colnorm2(const double* X, int n, int ld, int ncols, double* out, hipStream_t s);
/* Per instance segment: out[p + nsplit*(s + c*nseg)] = sum of squares of
 * the p-th of nsplit equal parts of rows [row0[s], row0[s+1]) of column c
 * (the caller adds the parts in order); segscale multiplies those rows by
 * sc[s + c*nseg].  row0 has nseg + 1 entries (device). */
void launch_segnorm2(const double* X, int ld, const int* row0, int nseg, int ncols, int nsplit,
                     double* out, hipStream_t s);
void launch_segscale(double* X, int ld, const int* row0, int nseg, int ncols, const double* sc,
                     hipStream_t s);
/* Y <- X on the rows of segment s of column c where mask[s + c*nseg] != 0
 * (mask on the device; the rows split into nsplit parts per workgroup) */
void launch_segcopy(const double* X, double* Y, int ld, const int* row0, int nseg, int ncols,
                    int nsplit, const int* mask, hipStream_t s);
/* X += alpha * D (FP32 correction) over an n x ncols block */
void launch_add_f32(double* X, int ldx, const float* D, int ldd, int n, int ncols, hipStream_t s);
/* Batched column operations of the block GMRES: dots of w with nvec basis
 * blocks V + i*vstride (out is nvec x ncols), the combination
 * w = (assign ? 0 : w) + sign * sum_i V_i coef(i,:), and column scaling. */
void launch_multi_dot(const double* V, size_t vstride, int nvec, const double* w, int n, int ld,
                      int ncols, double* out, hipStream_t s);
void launch_multi_axpy(const double* V, size_t vstride, int nvec, const double* coef, double* w,
                       int n, int ld, int ncols, double sign, int assign, hipStream_t s);
void launch_col_scale(double* X, int ld, int n, int ncols, const double* sc, hipStream_t s);
/* x <- fl(x + d); d <- (x_new - x_old) up to one rounding (n x ncols, ld) */
void launch_step_update(double* x, double* d, int ld, int n, int ncols, hipStream_t s);
/* Adds the number of non-finite entries of x[0, n) to *out (diagnostics). */
template <class T>
void launch_nonfinite(const T* x, size_t n, unsigned long long* out, hipStream_t s);
/* R = B - Y over an n x ncols block */
void launch_sub(const double* B, int ldb, const double* Y, int ldy, double* R, int ldr, int n,
                int ncols, hipStream_t s);

}  // namespace hodlr

#endif
