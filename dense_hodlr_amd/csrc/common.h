/*
 * Shared definitions of the HODLR engine: error handling and the descriptor
 * structs that host code fills and device kernels read.
 */
#ifndef DENSE_HODLR_COMMON_H
#define DENSE_HODLR_COMMON_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "kernel_fn.h"

namespace hodlr {

struct HipError : std::runtime_error {
  explicit HipError(const std::string& s) : std::runtime_error(s) {}
};

/* Refinement ended above its target: the output holds the best iterate
 * (column by column) and the stats its residual. */
struct NotConverged : std::runtime_error {
  explicit NotConverged(const std::string& s) : std::runtime_error(s) {}
};

#define HODLR_HIP_CHECK(expr)                                                          \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      throw ::hodlr::HipError(std::string("HIP error ") + hipGetErrorString(e_) +      \
                              " at " + __FILE__ + ":" + std::to_string(__LINE__) +     \
                              " in " + #expr);                                         \
  } while (0)

/* Pointers read from descriptors are generic (flat) to the compiler, and
 * flat memory instructions count against the LDS wait counter as well: a
 * wait for an LDS read then also waits for every global load or store in
 * flight (a register prefetch, a store before a barrier).  Every device
 * buffer of the engine is global memory; kernels re-type the descriptor
 * pointers they access with to_global. */
#ifdef __HIPCC__
template <class T>
using gptr = T __attribute__((address_space(1)))*;
template <class T>
__device__ __forceinline__ gptr<T> to_global(T* p) {
  return (gptr<T>)p;
}
#endif

/* Source of matrix entries for one instance (problem) of a forest. */
struct EntrySource {
  int kind;            /* 0 typed kernel, 1 dense column-major matrix */
  int n;               /* instance size */
  int ld;              /* dense leading dimension */
  int pad_;
  const double* dense; /* device pointer (kind 1) */
  HodlrKernelParams kp; /* device pointers to coordinates (kind 0) */
};

__host__ __device__ inline double entry_at(const EntrySource& s, int i, int j) {
  if (s.kind == 1) return s.dense[(size_t)i + (size_t)j * s.ld];
  return hodlr_kernel_entry(s.kp, i, j);
}

/* One off-diagonal block compressed by ACA.  Block rows are instance rows
 * [i0, i0+m) and columns [j0, j0+n); its factors live in the level panels:
 * W rows start at global row w_row0, V rows at global row v_row0. */
struct AcaBlock {
  int inst;
  int level;
  int i0, j0;
  int m, n;
  int w_row0, v_row0;
  int row_wg0, row_nwg; /* workgroups of the row phase */
  int col_wg0, col_nwg; /* workgroups of the column phase */
  int bits_row0;        /* word offset of the chosen-row bitmap */
  int bits_col0;        /* word offset of the chosen-column bitmap */
  int part_off;         /* offset (in doubles) of this block's partial slots */
  int lidx;             /* index of the block within its level (Gram slot) */
  int min_rank;         /* the node's rank bounds (-1: none; max_rank 0: a zero */
  int max_rank;         /* rank-1 block without iterating, lowRank.cpp:120-125) */
  int skip;             /* 0 ACA iterates; 1 max_rank 0 (no iterations); 2 another method */
  int pad_;
};

/* Per-block iteration state of ACA (device resident). */
struct AcaState {
  int k;        /* accepted terms */
  int cur_row;  /* row for the next iteration */
  int status;   /* 0 active, 1 done, 2 paused: capacity reached */
  int cur_col;  /* column chosen in the current iteration (-1: none) */
  double frob2; /* incremental ||W V^T||_F^2 */
  double eps;   /* last ||w|| ||v|| / ||W V^T||_F */
  int iters;    /* iterations executed (including row jumps) */
  int pending;  /* 1: term k-1 is stored, its epsilon is completed next iteration */
  int fin_chunks; /* finalize chunks with dot products this iteration (column phase) */
  int pad_;
};

/* One problem of a grouped GEMM: C = alpha * op(A) * B + beta * C,
 * op(A) = A (trans_a 0, A is M x K) or A^T (trans_a 1, A is K x M).
 * Element type T: double (FP64 path) or float (FP32 bulk of the mixed path). */
template <class T>
struct GemmDescT {
  int M, N, K;
  int trans_a;
  int lda, ldb, ldc;
  int ksplit;       /* number of K slices (>=1) */
  int tile0;        /* first tile index of this problem */
  long long ws_off; /* workspace offset (elements) for split-K partials */
  T alpha, beta;
  const T* A;
  const T* B;
  T* C;
};
using GemmDesc = GemmDescT<double>;
using GemmDescF = GemmDescT<float>;

}  // namespace hodlr

#endif
