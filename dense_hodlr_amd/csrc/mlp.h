/*
 * Rank-predictor network: an MLP 8 -> 64 -> 64 -> 1 with GELU activations
 * (0.5 x (1 + erf(x / sqrt 2))) that maps geometric features of an
 * off-diagonal block to the log of its ACA rank.  FP32 parameters, training
 * by Adam on the mean squared error, inference in BF16 (bf16 weights and
 * activations, FP32 accumulation).
 *
 * Parameter vector (FP32, MLP_NPARAM entries), row-major weights:
 *   W1[64][8], b1[64], W2[64][64], b2[64], W3[64], b3
 */
#ifndef DENSE_HODLR_MLP_H
#define DENSE_HODLR_MLP_H

#include "common.h"

namespace hodlr {

constexpr int MLP_IN = 8;
constexpr int MLP_H = 64;
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_H * MLP_IN;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_H;
constexpr int MLP_NPARAM = MLP_OFF_B3 + 1;  /* 4801 */
constexpr int MLP_CHUNK = 512;              /* samples per gradient partial */

/* Forward in FP32: X is n x 8 (row-major), Z1/Z2 (64 x n, feature-major)
 * receive the pre-activations when not null, y the n outputs. */
void launch_mlp_forward(const float* P, const float* X, int n, float* Z1, float* Z2, float* y,
                        hipStream_t s);
/* Backward of the loss mean((y - t)^2): per-sample deltas and
 * post-activations into work (mlp_work_floats(n) floats), then gradient
 * partials per chunk of MLP_CHUNK samples reduced in chunk order into G
 * (MLP_NPARAM); part holds mlp_chunks(n) * MLP_NPARAM floats. */
inline size_t mlp_work_floats(int n) { return (size_t)(4 * MLP_H + 1) * n; }
inline int mlp_chunks(int n) { return (n + MLP_CHUNK - 1) / MLP_CHUNK; }
void launch_mlp_backward(const float* P, const float* X, const float* t, int n, const float* Z1,
                         const float* Z2, const float* y, float* work, float* part, float* G,
                         hipStream_t s);
/* Adam step (bias-corrected), step number t >= 1. */
void launch_mlp_adam(float* P, const float* G, float* M, float* V, float lr, float b1, float b2,
                     float eps, int t, hipStream_t s);
/* Inference in BF16: parameters and activations rounded to bfloat16 (round
 * to nearest even), products accumulated in FP32. */
void launch_mlp_infer_bf16(const float* P, const float* X, int n, float* y, hipStream_t s);

/* Geometric features of off-diagonal blocks (points of one instance):
 * one descriptor per block. */
struct FeatDesc {
  const double* x;
  const double* y;
  const double* z;
  int r0, nr;    /* row points [r0, r0 + nr) */
  int c0, nc;    /* column points */
  int dim;
  int level;
  float length_scale;
  float log10_tol;
  float family;
  int pad_;
};
/* out[b*8 + f]: log(nr + nc), |c_r - c_c| / (d_r + d_c), d_r / l, d_c / l,
 * log(l), log10(tol), level, family, where c are centroids and d the largest
 * distance of a point from its centroid. */
void launch_block_features(const FeatDesc* d, int count, float* out, hipStream_t s);

}  // namespace hodlr

#endif
