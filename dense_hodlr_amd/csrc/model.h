/*
 * Rank-predictor model object: the MLP parameters and Adam state on the
 * device, training and inference through the kernels of mlp.hip.
 */
#ifndef DENSE_HODLR_MODEL_H
#define DENSE_HODLR_MODEL_H

#include <cstdint>
#include <vector>

#include "engine.h"
#include "mlp.h"

namespace hodlr {

class RankModel {
 public:
  explicit RankModel(uint64_t seed);
  ~RankModel();
  RankModel(const RankModel&) = delete;
  RankModel& operator=(const RankModel&) = delete;

  void get_params(float* out);
  void set_params(const float* in);
  const std::vector<float>& host_params();
  /* y = f(X), FP32 (n x 8 row-major inputs, host pointers) */
  void forward(const float* X, int n, float* y);
  /* y = f(X) in BF16 arithmetic (the inference used for rank sizing) */
  void predict_bf16(const float* X, int n, float* y);
  /* gradient of mean((f(X) - t)^2) over all n samples; returns the loss */
  float gradient(const float* X, const float* t, int n, float* grad);
  /* Adam on mini-batches of a seeded permutation per epoch; epoch_loss
   * (epochs entries, may be null) receives the full-data loss after each */
  void train(const float* X, const float* t, int n, int epochs, int batch, float lr, uint64_t seed,
             float* epoch_loss);

 private:
  void ensure(int n);
  void fwd_bwd(const float* dX, const float* dt, int n);
  hipStream_t s_ = nullptr;
  DBuf<float> P_, M_, V_, G_;
  DBuf<float> X_, T_, Z1_, Z2_, Y_, work_, part_;
  int cap_ = 0;
  int step_ = 0;
  std::vector<float> host_;
};

}  // namespace hodlr

#endif
