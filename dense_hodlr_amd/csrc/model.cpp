/*
 * Rank-predictor model object (see model.h).
 */
#include "model.h"

#include <algorithm>
#include <cmath>

namespace hodlr {

namespace {

/* SplitMix64: next 64-bit output and a double in [0, 1) from its top 53 bits */
uint64_t splitmix_next(uint64_t& s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
double splitmix_uniform(uint64_t& s) { return (double)(splitmix_next(s) >> 11) * 0x1p-53; }

}  // namespace

/* Glorot-uniform weights U(-a, a), a = sqrt(6 / (fan_in + fan_out)), zero
 * biases, drawn in parameter order from SplitMix64(seed). */
RankModel::RankModel(uint64_t seed) {
  HODLR_HIP_CHECK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
  host_.assign(MLP_NPARAM, 0.0f);
  uint64_t st = seed;
  auto fill = [&](int off, int cnt, int fin, int fout) {
    const double a = std::sqrt(6.0 / (fin + fout));
    for (int i = 0; i < cnt; i++) host_[off + i] = (float)((2.0 * splitmix_uniform(st) - 1.0) * a);
  };
  fill(MLP_OFF_W1, MLP_H * MLP_IN, MLP_IN, MLP_H);
  fill(MLP_OFF_W2, MLP_H * MLP_H, MLP_H, MLP_H);
  fill(MLP_OFF_W3, MLP_H, MLP_H, 1);
  P_.alloc(MLP_NPARAM);
  M_.alloc(MLP_NPARAM);
  V_.alloc(MLP_NPARAM);
  G_.alloc(MLP_NPARAM);
  set_params(host_.data());
}

RankModel::~RankModel() {
  if (s_) {
    (void)hipStreamSynchronize(s_);
    (void)hipStreamDestroy(s_);
  }
}

void RankModel::get_params(float* out) {
  HODLR_HIP_CHECK(hipMemcpyAsync(out, P_.p, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, s_));
  HODLR_HIP_CHECK(hipStreamSynchronize(s_));
}

void RankModel::set_params(const float* in) {
  host_.assign(in, in + MLP_NPARAM);
  HODLR_HIP_CHECK(hipMemcpyAsync(P_.p, host_.data(), MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice, s_));
  HODLR_HIP_CHECK(hipMemsetAsync(M_.p, 0, M_.bytes(), s_));
  HODLR_HIP_CHECK(hipMemsetAsync(V_.p, 0, V_.bytes(), s_));
  HODLR_HIP_CHECK(hipStreamSynchronize(s_));
  step_ = 0;
}

const std::vector<float>& RankModel::host_params() {
  get_params(host_.data());
  return host_;
}

void RankModel::ensure(int n) {
  if (n <= cap_) return;
  HODLR_HIP_CHECK(hipStreamSynchronize(s_));
  X_.alloc((size_t)n * MLP_IN);
  T_.alloc(n);
  Z1_.alloc((size_t)MLP_H * n);
  Z2_.alloc((size_t)MLP_H * n);
  Y_.alloc(n);
  work_.alloc(mlp_work_floats(n));
  part_.alloc((size_t)mlp_chunks(n) * MLP_NPARAM);
  cap_ = n;
}

void RankModel::forward(const float* X, int n, float* y) {
  if (n <= 0) return;
  ensure(n);
  HODLR_HIP_CHECK(hipMemcpyAsync(X_.p, X, (size_t)n * MLP_IN * sizeof(float), hipMemcpyHostToDevice, s_));
  launch_mlp_forward(P_.p, X_.p, n, nullptr, nullptr, Y_.p, s_);
  HODLR_HIP_CHECK(hipMemcpyAsync(y, Y_.p, n * sizeof(float), hipMemcpyDeviceToHost, s_));
  HODLR_HIP_CHECK(hipStreamSynchronize(s_));
}

void RankModel::predict_bf16(const float* X, int n, float* y) {
  if (n <= 0) return;
  ensure(n);
  HODLR_HIP_CHECK(hipMemcpyAsync(X_.p, X, (size_t)n * MLP_IN * sizeof(float), hipMemcpyHostToDevice, s_));
  launch_mlp_infer_bf16(P_.p, X_.p, n, Y_.p, s_);
  HODLR_HIP_CHECK(hipMemcpyAsync(y, Y_.p, n * sizeof(float), hipMemcpyDeviceToHost, s_));
  HODLR_HIP_CHECK(hipStreamSynchronize(s_));
}

void RankModel::fwd_bwd(const float* dX, const float* dt, int n) {
  launch_mlp_forward(P_.p, dX, n, Z1_.p, Z2_.p, Y_.p, s_);
  launch_mlp_backward(P_.p, dX, dt, n, Z1_.p, Z2_.p, Y_.p, work_.p, part_.p, G_.p, s_);
}

float RankModel::gradient(const float* X, const float* t, int n, float* grad) {
  if (n <= 0) throw std::invalid_argument("no samples");
  ensure(n);
  HODLR_HIP_CHECK(hipMemcpyAsync(X_.p, X, (size_t)n * MLP_IN * sizeof(float), hipMemcpyHostToDevice, s_));
  HODLR_HIP_CHECK(hipMemcpyAsync(T_.p, t, n * sizeof(float), hipMemcpyHostToDevice, s_));
  fwd_bwd(X_.p, T_.p, n);
  std::vector<float> y(n);
  HODLR_HIP_CHECK(hipMemcpyAsync(grad, G_.p, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, s_));
  HODLR_HIP_CHECK(hipMemcpyAsync(y.data(), Y_.p, n * sizeof(float), hipMemcpyDeviceToHost, s_));
  HODLR_HIP_CHECK(hipStreamSynchronize(s_));
  double l = 0.0;
  for (int i = 0; i < n; i++) l += ((double)y[i] - t[i]) * ((double)y[i] - t[i]);
  return (float)(l / n);
}

void RankModel::train(const float* X, const float* t, int n, int epochs, int batch, float lr,
                      uint64_t seed, float* epoch_loss) {
  if (n <= 0 || epochs <= 0 || batch <= 0 || !(lr > 0.0f)) throw std::invalid_argument("bad training arguments");
  ensure(std::max(n, batch));
  DBuf<float> dX, dt;
  dX.alloc((size_t)n * MLP_IN);
  dt.alloc(n);
  std::vector<int> perm(n);
  std::vector<float> px((size_t)n * MLP_IN), pt(n), y(n);
  uint64_t st = seed;
  for (int e = 0; e < epochs; e++) {
    for (int i = 0; i < n; i++) perm[i] = i;
    for (int i = n - 1; i > 0; i--) {  // Fisher-Yates
      const int j = (int)(splitmix_next(st) % (uint64_t)(i + 1));
      std::swap(perm[i], perm[j]);
    }
    for (int i = 0; i < n; i++) {
      std::copy(X + (size_t)perm[i] * MLP_IN, X + (size_t)perm[i] * MLP_IN + MLP_IN, px.begin() + (size_t)i * MLP_IN);
      pt[i] = t[perm[i]];
    }
    HODLR_HIP_CHECK(hipMemcpyAsync(dX.p, px.data(), px.size() * sizeof(float), hipMemcpyHostToDevice, s_));
    HODLR_HIP_CHECK(hipMemcpyAsync(dt.p, pt.data(), pt.size() * sizeof(float), hipMemcpyHostToDevice, s_));
    for (int b0 = 0; b0 < n; b0 += batch) {
      const int nb = std::min(batch, n - b0);
      fwd_bwd(dX.p + (size_t)b0 * MLP_IN, dt.p + b0, nb);
      launch_mlp_adam(P_.p, G_.p, M_.p, V_.p, lr, 0.9f, 0.999f, 1e-8f, ++step_, s_);
    }
    if (epoch_loss) {
      launch_mlp_forward(P_.p, dX.p, n, nullptr, nullptr, Y_.p, s_);
      HODLR_HIP_CHECK(hipMemcpyAsync(y.data(), Y_.p, n * sizeof(float), hipMemcpyDeviceToHost, s_));
      HODLR_HIP_CHECK(hipStreamSynchronize(s_));
      double l = 0.0;
      for (int i = 0; i < n; i++) l += ((double)y[i] - pt[i]) * ((double)y[i] - pt[i]);
      epoch_loss[e] = (float)(l / n);
    }
  }
  HODLR_HIP_CHECK(hipStreamSynchronize(s_));
}

}  // namespace hodlr
