/*
 * Rank-predictor MLP kernels (see mlp.h).  Sizes are small (4801
 * parameters, up to a few hundred thousand samples), so the kernels favour
 * determinism over peak rate: one sample per thread with the parameters in
 * LDS for the forward and delta passes, and gradient partials per chunk of
 * samples summed in chunk order (no atomics), so repeated runs give
 * bit-identical gradients.
 */
#include "mlp.h"

namespace hodlr {

namespace {

constexpr int TPB = 256;

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }

__device__ __forceinline__ float gelu_grad(float x) {
  const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
  return cdf + x * pdf;
}

/* x rounded to the nearest bfloat16 (ties to even), returned as float */
__device__ __forceinline__ float bf16r(float x) {
  unsigned u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return x;  // inf / nan unchanged
  u += 0x7fffu + ((u >> 16) & 1u);
  return __uint_as_float(u & 0xffff0000u);
}

template <bool BF16>
__device__ __forceinline__ void load_params(const float* __restrict__ P, float* sp) {
  for (int i = threadIdx.x; i < MLP_NPARAM; i += TPB) sp[i] = BF16 ? bf16r(P[i]) : P[i];
  __syncthreads();
}

template <bool BF16>
__global__ void __launch_bounds__(TPB)
    mlp_forward_kernel(const float* __restrict__ P, const float* __restrict__ X, int n,
                       float* __restrict__ Z1, float* __restrict__ Z2, float* __restrict__ y) {
  __shared__ float sp[MLP_NPARAM];
  load_params<BF16>(P, sp);
  const int s = blockIdx.x * TPB + threadIdx.x;
  if (s >= n) return;
  float x[MLP_IN];
#pragma unroll
  for (int k = 0; k < MLP_IN; k++) x[k] = BF16 ? bf16r(X[(size_t)s * MLP_IN + k]) : X[(size_t)s * MLP_IN + k];
  float a1[MLP_H];
#pragma unroll
  for (int j = 0; j < MLP_H; j++) {
    float z = sp[MLP_OFF_B1 + j];
#pragma unroll
    for (int k = 0; k < MLP_IN; k++) z = fmaf(sp[MLP_OFF_W1 + j * MLP_IN + k], x[k], z);
    if (Z1) Z1[(size_t)j * n + s] = z;
    a1[j] = BF16 ? bf16r(gelu(z)) : gelu(z);
  }
  float out = sp[MLP_OFF_B3];
  for (int i = 0; i < MLP_H; i++) {
    float z = sp[MLP_OFF_B2 + i];
#pragma unroll
    for (int j = 0; j < MLP_H; j++) z = fmaf(sp[MLP_OFF_W2 + i * MLP_H + j], a1[j], z);
    if (Z2) Z2[(size_t)i * n + s] = z;
    const float a2 = BF16 ? bf16r(gelu(z)) : gelu(z);
    out = fmaf(sp[MLP_OFF_W3 + i], a2, out);
  }
  y[s] = out;
}

/* per-sample deltas and post-activations; work = [D1 | D2 | A1 | A2] (64 x n
 * each, feature-major) followed by D3 (n) */
__global__ void __launch_bounds__(TPB)
    mlp_delta_kernel(const float* __restrict__ P, const float* __restrict__ t, int n,
                     const float* __restrict__ Z1, const float* __restrict__ Z2,
                     const float* __restrict__ y, float* __restrict__ work) {
  __shared__ float sp[MLP_NPARAM];
  load_params<false>(P, sp);
  const int s = blockIdx.x * TPB + threadIdx.x;
  if (s >= n) return;
  float* D1 = work;
  float* D2 = work + (size_t)MLP_H * n;
  float* A1 = work + (size_t)2 * MLP_H * n;
  float* A2 = work + (size_t)3 * MLP_H * n;
  float* D3 = work + (size_t)4 * MLP_H * n;
  const float d3 = 2.0f * (y[s] - t[s]) / (float)n;  // d mean((y-t)^2) / dy
  D3[s] = d3;
  float d2[MLP_H];
#pragma unroll
  for (int i = 0; i < MLP_H; i++) {
    const float z = Z2[(size_t)i * n + s];
    A2[(size_t)i * n + s] = gelu(z);
    d2[i] = d3 * sp[MLP_OFF_W3 + i] * gelu_grad(z);
    D2[(size_t)i * n + s] = d2[i];
  }
  for (int j = 0; j < MLP_H; j++) {
    float acc = 0.0f;
#pragma unroll
    for (int i = 0; i < MLP_H; i++) acc = fmaf(sp[MLP_OFF_W2 + i * MLP_H + j], d2[i], acc);
    const float z = Z1[(size_t)j * n + s];
    A1[(size_t)j * n + s] = gelu(z);
    D1[(size_t)j * n + s] = acc * gelu_grad(z);
  }
}

/* gradient partial of parameter p over samples of chunk blockIdx.x */
__global__ void __launch_bounds__(TPB)
    mlp_grad_part_kernel(const float* __restrict__ X, int n, const float* __restrict__ work,
                         float* __restrict__ part) {
  const int p = blockIdx.y * TPB + threadIdx.x;
  if (p >= MLP_NPARAM) return;
  const int c = blockIdx.x;
  const int s0 = c * MLP_CHUNK, s1 = min(n, s0 + MLP_CHUNK);
  const float* D1 = work;
  const float* D2 = work + (size_t)MLP_H * n;
  const float* A1 = work + (size_t)2 * MLP_H * n;
  const float* A2 = work + (size_t)3 * MLP_H * n;
  const float* D3 = work + (size_t)4 * MLP_H * n;
  float acc = 0.0f;
  if (p < MLP_OFF_B1) {  // W1[j][k]
    const int j = p / MLP_IN, k = p % MLP_IN;
    for (int s = s0; s < s1; s++) acc = fmaf(D1[(size_t)j * n + s], X[(size_t)s * MLP_IN + k], acc);
  } else if (p < MLP_OFF_W2) {  // b1[j]
    const int j = p - MLP_OFF_B1;
    for (int s = s0; s < s1; s++) acc += D1[(size_t)j * n + s];
  } else if (p < MLP_OFF_B2) {  // W2[i][j]
    const int q = p - MLP_OFF_W2, i = q / MLP_H, j = q % MLP_H;
    for (int s = s0; s < s1; s++) acc = fmaf(D2[(size_t)i * n + s], A1[(size_t)j * n + s], acc);
  } else if (p < MLP_OFF_W3) {  // b2[i]
    const int i = p - MLP_OFF_B2;
    for (int s = s0; s < s1; s++) acc += D2[(size_t)i * n + s];
  } else if (p < MLP_OFF_B3) {  // W3[i]
    const int i = p - MLP_OFF_W3;
    for (int s = s0; s < s1; s++) acc = fmaf(D3[s], A2[(size_t)i * n + s], acc);
  } else {  // b3
    for (int s = s0; s < s1; s++) acc += D3[s];
  }
  part[(size_t)c * MLP_NPARAM + p] = acc;
}

__global__ void __launch_bounds__(TPB)
    mlp_grad_reduce_kernel(const float* __restrict__ part, int nchunks, float* __restrict__ G) {
  const int p = blockIdx.x * TPB + threadIdx.x;
  if (p >= MLP_NPARAM) return;
  float acc = 0.0f;
  for (int c = 0; c < nchunks; c++) acc += part[(size_t)c * MLP_NPARAM + p];
  G[p] = acc;
}

__global__ void __launch_bounds__(TPB)
    mlp_adam_kernel(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ M,
                    float* __restrict__ V, float lr, float b1, float b2, float eps, float c1,
                    float c2) {
  const int p = blockIdx.x * TPB + threadIdx.x;
  if (p >= MLP_NPARAM) return;
  const float g = G[p];
  const float m = b1 * M[p] + (1.0f - b1) * g;
  const float v = b2 * V[p] + (1.0f - b2) * g * g;
  M[p] = m;
  V[p] = v;
  P[p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

__global__ void __launch_bounds__(TPB) block_features_kernel(const FeatDesc* __restrict__ descs,
                                                             float* __restrict__ out) {
  __shared__ double red[6][TPB];
  __shared__ double cen[6];
  const FeatDesc d = descs[blockIdx.x];
  const int t = threadIdx.x;
  auto coord = [&](int i, int a) { return a == 0 ? d.x[i] : (a == 1 ? d.y[i] : d.z[i]); };
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int i = t; i < d.nr; i += TPB)
    for (int a = 0; a < d.dim; a++) acc[a] += coord(d.r0 + i, a);
  for (int i = t; i < d.nc; i += TPB)
    for (int a = 0; a < d.dim; a++) acc[3 + a] += coord(d.c0 + i, a);
  for (int q = 0; q < 6; q++) red[q][t] = acc[q];
  __syncthreads();
  for (int w = TPB / 2; w > 0; w >>= 1) {
    if (t < w)
      for (int q = 0; q < 6; q++) red[q][t] += red[q][t + w];
    __syncthreads();
  }
  if (t < 6) cen[t] = red[t][0] / (double)(t < 3 ? d.nr : d.nc);
  __syncthreads();
  double dr = 0.0, dc = 0.0;
  for (int i = t; i < d.nr; i += TPB) {
    double s = 0.0;
    for (int a = 0; a < d.dim; a++) {
      const double v = coord(d.r0 + i, a) - cen[a];
      s += v * v;
    }
    dr = fmax(dr, s);
  }
  for (int i = t; i < d.nc; i += TPB) {
    double s = 0.0;
    for (int a = 0; a < d.dim; a++) {
      const double v = coord(d.c0 + i, a) - cen[3 + a];
      s += v * v;
    }
    dc = fmax(dc, s);
  }
  __syncthreads();
  red[0][t] = dr;
  red[1][t] = dc;
  __syncthreads();
  for (int w = TPB / 2; w > 0; w >>= 1) {
    if (t < w) {
      red[0][t] = fmax(red[0][t], red[0][t + w]);
      red[1][t] = fmax(red[1][t], red[1][t + w]);
    }
    __syncthreads();
  }
  if (t == 0) {
    const double rr = sqrt(red[0][0]), rc = sqrt(red[1][0]);
    double sep = 0.0;
    for (int a = 0; a < d.dim; a++) sep += (cen[a] - cen[3 + a]) * (cen[a] - cen[3 + a]);
    sep = sqrt(sep);
    const double l = d.length_scale;
    float* o = out + (size_t)blockIdx.x * MLP_IN;
    o[0] = (float)log((double)(d.nr + d.nc));
    o[1] = (float)(rr + rc > 0.0 ? sep / (rr + rc) : 0.0);
    o[2] = (float)(rr / l);
    o[3] = (float)(rc / l);
    o[4] = (float)log(l);
    o[5] = d.log10_tol;
    o[6] = (float)d.level;
    o[7] = d.family;
  }
}

}  // namespace

void launch_mlp_forward(const float* P, const float* X, int n, float* Z1, float* Z2, float* y,
                        hipStream_t s) {
  if (n <= 0) return;
  mlp_forward_kernel<false><<<(n + TPB - 1) / TPB, TPB, 0, s>>>(P, X, n, Z1, Z2, y);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_mlp_infer_bf16(const float* P, const float* X, int n, float* y, hipStream_t s) {
  if (n <= 0) return;
  mlp_forward_kernel<true><<<(n + TPB - 1) / TPB, TPB, 0, s>>>(P, X, n, nullptr, nullptr, y);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_mlp_backward(const float* P, const float* X, const float* t, int n, const float* Z1,
                         const float* Z2, const float* y, float* work, float* part, float* G,
                         hipStream_t s) {
  if (n <= 0) return;
  mlp_delta_kernel<<<(n + TPB - 1) / TPB, TPB, 0, s>>>(P, t, n, Z1, Z2, y, work);
  HODLR_HIP_CHECK(hipGetLastError());
  const int nchunks = (n + MLP_CHUNK - 1) / MLP_CHUNK;
  dim3 g(nchunks, (MLP_NPARAM + TPB - 1) / TPB);
  mlp_grad_part_kernel<<<g, TPB, 0, s>>>(X, n, work, part);
  HODLR_HIP_CHECK(hipGetLastError());
  mlp_grad_reduce_kernel<<<(MLP_NPARAM + TPB - 1) / TPB, TPB, 0, s>>>(part, nchunks, G);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_mlp_adam(float* P, const float* G, float* M, float* V, float lr, float b1, float b2,
                     float eps, int t, hipStream_t s) {
  const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
  mlp_adam_kernel<<<(MLP_NPARAM + TPB - 1) / TPB, TPB, 0, s>>>(P, G, M, V, lr, b1, b2, eps, c1, c2);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_block_features(const FeatDesc* d, int count, float* out, hipStream_t s) {
  if (count <= 0) return;
  block_features_kernel<<<count, TPB, 0, s>>>(d, out);
  HODLR_HIP_CHECK(hipGetLastError());
}

}  // namespace hodlr
