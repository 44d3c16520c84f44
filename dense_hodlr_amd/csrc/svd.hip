/*
 * One-sided Jacobi (Hestenes) SVD kernels, batched over all off-diagonal
 * blocks (see svd.h).  The rotation of a column pair (p < q) follows the
 * textbook one-sided Jacobi step: alpha = |a_p|^2, beta = |a_q|^2,
 * gamma = a_p . a_q; skipped when |gamma| <= eps sqrt(alpha beta) or when
 * |gamma| <= (eps * largest initial column norm)^2 (rounding noise);
 * zeta = (beta - alpha) / (2 gamma), t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)),
 * c = 1 / sqrt(1 + t^2), s = c t; a_p <- c a_p - s a_q, a_q <- s a_p + c a_q,
 * and the same rotation on the columns of V.  Sweeps repeat until a sweep
 * rotates nothing.
 */
#include "svd.h"

#include <algorithm>

namespace hodlr {

namespace {

constexpr int TPB = 256;

__device__ __forceinline__ int find_svd(const SvdDesc* __restrict__ d, int count) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].wg0 <= (int)blockIdx.x)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__global__ void __launch_bounds__(TPB)
    svd_step_kernel(const SvdDesc* __restrict__ descs, int count, int step,
                    unsigned* __restrict__ rotated) {
  __shared__ double sa[TPB], sb[TPB], sg[TPB];
  const SvdDesc d = descs[find_svd(descs, count)];
  const int ne = d.n + (d.n & 1);
  if (ne < 2 || step >= ne - 1) return;
  const int pr = (int)blockIdx.x - d.wg0;
  auto pos = [&](int k) { return k == 0 ? 0 : 1 + (k - 1 + step) % (ne - 1); };
  const int c1 = pos(pr), c2 = pos(ne - 1 - pr);
  const int cp = min(c1, c2), cq = max(c1, c2);
  if (cq >= d.n) return;  // the padding column of an odd n
  double* ap = d.A + (size_t)cp * d.m;
  double* aq = d.A + (size_t)cq * d.m;
  const int t = threadIdx.x;
  double a = 0.0, b = 0.0, g = 0.0;
  for (int i = t; i < d.m; i += TPB) {
    const double x = ap[i], y = aq[i];
    a += x * x;
    b += y * y;
    g += x * y;
  }
  sa[t] = a;
  sb[t] = b;
  sg[t] = g;
  __syncthreads();
  for (int w = TPB / 2; w > 0; w >>= 1) {
    if (t < w) {
      sa[t] += sa[t + w];
      sb[t] += sb[t + w];
      sg[t] += sg[t + w];
    }
    __syncthreads();
  }
  const double alpha = sa[0], beta = sb[0], gamma = sg[0];
  const double eps = 2.220446049250313e-16;
  if (gamma == 0.0 || fabs(gamma) <= eps * sqrt(alpha * beta) || fabs(gamma) <= d.floor2) return;
  const double zeta = (beta - alpha) / (2.0 * gamma);
  const double tt = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
  const double c = 1.0 / sqrt(1.0 + tt * tt), s = c * tt;
  for (int i = t; i < d.m; i += TPB) {
    const double x = ap[i], y = aq[i];
    ap[i] = c * x - s * y;
    aq[i] = s * x + c * y;
  }
  double* vp = d.V + (size_t)cp * d.n;
  double* vq = d.V + (size_t)cq * d.n;
  for (int i = t; i < d.n; i += TPB) {
    const double x = vp[i], y = vq[i];
    vp[i] = c * x - s * y;
    vq[i] = s * x + c * y;
  }
  if (t == 0) atomicAdd(rotated, 1u);
}

__global__ void __launch_bounds__(TPB)
    svd_norms_kernel(const SvdDesc* __restrict__ descs, const int* __restrict__ off,
                     double* __restrict__ out) {
  __shared__ double sh[TPB];
  const SvdDesc d = descs[blockIdx.y];
  const int c = blockIdx.x;
  if (c >= d.n) return;
  const double* a = d.A + (size_t)c * d.m;
  double s = 0.0;
  for (int i = threadIdx.x; i < d.m; i += TPB) s += a[i] * a[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = TPB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[off[blockIdx.y] + c] = sqrt(sh[0]);
}

__global__ void __launch_bounds__(TPB) col_copy_kernel(const ColCopyDesc* __restrict__ descs) {
  const ColCopyDesc d = descs[blockIdx.y];
  for (int i = blockIdx.x * TPB + threadIdx.x; i < d.rows; i += gridDim.x * TPB) {
    const double v = d.src[i];
    d.dst[i] = d.mode == 0 ? v * d.f : (d.f != 0.0 ? v / d.f : 0.0);
  }
}

}  // namespace

int svd_step_wgs(const SvdDesc& d) { return (d.n + (d.n & 1)) / 2; }

void launch_svd_step(const SvdDesc* d, int count, int nwg, int step, unsigned* rotated,
                     hipStream_t s) {
  if (count == 0 || nwg == 0) return;
  svd_step_kernel<<<nwg, TPB, 0, s>>>(d, count, step, rotated);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_svd_norms(const SvdDesc* d, int count, const int* off, int maxn, double* out,
                      hipStream_t s) {
  if (count == 0 || maxn == 0) return;
  for (int c0 = 0; c0 < count; c0 += 65535) {  // grid.y <= 65535
    dim3 g(maxn, std::min(65535, count - c0));
    svd_norms_kernel<<<g, TPB, 0, s>>>(d + c0, off + c0, out);
    HODLR_HIP_CHECK(hipGetLastError());
  }
}

void launch_col_copy(const ColCopyDesc* d, int count, hipStream_t s) {
  if (count == 0) return;
  for (int c0 = 0; c0 < count; c0 += 65535) {
    dim3 g(8, std::min(65535, count - c0));
    col_copy_kernel<<<g, TPB, 0, s>>>(d + c0);
    HODLR_HIP_CHECK(hipGetLastError());
  }
}

}  // namespace hodlr
