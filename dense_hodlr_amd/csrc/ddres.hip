/*
 * Double-double residual kernels (see ddres.h).
 *
 * Error-free transformations: TwoSum (Knuth) gives s + e = a + b exactly;
 * TwoProd via FMA gives p + e = a * b exactly.  A Dot2 accumulator keeps the
 * running sum in hi and the accumulated rounding errors in lo.
 *
 *   projection  T = V^T X per node half and level: 16 x 16 (c, j) output tile
 *               per workgroup over a chunk of DD_CHUNK rows, operands staged
 *               through LDS 64 rows at a time; chunk partials reduced in
 *               chunk order by a second kernel (deterministic).
 *   apply       Y += A T (leaves: A = leaf block, T = X; low rank: A = U,
 *               T = projection): one row per thread, 16 right-hand sides per
 *               workgroup, T staged through LDS 128 terms at a time.
 */
#include "ddres.h"

namespace hodlr {

namespace {

constexpr int TPB = 256;

__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
#pragma clang fp contract(off)
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

/* (sh, sl) += a * (th, tl) */
__device__ __forceinline__ void dot2_step(double a, double th, double tl, double& sh,
                                          double& sl) {
#pragma clang fp contract(off)
  const double p = a * th;
  const double pe = __builtin_fma(a, th, -p);
  double s, q;
  two_sum(sh, p, s, q);
  sh = s;
  sl += (q + pe) + a * tl;
}

template <class D>
__device__ __forceinline__ int find_desc(const D* __restrict__ d, int count, int key,
                                         int D::*field) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].*field <= key)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

constexpr int PT = 16;  // projection tile: PT columns c x PT right-hand sides j
constexpr int PR = 64;  // rows staged per step

__global__ void __launch_bounds__(TPB)
    dd_proj_kernel(const DdProjDesc* __restrict__ descs, int count, double* __restrict__ ph,
                   double* __restrict__ pl) {
  __shared__ double Vs[PR][PT + 1];
  __shared__ double Xs[PR][PT + 1];
  const DdProjDesc d = descs[find_desc(descs, count, (int)blockIdx.x, &DdProjDesc::wg0)];
  const int nct = (d.k + PT - 1) / PT, njt = (d.nrhs + PT - 1) / PT;
  int w = (int)blockIdx.x - d.wg0;
  const int chunk = w / (nct * njt);
  w -= chunk * nct * njt;
  const int c0 = (w % nct) * PT, j0 = (w / nct) * PT;
  const int t = threadIdx.x, c = t % PT, j = t / PT;
  const int rbeg = chunk * DD_CHUNK, rend = min(d.rows, rbeg + DD_CHUNK);
  double sh = 0.0, sl = 0.0;
  for (int r0 = rbeg; r0 < rend; r0 += PR) {
#pragma unroll
    for (int q = 0; q < PR * PT / TPB; q++) {
      const int e = t + q * TPB;
      const int rr = e % PR, cc = e / PR;
      const int gr = r0 + rr;
      Vs[rr][cc] = (gr < rend && c0 + cc < d.k) ? d.V[(size_t)gr + (size_t)(c0 + cc) * d.ldv] : 0.0;
      Xs[rr][cc] = (gr < rend && j0 + cc < d.nrhs) ? d.X[(size_t)gr + (size_t)(j0 + cc) * d.ldx] : 0.0;
    }
    __syncthreads();
    const int nr = min(PR, rend - r0);
    for (int rr = 0; rr < nr; rr++) dot2_step(Vs[rr][c], Xs[rr][j], 0.0, sh, sl);
    __syncthreads();
  }
  if (c0 + c < d.k && j0 + j < d.nrhs) {
    const size_t o = (size_t)d.part_off + ((size_t)chunk * d.nrhs + (j0 + j)) * d.k + (c0 + c);
    ph[o] = sh;
    pl[o] = sl;
  }
}

__global__ void __launch_bounds__(TPB)
    dd_reduce_kernel(const DdProjDesc* __restrict__ descs, int count,
                     const double* __restrict__ ph, const double* __restrict__ pl,
                     double* __restrict__ th, double* __restrict__ tl) {
  const DdProjDesc d = descs[find_desc(descs, count, (int)blockIdx.x, &DdProjDesc::rwg0)];
  const int e = ((int)blockIdx.x - d.rwg0) * TPB + threadIdx.x;
  const int kn = d.k * d.nrhs;
  if (e >= kn) return;
  double sh = 0.0, sl = 0.0;
  for (int ch = 0; ch < d.nchunks; ch++) {
    const size_t o = (size_t)d.part_off + (size_t)ch * kn + e;
    double s, q;
    two_sum(sh, ph[o], s, q);
    sh = s;
    sl += q + pl[o];
  }
  double h, l;
  two_sum(sh, sl, h, l);
  th[d.t_off + e] = h;
  tl[d.t_off + e] = l;
}

constexpr int AJ = 16;   // right-hand sides per apply workgroup
constexpr int AK = 128;  // terms staged per step

__global__ void __launch_bounds__(TPB)
    dd_apply_kernel(const DdApplyDesc* __restrict__ descs, int count) {
  __shared__ double Th[AK][AJ];
  __shared__ double Tl[AK][AJ];
  const DdApplyDesc d = descs[find_desc(descs, count, (int)blockIdx.x, &DdApplyDesc::wg0)];
  const int nrt = (d.rows + TPB - 1) / TPB;
  const int w = (int)blockIdx.x - d.wg0;
  const int row0 = (w % nrt) * TPB, j0 = (w / nrt) * AJ;
  const int t = threadIdx.x, row = row0 + t;
  const int nj = min(AJ, d.nrhs - j0);
  double sh[AJ], sl[AJ];
#pragma unroll
  for (int jj = 0; jj < AJ; jj++) {
    sh[jj] = 0.0;
    sl[jj] = 0.0;
  }
  for (int c0 = 0; c0 < d.k; c0 += AK) {
    const int nc = min(AK, d.k - c0);
#pragma unroll
    for (int q = 0; q < AK * AJ / TPB; q++) {
      const int e = t + q * TPB;
      const int cc = e % AK, jj = e / AK;
      const bool ok = cc < nc && jj < nj;
      const size_t o = (size_t)(c0 + cc) + (size_t)(j0 + jj) * d.ldt;
      Th[cc][jj] = ok ? d.t_hi[o] : 0.0;
      Tl[cc][jj] = (ok && d.t_lo) ? d.t_lo[o] : 0.0;
    }
    __syncthreads();
    if (row < d.rows) {
      const double* a = d.A + row + (size_t)c0 * d.lda;
      for (int cc = 0; cc < nc; cc++) {
        const double av = a[(size_t)cc * d.lda];
#pragma unroll
        for (int jj = 0; jj < AJ; jj++) dot2_step(av, Th[cc][jj], Tl[cc][jj], sh[jj], sl[jj]);
      }
    }
    __syncthreads();
  }
  if (row >= d.rows) return;
#pragma unroll
  for (int jj = 0; jj < AJ; jj++) {
    if (jj >= nj) break;
    const size_t o = (size_t)row + (size_t)(j0 + jj) * d.ldy;
    double h = sh[jj], l = sl[jj];
    if (d.accumulate) {
      double s, q;
      two_sum(d.y_hi[o], h, s, q);
      h = s;
      l = q + (l + d.y_lo[o]);
    }
    double hh, ll;
    two_sum(h, l, hh, ll);
    d.y_hi[o] = hh;
    d.y_lo[o] = ll;
  }
}

__global__ void __launch_bounds__(TPB)
    dd_residual_kernel(const double* __restrict__ B, int ldb, const double* __restrict__ yh,
                       const double* __restrict__ yl, int ldy, double* __restrict__ R, int ldr,
                       int n, int ncols) {
  const size_t tot = (size_t)n * ncols;
  for (size_t e = (size_t)blockIdx.x * TPB + threadIdx.x; e < tot; e += (size_t)gridDim.x * TPB) {
    const int i = (int)(e % n), j = (int)(e / n);
    const size_t oy = (size_t)i + (size_t)j * ldy;
    double s, q;
    two_sum(B[(size_t)i + (size_t)j * ldb], -yh[oy], s, q);
    R[(size_t)i + (size_t)j * ldr] = s + (q - yl[oy]);
  }
}

}  // namespace

int dd_proj_wgs(const DdProjDesc& d) {
  return d.nchunks * ((d.k + PT - 1) / PT) * ((d.nrhs + PT - 1) / PT);
}
int dd_reduce_wgs(const DdProjDesc& d) { return (d.k * d.nrhs + TPB - 1) / TPB; }
int dd_apply_wgs(const DdApplyDesc& d) {
  return ((d.rows + TPB - 1) / TPB) * ((d.nrhs + AJ - 1) / AJ);
}

void launch_dd_proj(const DdProjDesc* d, int count, int nwg, double* part_hi, double* part_lo,
                    hipStream_t s) {
  if (count == 0 || nwg == 0) return;
  dd_proj_kernel<<<nwg, TPB, 0, s>>>(d, count, part_hi, part_lo);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_dd_reduce(const DdProjDesc* d, int count, int nwg, const double* part_hi,
                      const double* part_lo, double* t_hi, double* t_lo, hipStream_t s) {
  if (count == 0 || nwg == 0) return;
  dd_reduce_kernel<<<nwg, TPB, 0, s>>>(d, count, part_hi, part_lo, t_hi, t_lo);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_dd_apply(const DdApplyDesc* d, int count, int nwg, hipStream_t s) {
  if (count == 0 || nwg == 0) return;
  dd_apply_kernel<<<nwg, TPB, 0, s>>>(d, count);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_dd_residual(const double* B, int ldb, const double* y_hi, const double* y_lo, int ldy,
                        double* R, int ldr, int n, int ncols, hipStream_t s) {
  if ((size_t)n * ncols == 0) return;
  dd_residual_kernel<<<2048, TPB, 0, s>>>(B, ldb, y_hi, y_lo, ldy, R, ldr, n, ncols);
  HODLR_HIP_CHECK(hipGetLastError());
}

}  // namespace hodlr
