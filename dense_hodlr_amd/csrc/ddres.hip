/*
 * Double-double residual kernels (see ddres.h).
 *
 * Error-free transformations: TwoSum (Knuth) gives s + e = a + b exactly;
 * TwoProd via FMA gives p + e = a * b exactly.  A Dot2 accumulator keeps the
 * running sum in hi and the accumulated rounding errors in lo.
 *
 *   projection  T = V^T X per node half and level: a 64 x 64 (c, j) output
 *               tile per workgroup (64 x 16 or 256 x 4 for few right-hand
 *               sides), 4 x 4 (or 4 x 1) outputs per thread, over a chunk of
 *               DD_CHUNK rows; operands staged through LDS 32 rows at a time
 *               with the next rows' loads in flight.  X is exact, so each
 *               term is TwoProd + TwoSum.  Chunk partials are reduced in chunk
 *               order by a second kernel (deterministic).
 *   apply       Y += A T (leaves: A = leaf block, T = X; low rank: A = U,
 *               T = projection): one row per thread, 16 right-hand sides per
 *               workgroup (4 for fewer), T staged through LDS 128 terms at a
 *               time.
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

/* (sh, sl) += a * b, b exact (no low part) */
__device__ __forceinline__ void dot2_step_exact(double a, double b, double& sh, double& sl) {
#pragma clang fp contract(off)
  const double p = a * b;
  const double pe = __builtin_fma(a, b, -p);
  double s, q;
  two_sum(sh, p, s, q);
  sh = s;
  sl += q + pe;
}

/* Projection tile shapes: TC threads along c, TPB / TC along j; each thread
 * owns CA x JB outputs (c = c0 + tc + TC a, j = j0 + tj + (TPB / TC) b), so a
 * row step reads CA + JB operands from LDS for CA * JB products.  PR rows are
 * staged per step, the next step's global loads in flight meanwhile. */
template <int TC, int CA, int JB>
struct ProjShape {
  static constexpr int TJ = TPB / TC;
  static constexpr int CT = TC * CA;  // c tile
  static constexpr int JT = TJ * JB;  // j tile
  static constexpr int PR = CT + JT >= 128 ? 16 : 32;
};
using ProjWide = ProjShape<16, 4, 4>;   // 64 x 64, many right-hand sides
using ProjMid = ProjShape<16, 4, 1>;    // 64 x 16
using ProjNarrow = ProjShape<64, 4, 1>; // 256 x 4, up to 4 right-hand sides

template <int TC, int CA, int JB>
__global__ void __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(3)))
    dd_proj_kernel(const DdProjDesc* __restrict__ descs, int count, double* __restrict__ ph,
                   double* __restrict__ pl) {
  using S = ProjShape<TC, CA, JB>;
  constexpr int CT = S::CT, JT = S::JT, PR = S::PR, TJ = S::TJ;
  constexpr int VL = PR * CT / TPB, XL = (PR * JT + TPB - 1) / TPB;
  __shared__ double Vs[PR][CT + 1];
  __shared__ double Xs[PR][JT + 1];
  const DdProjDesc d = descs[find_desc(descs, count, (int)blockIdx.x, &DdProjDesc::wg0)];
  const gptr<const double> dV = to_global(d.V), dX = to_global(d.X);
  const int nct = (d.k + CT - 1) / CT, njt = (d.nrhs + JT - 1) / JT;
  int w = (int)blockIdx.x - d.wg0;
  const int chunk = w / (nct * njt);
  w -= chunk * nct * njt;
  const int c0 = (w % nct) * CT, j0 = (w / nct) * JT;
  const int t = threadIdx.x, tc = t % TC, tj = t / TC;
  const int rbeg = chunk * DD_CHUNK, rend = min(d.rows, rbeg + DD_CHUNK);
  double sh[CA][JB], sl[CA][JB];
#pragma unroll
  for (int a = 0; a < CA; a++)
#pragma unroll
    for (int b = 0; b < JB; b++) {
      sh[a][b] = 0.0;
      sl[a][b] = 0.0;
    }
  double rv[VL], rx[XL];
  auto load = [&](int r0) {
#pragma unroll
    for (int q = 0; q < VL; q++) {
      const int e = t + q * TPB;
      const int rr = e % PR, cc = e / PR, gr = r0 + rr;
      rv[q] = (gr < rend && c0 + cc < d.k) ? dV[(size_t)gr + (size_t)(c0 + cc) * d.ldv] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < XL; q++) {
      const int e = t + q * TPB;
      const int rr = e % PR, cc = e / PR, gr = r0 + rr;
      rx[q] = (e < PR * JT && gr < rend && j0 + cc < d.nrhs)
                  ? dX[(size_t)gr + (size_t)(j0 + cc) * d.ldx]
                  : 0.0;
    }
  };
  if (rbeg < rend) load(rbeg);
  for (int r0 = rbeg; r0 < rend; r0 += PR) {
#pragma unroll
    for (int q = 0; q < VL; q++) {
      const int e = t + q * TPB;
      Vs[e % PR][e / PR] = rv[q];
    }
#pragma unroll
    for (int q = 0; q < XL; q++) {
      const int e = t + q * TPB;
      if (e < PR * JT) Xs[e % PR][e / PR] = rx[q];
    }
    __syncthreads();
    if (r0 + PR < rend) load(r0 + PR);
    const int nr = min(PR, rend - r0);
    // two rows per trip: the pair (sh, sl) alternates between two register
    // sets instead of being copied back every row
#pragma unroll 2
    for (int rr = 0; rr < nr; rr++) {
      double va[CA], xb[JB];
#pragma unroll
      for (int a = 0; a < CA; a++) va[a] = Vs[rr][tc + TC * a];
#pragma unroll
      for (int b = 0; b < JB; b++) xb[b] = Xs[rr][tj + TJ * b];
#pragma unroll
      for (int a = 0; a < CA; a++)
#pragma unroll
        for (int b = 0; b < JB; b++) dot2_step_exact(va[a], xb[b], sh[a][b], sl[a][b]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < CA; a++)
#pragma unroll
    for (int b = 0; b < JB; b++) {
      const int c = c0 + tc + TC * a, j = j0 + tj + TJ * b;
      if (c < d.k && j < d.nrhs) {
        const size_t o = (size_t)d.part_off + ((size_t)chunk * d.nrhs + j) * d.k + c;
        ph[o] = sh[a][b];
        pl[o] = sl[a][b];
      }
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

constexpr int AK = 128;  // terms staged per step

/* (sh, sl) += a * (th, tl): the product a * th exactly (TwoProd), its sum
 * with sh exactly (TwoSum); the low-order terms, a * tl included, are
 * accumulated in sl with a fused multiply-add (their own rounding is far
 * below the pair's precision). */
__device__ __forceinline__ void dot2_step_lo(double a, double th, double tl, double& sh,
                                             double& sl) {
  double p, pe, s, q;
  {
#pragma clang fp contract(off)
    p = a * th;
    pe = __builtin_fma(a, th, -p);
    two_sum(sh, p, s, q);
  }
  sh = s;
  sl = __builtin_fma(a, tl, sl + (q + pe));
}

template <int AJ>
__global__ void __launch_bounds__(TPB)
    dd_apply_kernel(const DdApplyDesc* __restrict__ descs, int count) {
  __shared__ double Th[AK][AJ];
  __shared__ double Tl[AK][AJ];
  const DdApplyDesc d = descs[find_desc(descs, count, (int)blockIdx.x, &DdApplyDesc::wg0)];
  const gptr<const double> tH = to_global(d.t_hi), tL = to_global(d.t_lo);
  const gptr<double> yH = to_global(d.y_hi), yL = to_global(d.y_lo);
  const int w = (int)blockIdx.x - d.wg0;
  // the right-hand-side groups of one row tile are consecutive workgroups,
  // which run together and share the tile's rows of A through the L2
  const int njt = (d.nrhs + AJ - 1) / AJ;
  const int row0 = (w / njt) * TPB, j0 = (w % njt) * AJ;
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
    for (int q = 0; q < (AK * AJ + TPB - 1) / TPB; q++) {
      const int e = t + q * TPB;
      if (e < AK * AJ) {
        const int cc = e % AK, jj = e / AK;
        const bool ok = cc < nc && jj < nj;
        const size_t o = (size_t)(c0 + cc) + (size_t)(j0 + jj) * d.ldt;
        Th[cc][jj] = ok ? tH[o] : 0.0;
        Tl[cc][jj] = (ok && d.t_lo) ? tL[o] : 0.0;
      }
    }
    __syncthreads();
    if (row < d.rows) {
      // the multipliers of four terms in registers while the next four load
      // (a load per term would wait out a memory latency every term).  The
      // terms in order; the group's terms past nc have a zero multiplier and
      // zero staged terms, which leave (sh, sl) exactly unchanged (neither
      // is ever -0.0: both start at +0.0 and round to nearest), so the
      // groups need no test per term and the pairs alternate between
      // register sets instead of being copied back every term.
      static_assert(AK % 4 == 0, "staged terms in groups of four");
      const gptr<const double> a = to_global(d.A) + row + (size_t)c0 * d.lda;
      auto load4 = [&](int c, double(&v)[4]) {
#pragma unroll
        for (int u = 0; u < 4; u++) v[u] = c + u < nc ? a[(size_t)(c + u) * d.lda] : 0.0;
      };
      double cur[4], nxt[4];
      load4(0, cur);
      for (int cc = 0; cc < nc; cc += 4) {
        load4(cc + 4, nxt);
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
          for (int jj = 0; jj < AJ; jj++)
            dot2_step_lo(cur[u], Th[cc + u][jj], Tl[cc + u][jj], sh[jj], sl[jj]);
#pragma unroll
        for (int u = 0; u < 4; u++) cur[u] = nxt[u];
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
      two_sum(yH[o], h, s, q);
      h = s;
      l = q + (l + yL[o]);
    }
    double hh, ll;
    two_sum(h, l, hh, ll);
    yH[o] = hh;
    yL[o] = ll;
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

namespace {
// tile shape by right-hand side count (uniform over one launch)
template <class F>
auto with_proj_shape(int nrhs, F&& f) {
  if (nrhs > 16) return f(ProjWide{});
  if (nrhs > 4) return f(ProjMid{});
  return f(ProjNarrow{});
}
int apply_cols(int nrhs) { return nrhs > 4 ? 16 : 4; }
}  // namespace

int dd_proj_wgs(const DdProjDesc& d) {
  return with_proj_shape(d.nrhs, [&](auto sh) {
    using S = decltype(sh);
    return d.nchunks * ((d.k + S::CT - 1) / S::CT) * ((d.nrhs + S::JT - 1) / S::JT);
  });
}
int dd_reduce_wgs(const DdProjDesc& d) { return (d.k * d.nrhs + TPB - 1) / TPB; }
int dd_apply_wgs(const DdApplyDesc& d) {
  const int aj = apply_cols(d.nrhs);
  return ((d.rows + TPB - 1) / TPB) * ((d.nrhs + aj - 1) / aj);
}

void launch_dd_proj(const DdProjDesc* d, int count, int nwg, int nrhs, double* part_hi,
                    double* part_lo, hipStream_t s) {
  if (count == 0 || nwg == 0) return;
  if (nrhs > 16)
    dd_proj_kernel<16, 4, 4><<<nwg, TPB, 0, s>>>(d, count, part_hi, part_lo);
  else if (nrhs > 4)
    dd_proj_kernel<16, 4, 1><<<nwg, TPB, 0, s>>>(d, count, part_hi, part_lo);
  else
    dd_proj_kernel<64, 4, 1><<<nwg, TPB, 0, s>>>(d, count, part_hi, part_lo);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_dd_reduce(const DdProjDesc* d, int count, int nwg, const double* part_hi,
                      const double* part_lo, double* t_hi, double* t_lo, hipStream_t s) {
  if (count == 0 || nwg == 0) return;
  dd_reduce_kernel<<<nwg, TPB, 0, s>>>(d, count, part_hi, part_lo, t_hi, t_lo);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_dd_apply(const DdApplyDesc* d, int count, int nwg, int nrhs, hipStream_t s) {
  if (count == 0 || nwg == 0) return;
  const int aj = apply_cols(nrhs);
  if (aj == 16)
    dd_apply_kernel<16><<<nwg, TPB, 0, s>>>(d, count);
  else
    dd_apply_kernel<4><<<nwg, TPB, 0, s>>>(d, count);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_dd_residual(const double* B, int ldb, const double* y_hi, const double* y_lo, int ldy,
                        double* R, int ldr, int n, int ncols, hipStream_t s) {
  if ((size_t)n * ncols == 0) return;
  dd_residual_kernel<<<2048, TPB, 0, s>>>(B, ldb, y_hi, y_lo, ldy, R, ldr, n, ncols);
  HODLR_HIP_CHECK(hipGetLastError());
}

}  // namespace hodlr
