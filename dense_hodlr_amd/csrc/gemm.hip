/*
 * Grouped FP64 GEMM: every product of a tree level (all nodes, both halves)
 * in one launch.  C = alpha * op(A) * B + beta * C with op(A) = A or A^T.
 *
 * 64 x 64 output tiles, K staged through LDS in slices of 16, 256 threads with
 * a 4 x 4 register tile each.  Problems whose reduction dimension is long and
 * whose output is small (the V^T X products of the upper tree levels) are
 * split along K; the slices land in a workspace and a second kernel sums them
 * in slice order, so results do not depend on scheduling.
 */
#include "gemm.h"

#include <algorithm>

namespace hodlr {

namespace {

constexpr int TPB = 256;
constexpr int BM = 64, BN = 64, BK = 16;

__global__ void __launch_bounds__(TPB)
    gemm_kernel(const GemmDesc* __restrict__ descs, int count, double* __restrict__ ws) {
  __shared__ double As[BK][BM + 1];
  __shared__ double Bs[BK][BN + 1];
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (descs[mid].tile0 <= (int)blockIdx.x)
      lo = mid;
    else
      hi = mid - 1;
  }
  const GemmDesc g = descs[lo];
  const int tm = (g.M + BM - 1) / BM, tn = (g.N + BN - 1) / BN;
  int tile = blockIdx.x - g.tile0;
  const int split = tile / (tm * tn);
  tile -= split * tm * tn;
  const int m0 = (tile % tm) * BM, n0 = (tile / tm) * BN;
  const int kchunk = ((g.K + g.ksplit - 1) / g.ksplit + BK - 1) / BK * BK;
  const int kbeg = split * kchunk, kend = min(g.K, kbeg + kchunk);
  const int t = threadIdx.x, tx = t & 15, ty = t >> 4;
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) acc[a][b] = 0.0;

  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    // A tile: op(A)(m0 + mm, k0 + kk)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int e = t + r * TPB;
      int mm, kk;
      if (g.trans_a) {
        kk = e & 15;
        mm = e >> 4;
      } else {
        mm = e & 63;
        kk = e >> 6;
      }
      const int gm = m0 + mm, gk = k0 + kk;
      double v = 0.0;
      if (gm < g.M && gk < kend)
        v = g.trans_a ? g.A[(size_t)gk + (size_t)gm * g.lda] : g.A[(size_t)gm + (size_t)gk * g.lda];
      As[kk][mm] = v;
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int e = t + r * TPB;
      const int kk = e & 15, nn = e >> 4;
      const int gn = n0 + nn, gk = k0 + kk;
      Bs[kk][nn] = (gn < g.N && gk < kend) ? g.B[(size_t)gk + (size_t)gn * g.ldb] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk++) {
      double av[4], bv[4];
#pragma unroll
      for (int a = 0; a < 4; a++) av[a] = As[kk][tx + 16 * a];
#pragma unroll
      for (int b = 0; b < 4; b++) bv[b] = Bs[kk][ty + 16 * b];
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = fma(av[a], bv[b], acc[a][b]);
    }
    __syncthreads();
  }
  if (g.ksplit == 1) {
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b = 0; b < 4; b++) {
        const int gm = m0 + tx + 16 * a, gn = n0 + ty + 16 * b;
        if (gm < g.M && gn < g.N) {
          double* c = g.C + (size_t)gm + (size_t)gn * g.ldc;
          *c = (g.beta == 0.0) ? g.alpha * acc[a][b] : g.alpha * acc[a][b] + g.beta * *c;
        }
      }
  } else {
    double* w = ws + g.ws_off + (size_t)split * g.M * g.N;
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b = 0; b < 4; b++) {
        const int gm = m0 + tx + 16 * a, gn = n0 + ty + 16 * b;
        if (gm < g.M && gn < g.N) w[(size_t)gm + (size_t)gn * g.M] = acc[a][b];
      }
  }
}

__global__ void __launch_bounds__(TPB)
    splitk_reduce_kernel(const GemmDesc* __restrict__ descs, const int* __restrict__ ids,
                         const double* __restrict__ ws) {
  const GemmDesc g = descs[ids[blockIdx.y]];
  const size_t mn = (size_t)g.M * g.N;
  for (size_t e = (size_t)blockIdx.x * TPB + threadIdx.x; e < mn; e += (size_t)gridDim.x * TPB) {
    double s = 0.0;
    for (int sp = 0; sp < g.ksplit; sp++) s += ws[g.ws_off + (size_t)sp * mn + e];
    const int gm = (int)(e % g.M), gn = (int)(e / g.M);
    double* c = g.C + (size_t)gm + (size_t)gn * g.ldc;
    *c = (g.beta == 0.0) ? g.alpha * s : g.alpha * s + g.beta * *c;
  }
}

}  // namespace

size_t gemm_plan(std::vector<GemmDesc>& descs, std::vector<int>& split_ids, int& ntiles,
                 int target_tiles) {
  // Choose K splits so that problems with few output tiles and a long K are
  // spread over more workgroups; assign tile and workspace offsets.
  size_t tiles_out = 0;
  for (auto& g : descs) {
    const size_t tm = (g.M + BM - 1) / BM, tn = (g.N + BN - 1) / BN;
    tiles_out += tm * tn;
  }
  size_t ws = 0;
  int tile = 0;
  split_ids.clear();
  for (size_t p = 0; p < descs.size(); p++) {
    GemmDesc& g = descs[p];
    const int tm = (g.M + BM - 1) / BM, tn = (g.N + BN - 1) / BN;
    int ks = 1;
    if (g.K >= 4 * 256 && tiles_out < (size_t)target_tiles) {
      const int want = (int)std::min<size_t>(64, (target_tiles + tiles_out - 1) / std::max<size_t>(tiles_out, 1));
      ks = std::max(1, std::min(want, g.K / 256));
    }
    g.ksplit = ks;
    g.tile0 = tile;
    tile += tm * tn * ks;
    if (ks > 1) {
      g.ws_off = (long long)ws;
      ws += (size_t)ks * g.M * g.N;
      split_ids.push_back((int)p);
    } else {
      g.ws_off = 0;
    }
  }
  ntiles = tile;
  return ws;
}

void launch_gemm(const GemmDesc* d_descs, int count, int ntiles, double* d_ws,
                 const int* d_split_ids, int nsplit, hipStream_t s) {
  if (count == 0 || ntiles == 0) return;
  gemm_kernel<<<ntiles, TPB, 0, s>>>(d_descs, count, d_ws);
  HODLR_HIP_CHECK(hipGetLastError());
  if (nsplit > 0) {
    dim3 g(64, nsplit);
    splitk_reduce_kernel<<<g, TPB, 0, s>>>(d_descs, d_split_ids, d_ws);
    HODLR_HIP_CHECK(hipGetLastError());
  }
}

}  // namespace hodlr
