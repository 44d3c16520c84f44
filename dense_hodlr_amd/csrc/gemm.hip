/*
 * Grouped FP64 GEMM: every product of a tree level (all nodes, both halves)
 * in one launch.  C = alpha * op(A) * B + beta * C with op(A) = A or A^T.
 *
 * Each thread owns a TM x TN register tile (8 x 4), so a step over the
 * reduction index reads TM + TN operands from LDS for TM * TN fused
 * multiply-adds.  Tiles of A and B are staged through double-buffered LDS in
 * slices of BK = 16, the next slice's global loads in flight while the
 * current one is consumed.  Two tile shapes: 128 x 64 (256 threads) and, for
 * groups whose products have few rows, 64 x 64 (128 threads).  FP64 on
 * gfx950 has the same peak on the vector and matrix units, so the register
 * tiled vector FMA is the FP64 path.
 *
 * Problems whose reduction dimension is long and whose output is small (the
 * V^T X products of the upper tree levels) are split along K; the slices land
 * in a workspace and a second kernel sums them in slice order, so results do
 * not depend on scheduling.
 */
#include "gemm.h"

#include <algorithm>

namespace hodlr {

namespace {

constexpr int BK = 16;

template <int BM, int BN>
__device__ __forceinline__ int find_problem(const GemmDesc* __restrict__ descs, int count) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (descs[mid].tile0 <= (int)blockIdx.x)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

template <int BM, int BN>
__global__ void __launch_bounds__((BM / 8) * (BN / 4))
    gemm_kernel(const GemmDesc* __restrict__ descs, int count, double* __restrict__ ws) {
  constexpr int TM = 8, TN = 4;
  constexpr int NT = (BM / TM) * (BN / TN);
  constexpr int RX = BM / TM;  // threads along M
  constexpr int A_PER = BM * BK / NT;
  constexpr int B_PER = BN * BK / NT;
  __shared__ double As[2][BK][BM + 1];
  __shared__ double Bs[2][BK][BN + 1];
  const GemmDesc g = descs[find_problem<BM, BN>(descs, count)];
  const int tm = (g.M + BM - 1) / BM, tn = (g.N + BN - 1) / BN;
  int tile = blockIdx.x - g.tile0;
  const int split = tile / (tm * tn);
  tile -= split * tm * tn;
  const int m0 = (tile % tm) * BM, n0 = (tile / tm) * BN;
  const int kchunk = ((g.K + g.ksplit - 1) / g.ksplit + BK - 1) / BK * BK;
  const int kbeg = split * kchunk, kend = min(g.K, kbeg + kchunk);
  const int t = threadIdx.x, tx = t % RX, ty = t / RX;

  double acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; a++)
#pragma unroll
    for (int b = 0; b < TN; b++) acc[a][b] = 0.0;

  double ra[A_PER], rb[B_PER];
  auto load = [&](int k0) {
#pragma unroll
    for (int r = 0; r < A_PER; r++) {
      const int e = t + r * NT;
      int mm, kk;
      if (g.trans_a) {
        kk = e % BK;
        mm = e / BK;
      } else {
        mm = e % BM;
        kk = e / BM;
      }
      const int gm = m0 + mm, gk = k0 + kk;
      ra[r] = (gm < g.M && gk < kend)
                  ? (g.trans_a ? g.A[(size_t)gk + (size_t)gm * g.lda] : g.A[(size_t)gm + (size_t)gk * g.lda])
                  : 0.0;
    }
#pragma unroll
    for (int r = 0; r < B_PER; r++) {
      const int e = t + r * NT;
      const int kk = e % BK, nn = e / BK;
      const int gn = n0 + nn, gk = k0 + kk;
      rb[r] = (gn < g.N && gk < kend) ? g.B[(size_t)gk + (size_t)gn * g.ldb] : 0.0;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int r = 0; r < A_PER; r++) {
      const int e = t + r * NT;
      int mm, kk;
      if (g.trans_a) {
        kk = e % BK;
        mm = e / BK;
      } else {
        mm = e % BM;
        kk = e / BM;
      }
      As[buf][kk][mm] = ra[r];
    }
#pragma unroll
    for (int r = 0; r < B_PER; r++) {
      const int e = t + r * NT;
      Bs[buf][e % BK][e / BK] = rb[r];
    }
  };

  int buf = 0;
  if (kbeg < kend) {
    load(kbeg);
    store(0);
  }
  __syncthreads();
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    const bool more = k0 + BK < kend;
    if (more) load(k0 + BK);
#pragma unroll 2
    for (int kk = 0; kk < BK; kk++) {
      double av[TM], bv[TN];
#pragma unroll
      for (int a = 0; a < TM; a++) av[a] = As[buf][kk][tx + RX * a];
#pragma unroll
      for (int b = 0; b < TN; b++) bv[b] = Bs[buf][kk][ty + (BN / TN) * b];
#pragma unroll
      for (int a = 0; a < TM; a++)
#pragma unroll
        for (int b = 0; b < TN; b++) acc[a][b] = fma(av[a], bv[b], acc[a][b]);
    }
    if (more) store(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  if (g.ksplit == 1) {
#pragma unroll
    for (int a = 0; a < TM; a++)
#pragma unroll
      for (int b = 0; b < TN; b++) {
        const int gm = m0 + tx + RX * a, gn = n0 + ty + (BN / TN) * b;
        if (gm < g.M && gn < g.N) {
          double* c = g.C + (size_t)gm + (size_t)gn * g.ldc;
          *c = (g.beta == 0.0) ? g.alpha * acc[a][b] : g.alpha * acc[a][b] + g.beta * *c;
        }
      }
  } else {
    double* w = ws + g.ws_off + (size_t)split * g.M * g.N;
#pragma unroll
    for (int a = 0; a < TM; a++)
#pragma unroll
      for (int b = 0; b < TN; b++) {
        const int gm = m0 + tx + RX * a, gn = n0 + ty + (BN / TN) * b;
        if (gm < g.M && gn < g.N) w[(size_t)gm + (size_t)gn * g.M] = acc[a][b];
      }
  }
}

__global__ void __launch_bounds__(256)
    splitk_reduce_kernel(const GemmDesc* __restrict__ descs, const int* __restrict__ ids,
                         const double* __restrict__ ws) {
  const GemmDesc g = descs[ids[blockIdx.y]];
  const size_t mn = (size_t)g.M * g.N;
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < mn; e += (size_t)gridDim.x * 256) {
    double s = 0.0;
    for (int sp = 0; sp < g.ksplit; sp++) s += ws[g.ws_off + (size_t)sp * mn + e];
    const int gm = (int)(e % g.M), gn = (int)(e / g.M);
    double* c = g.C + (size_t)gm + (size_t)gn * g.ldc;
    *c = (g.beta == 0.0) ? g.alpha * s : g.alpha * s + g.beta * *c;
  }
}

}  // namespace

int gemm_tile_m(const std::vector<GemmDesc>& descs) {
  // the wide tile when most of the work has more than 64 rows
  double big = 0, small = 0;
  for (auto& g : descs) {
    const double w = (double)g.M * g.N * g.K;
    if (g.M > 64)
      big += w;
    else
      small += w;
  }
  return big >= small ? 128 : 64;
}

size_t gemm_plan(std::vector<GemmDesc>& descs, std::vector<int>& split_ids, int& ntiles,
                 int target_tiles, int bm) {
  const int BM = bm, BN = 64;
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
      const int want = (int)std::min<size_t>(
          64, (target_tiles + tiles_out - 1) / std::max<size_t>(tiles_out, 1));
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
                 const int* d_split_ids, int nsplit, int bm, hipStream_t s) {
  if (count == 0 || ntiles == 0) return;
  if (bm == 128)
    gemm_kernel<128, 64><<<ntiles, 256, 0, s>>>(d_descs, count, d_ws);
  else
    gemm_kernel<64, 64><<<ntiles, 128, 0, s>>>(d_descs, count, d_ws);
  HODLR_HIP_CHECK(hipGetLastError());
  if (nsplit > 0) {
    dim3 g(64, nsplit);
    splitk_reduce_kernel<<<g, 256, 0, s>>>(d_descs, d_split_ids, d_ws);
    HODLR_HIP_CHECK(hipGetLastError());
  }
}

}  // namespace hodlr
