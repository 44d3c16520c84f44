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
#include <type_traits>

namespace hodlr {

namespace {

constexpr int BK = 16;

/* Workgroups go to the 8 XCDs of the chip round-robin (workgroup i to XCD
 * i % 8), each XCD with its own L2.  Tiles are renumbered so that every XCD
 * runs one contiguous range of them: neighbouring tiles, which share operand
 * columns and rows, then meet in the same L2.  A bijection of [0, n). */
__device__ __forceinline__ int xcd_tile() {
  const int n = (int)gridDim.x, i = (int)blockIdx.x;
  const int q = n / 8, r = n % 8, x = i % 8, p = i / 8;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + p;
}

template <class D>
__device__ __forceinline__ int find_problem(const D* __restrict__ descs, int count, int tile) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (descs[mid].tile0 <= tile)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

template <class T, int BM, int BN>
__global__ void __launch_bounds__((BM / 8) * (BN / 4))
    gemm_kernel(const GemmDescT<T>* __restrict__ descs, int count, T* __restrict__ ws) {
  constexpr int TM = 8, TN = 4;
  constexpr int NT = (BM / TM) * (BN / TN);
  constexpr int RX = BM / TM;  // threads along M
  constexpr int A_PER = BM * BK / NT;
  constexpr int B_PER = BN * BK / NT;
  __shared__ T As[2][BK][BM + 1];
  __shared__ T Bs[2][BK][BN + 1];
  const int gtile = xcd_tile();
  const GemmDescT<T> g = descs[find_problem(descs, count, gtile)];
  const gptr<const T> gA = to_global(g.A), gB = to_global(g.B);
  const gptr<T> gC = to_global(g.C);
  const int tm = (g.M + BM - 1) / BM, tn = (g.N + BN - 1) / BN;
  int tile = gtile - g.tile0;
  const int split = tile / (tm * tn);
  tile -= split * tm * tn;
  const int m0 = (tile % tm) * BM, n0 = (tile / tm) * BN;
  const int kchunk = ((g.K + g.ksplit - 1) / g.ksplit + BK - 1) / BK * BK;
  const int kbeg = split * kchunk, kend = min(g.K, kbeg + kchunk);
  const int t = threadIdx.x, tx = t % RX, ty = t / RX;

  T acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; a++)
#pragma unroll
    for (int b = 0; b < TN; b++) acc[a][b] = T(0);

  T ra[A_PER], rb[B_PER];
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
                  ? (g.trans_a ? gA[(size_t)gk + (size_t)gm * g.lda] : gA[(size_t)gm + (size_t)gk * g.lda])
                  : T(0);
    }
#pragma unroll
    for (int r = 0; r < B_PER; r++) {
      const int e = t + r * NT;
      const int kk = e % BK, nn = e / BK;
      const int gn = n0 + nn, gk = k0 + kk;
      rb[r] = (gn < g.N && gk < kend) ? gB[(size_t)gk + (size_t)gn * g.ldb] : T(0);
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
      T av[TM], bv[TN];
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
          gptr<T> c = gC + (size_t)gm + (size_t)gn * g.ldc;
          *c = (g.beta == T(0)) ? g.alpha * acc[a][b] : g.alpha * acc[a][b] + g.beta * *c;
        }
      }
  } else {
    T* w = ws + g.ws_off + (size_t)split * g.M * g.N;
#pragma unroll
    for (int a = 0; a < TM; a++)
#pragma unroll
      for (int b = 0; b < TN; b++) {
        const int gm = m0 + tx + RX * a, gn = n0 + ty + (BN / TN) * b;
        if (gm < g.M && gn < g.N) w[(size_t)gm + (size_t)gn * g.M] = acc[a][b];
      }
  }
}

/* The matrix cores: the descriptors and split-K contract of gemm_kernel,
 * tiles BM x BN of 128 x 64 or 64 x 128 (the latter for products with few
 * rows: the V^T W products of the factorization, whose row count is a rank),
 * each wave computing a 64 x 32 sub-tile as 4 x 2 blocks of
 * v_mfma_f64_16x16x4_f64 (FP64) or v_mfma_f32_16x16x4_f32 (FP32, the mixed
 * path's products, accumulated in FP32).  Operand layout of both
 * instructions (wave of 64): lane l holds A[l % 16][l / 16] of the 16 x 4
 * A block and B[l / 16][l % 16] of the 4 x 16 B block; the 16 x 16 result
 * is held in register i = 0..3 as D[l / 16 + 4 i][l % 16] (FP64) or
 * D[4 (l / 16) + i][l % 16] (FP32).  The blocks are multiplied transposed,
 * D = B^T A^T = (A B)^T, so that lane l holds (A B)[l % 16][Mfma<T>::col]:
 * the 16 lanes of a quarter wave then read and write 16 consecutive rows of
 * one column of the column-major C (a 64- or 128-byte run) instead of one
 * element in each of 16 columns. */
typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <class T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef v4d acc;
  static __device__ __forceinline__ acc mul(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // result column (of the 16) that register i of lane group g = l / 16 holds
  static __device__ __forceinline__ int col(int g, int i) { return g + 4 * i; }
};
template <>
struct Mfma<float> {
  typedef v4f acc;
  static __device__ __forceinline__ acc mul(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int col(int g, int i) { return 4 * g + i; }
};

template <class T, int BM, int BN>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3)))
    gemm_mfma_kernel(const GemmDescT<T>* __restrict__ descs, int count, T* __restrict__ ws) {
  constexpr int NT = 256, WM = BM / 64;  // waves along M
  static_assert(WM * (BN / 32) == 4, "four waves of 64 x 32");
  constexpr int A_PER = BM * BK / NT;
  constexpr int B_PER = BN * BK / NT;
  __shared__ T As[BK][BM + 2];
  __shared__ T Bs[BK][BN + 2];
  const int gtile = xcd_tile();
  const GemmDescT<T> g = descs[find_problem(descs, count, gtile)];
  const gptr<const T> gA = to_global(g.A), gB = to_global(g.B);
  const gptr<T> gC = to_global(g.C);
  const int tm = (g.M + BM - 1) / BM, tn = (g.N + BN - 1) / BN;
  int tile = gtile - g.tile0;
  const int split = tile / (tm * tn);
  tile -= split * tm * tn;
  const int m0 = (tile % tm) * BM, n0 = (tile / tm) * BN;
  const int kchunk = ((g.K + g.ksplit - 1) / g.ksplit + BK - 1) / BK * BK;
  const int kbeg = split * kchunk, kend = min(g.K, kbeg + kchunk);
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = (wave % WM) * 64, wn = (wave / WM) * 32;
  const int lr = lane & 15, lk = lane >> 4;

  // C += A B and C -= A B (the factorization's updates) start the sums from
  // +-C, so the loads of C overlap the first operand slices instead of
  // following the last one; alpha = +-1 keeps every step exact but the sums
  const bool cinit = g.ksplit == 1 && g.beta == T(1) && (g.alpha == T(1) || g.alpha == T(-1));
  typename Mfma<T>::acc acc[4][2];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int gm = m0 + wm + a * 16 + lr, gn = n0 + wn + b * 16 + Mfma<T>::col(lk, i);
        acc[a][b][i] = (cinit && gm < g.M && gn < g.N)
                           ? g.alpha * gC[(size_t)gm + (size_t)gn * g.ldc]
                           : T(0);
      }

  T ra[A_PER], rb[B_PER];
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
                  ? (g.trans_a ? gA[(size_t)gk + (size_t)gm * g.lda] : gA[(size_t)gm + (size_t)gk * g.lda])
                  : T(0);
    }
#pragma unroll
    for (int r = 0; r < B_PER; r++) {
      const int e = t + r * NT;
      const int kk = e % BK, nn = e / BK;
      const int gn = n0 + nn, gk = k0 + kk;
      rb[r] = (gn < g.N && gk < kend) ? gB[(size_t)gk + (size_t)gn * g.ldb] : T(0);
    }
  };
  auto store = [&]() {
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
      As[kk][mm] = ra[r];
    }
#pragma unroll
    for (int r = 0; r < B_PER; r++) {
      const int e = t + r * NT;
      Bs[e % BK][e / BK] = rb[r];
    }
  };

  // one LDS slice; the next slice's loads are in flight in registers while
  // the current one is multiplied (half the LDS of double buffering, so more
  // workgroups fit per CU)
  if (kbeg < kend) load(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    __syncthreads();
    store();
    __syncthreads();
    if (k0 + BK < kend) load(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T av[4], bv[2];
#pragma unroll
      for (int a = 0; a < 4; a++) av[a] = As[kk + lk][wm + a * 16 + lr];
#pragma unroll
      for (int b = 0; b < 2; b++) bv[b] = Bs[kk + lk][wn + b * 16 + lr];
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
          acc[a][b] = Mfma<T>::mul(bv[b], av[a], acc[a][b]);
    }
  }
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int gm = m0 + wm + a * 16 + lr, gn = n0 + wn + b * 16 + Mfma<T>::col(lk, i);
        if (gm < g.M && gn < g.N) {
          const T v = acc[a][b][i];
          if (cinit) {
            gC[(size_t)gm + (size_t)gn * g.ldc] = g.alpha * v;
          } else if (g.ksplit == 1) {
            gptr<T> c = gC + (size_t)gm + (size_t)gn * g.ldc;
            *c = (g.beta == T(0)) ? g.alpha * v : g.alpha * v + g.beta * *c;
          } else {
            ws[g.ws_off + (size_t)split * g.M * g.N + (size_t)gm + (size_t)gn * g.M] = v;
          }
        }
      }
}

/* The factorization's low-rank updates C -= A B (op(A) = A, alpha = -1 or
 * +1 and beta = 1, K <= 4 KS) on the matrix cores, streamed: a workgroup
 * takes 64 RB rows of C and UPD_NCH consecutive 64-column tiles.  Its slice
 * of A (64 RB x K) is loaded once into registers as the instruction's
 * operands; per column tile the K x 64 tile of B is staged through LDS, and
 * the next tile's C and B values are loaded while the current tile is
 * multiplied, so the read-modify-write of C streams instead of paying a
 * tile's latency.  Wave w computes rows 16 RB w .. 16 RB w + 16 RB - 1 of
 * the tile as RB x 4 blocks of 16 x 16, in the layout of gemm_mfma_kernel;
 * the accumulation (C scaled by alpha, then the K steps of four in
 * increasing order, padded to a multiple of 16 as there) is that kernel's,
 * so the results are bit-identical. */
constexpr int UPD_NCH = 8;

template <int KS, int RB>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
    gemm_update_kernel(const GemmDescT<double>* __restrict__ descs, int count) {
  constexpr int WR = 64 * RB;  // rows per workgroup
  __shared__ double Bs[4 * KS][64 + 2];
  const int wgl = xcd_tile();
  const GemmDescT<double> g = descs[find_problem(descs, count, wgl)];
  const gptr<const double> gA = to_global(g.A), gB = to_global(g.B);
  const gptr<double> gC = to_global(g.C);
  const int tm = (g.M + WR - 1) / WR, ntn = (g.N + 63) / 64;
  const int w = wgl - g.tile0, mb = w % tm, nchunk = w / tm;
  const int m0 = mb * WR, t0 = nchunk * UPD_NCH, t1 = min(ntn, t0 + UPD_NCH);
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int r0 = m0 + 16 * RB * wave;
  const int nks = (g.K + 15) / 16 * 4;  // K steps of four, as in gemm_mfma_kernel's slices
  const double alpha = g.alpha;
  // A as operands: av[a][s] = A(r0 + 16 a + lr, 4 s + lk)
  double av[RB][KS];
#pragma unroll
  for (int a = 0; a < RB; a++)
#pragma unroll
    for (int s = 0; s < KS; s++) {
      const int gm = r0 + 16 * a + lr, gk = 4 * s + lk;
      av[a][s] = (gm < g.M && gk < g.K) ? gA[(size_t)gm + (size_t)gk * g.lda] : 0.0;
    }
  // C of column tile tt: cn[a][b][i] = C(r0 + 16 a + lr, 64 tt + 16 b + lk + 4 i)
  double cn[RB][4][4];
  auto load_c = [&](int tt) {
#pragma unroll
    for (int a = 0; a < RB; a++)
#pragma unroll
      for (int b = 0; b < 4; b++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int gm = r0 + 16 * a + lr, gn = 64 * tt + 16 * b + lk + 4 * i;
          cn[a][b][i] = (gm < g.M && gn < g.N) ? gC[(size_t)gm + (size_t)gn * g.ldc] : 0.0;
        }
  };
  // B of column tile tt, this thread's share: rows e % (4 KS), column e / (4 KS)
  double rb[KS];
  auto load_b = [&](int tt) {
#pragma unroll
    for (int q = 0; q < KS; q++) {
      const int e = t + 256 * q, kk = e % (4 * KS), nn = e / (4 * KS);
      const int gn = 64 * tt + nn;
      rb[q] = (kk < g.K && gn < g.N) ? gB[(size_t)kk + (size_t)gn * g.ldb] : 0.0;
    }
  };
  if (t0 < t1) {
    load_c(t0);
    load_b(t0);
  }
  for (int tt = t0; tt < t1; tt++) {
    __syncthreads();  // every wave is done with the previous tile of B
#pragma unroll
    for (int q = 0; q < KS; q++) {
      const int e = t + 256 * q;
      Bs[e % (4 * KS)][e / (4 * KS)] = rb[q];
    }
    v4d acc[RB][4];
#pragma unroll
    for (int a = 0; a < RB; a++)
#pragma unroll
      for (int b = 0; b < 4; b++)
#pragma unroll
        for (int i = 0; i < 4; i++) acc[a][b][i] = alpha * cn[a][b][i];
    __syncthreads();
    if (tt + 1 < t1) {
      load_c(tt + 1);
      load_b(tt + 1);
    }
#pragma unroll
    for (int s = 0; s < KS; s++) {
      if (s < nks) {
        double bv[4];
#pragma unroll
        for (int b = 0; b < 4; b++) bv[b] = Bs[4 * s + lk][16 * b + lr];
#pragma unroll
        for (int a = 0; a < RB; a++)
#pragma unroll
          for (int b = 0; b < 4; b++)
            acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[b], av[a][s], acc[a][b], 0, 0, 0);
      }
    }
#pragma unroll
    for (int a = 0; a < RB; a++)
#pragma unroll
      for (int b = 0; b < 4; b++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int gm = r0 + 16 * a + lr, gn = 64 * tt + 16 * b + lk + 4 * i;
          if (gm < g.M && gn < g.N) gC[(size_t)gm + (size_t)gn * g.ldc] = alpha * acc[a][b][i];
        }
  }
}

template <class T>
__global__ void __launch_bounds__(256)
    splitk_reduce_kernel(const GemmDescT<T>* __restrict__ descs, const int* __restrict__ ids,
                         const T* __restrict__ ws) {
  const GemmDescT<T> g = descs[ids[blockIdx.y]];
  const gptr<T> gC = to_global(g.C);
  const size_t mn = (size_t)g.M * g.N;
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < mn; e += (size_t)gridDim.x * 256) {
    T s = T(0);
    for (int sp = 0; sp < g.ksplit; sp++) s += ws[g.ws_off + (size_t)sp * mn + e];
    const int gm = (int)(e % g.M), gn = (int)(e / g.M);
    gptr<T> c = gC + (size_t)gm + (size_t)gn * g.ldc;
    *c = (g.beta == T(0)) ? g.alpha * s : g.alpha * s + g.beta * *c;
  }
}

}  // namespace

template <class T>
int gemm_tile_m(const std::vector<GemmDescT<T>>& descs, bool mfma) {
  if (mfma) {
    // the same tile area either way: the shape that pads less
    double a128 = 0, a64 = 0;
    for (auto& g : descs) {
      a128 += (double)((g.M + 127) / 128) * ((g.N + 63) / 64) * g.K;
      a64 += (double)((g.M + 63) / 64) * ((g.N + 127) / 128) * g.K;
    }
    return a64 < a128 ? 64 : 128;
  }
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

template <class T>
int gemm_tile_n(int bm, bool mfma) {
  return (mfma && bm == 64) ? 128 : 64;
}

template <class T>
size_t gemm_plan(std::vector<GemmDescT<T>>& descs, std::vector<int>& split_ids, int& ntiles,
                 int target_tiles, int bm, bool mfma) {
  const int BM = bm, BN = gemm_tile_n<T>(bm, mfma);
  size_t tiles_out = 0;
  for (auto& g : descs) {
    const size_t tm = (g.M + BM - 1) / BM, tn = (g.N + BN - 1) / BN;
    tiles_out += tm * tn;
  }
  size_t ws = 0;
  int tile = 0;
  split_ids.clear();
  for (size_t p = 0; p < descs.size(); p++) {
    GemmDescT<T>& g = descs[p];
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

template <class T>
void launch_gemm(const GemmDescT<T>* d_descs, int count, int ntiles, T* d_ws,
                 const int* d_split_ids, int nsplit, int bm, hipStream_t s, bool mfma) {
  if (count == 0 || ntiles == 0) return;
  if (mfma) {
    if (bm == 128)
      gemm_mfma_kernel<T, 128, 64><<<ntiles, 256, 0, s>>>(d_descs, count, d_ws);
    else
      gemm_mfma_kernel<T, 64, 128><<<ntiles, 256, 0, s>>>(d_descs, count, d_ws);
    HODLR_HIP_CHECK(hipGetLastError());
    if (nsplit > 0) {
      dim3 g(64, nsplit);
      splitk_reduce_kernel<T><<<g, 256, 0, s>>>(d_descs, d_split_ids, d_ws);
      HODLR_HIP_CHECK(hipGetLastError());
    }
    return;
  }
  if (bm == 128)
    gemm_kernel<T, 128, 64><<<ntiles, 256, 0, s>>>(d_descs, count, d_ws);
  else
    gemm_kernel<T, 64, 64><<<ntiles, 128, 0, s>>>(d_descs, count, d_ws);
  HODLR_HIP_CHECK(hipGetLastError());
  if (nsplit > 0) {
    dim3 g(64, nsplit);
    splitk_reduce_kernel<T><<<g, 256, 0, s>>>(d_descs, d_split_ids, d_ws);
    HODLR_HIP_CHECK(hipGetLastError());
  }
}

int gemm_update_variant(int maxk) { return maxk <= 32 ? 0 : maxk <= 48 ? 1 : 2; }

int gemm_update_rows(int maxk) { return gemm_update_variant(maxk) == 0 ? 128 : 64; }

bool gemm_update_eligible(const GemmDesc& g) {
  return g.trans_a == 0 && g.beta == 1.0 && (g.alpha == 1.0 || g.alpha == -1.0) &&
         g.K <= GEMM_UPDATE_KMAX && g.N >= 128;
}

int gemm_update_plan(std::vector<GemmDesc>& descs, int rows) {
  int wg = 0;
  for (auto& g : descs) {
    const int tm = (g.M + rows - 1) / rows, ntn = (g.N + 63) / 64;
    g.ksplit = 1;
    g.ws_off = 0;
    g.tile0 = wg;
    wg += tm * ((ntn + UPD_NCH - 1) / UPD_NCH);
  }
  return wg;
}

void launch_gemm_update(const GemmDesc* d_descs, int count, int nwg, int maxk, hipStream_t s) {
  if (count == 0 || nwg == 0) return;
  switch (gemm_update_variant(maxk)) {
    case 0: gemm_update_kernel<8, 2><<<nwg, 256, 0, s>>>(d_descs, count); break;
    case 1: gemm_update_kernel<12, 1><<<nwg, 256, 0, s>>>(d_descs, count); break;
    default: gemm_update_kernel<16, 1><<<nwg, 256, 0, s>>>(d_descs, count);
  }
  HODLR_HIP_CHECK(hipGetLastError());
}

template int gemm_tile_m<double>(const std::vector<GemmDesc>&, bool);
template int gemm_tile_m<float>(const std::vector<GemmDescF>&, bool);
template int gemm_tile_n<double>(int, bool);
template int gemm_tile_n<float>(int, bool);
template size_t gemm_plan<double>(std::vector<GemmDesc>&, std::vector<int>&, int&, int, int, bool);
template size_t gemm_plan<float>(std::vector<GemmDescF>&, std::vector<int>&, int&, int, int, bool);
template void launch_gemm<double>(const GemmDesc*, int, int, double*, const int*, int, int,
                                  hipStream_t, bool);
template void launch_gemm<float>(const GemmDescF*, int, int, float*, const int*, int, int,
                                 hipStream_t, bool);

}  // namespace hodlr
