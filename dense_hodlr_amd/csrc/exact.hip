/*
 * The exact kernel operator Y = K X on the device, K(i, j) evaluated on the
 * fly (no matrix is stored): the operator the iterative solver iterates on
 * for kernel instances.  One workgroup per (instance, 64-row tile); the
 * kernel's columns are swept in blocks of 64:
 *
 *  NR <= 4 right-hand sides: each thread one row (four row tiles of 64 per
 *          workgroup) accumulates K(i, j) x_j with the entries evaluated in
 *          registers (the cost is the N^2 entry evaluations);
 *  NR 16/64: the block of K (64 x 64) is evaluated once into LDS, X's block
 *          (64 x NR) staged beside it, and each thread accumulates NR / 4
 *          columns of one row from there (entry cost shared by NR columns).
 *
 * Products are summed in column order j = 0, 1, ... per row and right-hand
 * side.
 */
#include "exact.h"

#include <algorithm>

namespace hodlr {

namespace {

constexpr int TPB = 256;
constexpr int KB = 64;  // kernel columns per block

__device__ int find_tile(const KopDesc* __restrict__ d, int count, int wg) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].tile0 <= wg)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

/* one row per thread, 256 rows per workgroup, NR <= 4 right-hand sides */
template <int NR>
__global__ void __launch_bounds__(TPB) kop_rows_kernel(const KopDesc* __restrict__ d, int count,
                                                       const EntrySource* __restrict__ srcs,
                                                       int nrhs) {
  __shared__ double Xs[KB][NR];
  const KopDesc p = d[find_tile(d, count, blockIdx.x)];
  const EntrySource src = srcs[p.inst];
  const int n = src.n;
  const int i = ((int)blockIdx.x - p.tile0) * TPB + (int)threadIdx.x;
  double acc[NR];
#pragma unroll
  for (int c = 0; c < NR; c++) acc[c] = 0.0;
  for (int j0 = 0; j0 < n; j0 += KB) {
    __syncthreads();
    for (int e = threadIdx.x; e < KB * NR; e += TPB) {
      const int jj = e / NR, c = e - jj * NR;
      Xs[jj][c] = (j0 + jj < n && c < nrhs) ? p.X[(size_t)(j0 + jj) + (size_t)c * p.ldx] : 0.0;
    }
    __syncthreads();
    if (i < n) {
      const int jn = min(KB, n - j0);
      for (int jj = 0; jj < jn; jj++) {
        const double k = entry_at(src, i, j0 + jj);
#pragma unroll
        for (int c = 0; c < NR; c++) acc[c] += k * Xs[jj][c];
      }
    }
  }
  if (i < n)
#pragma unroll
    for (int c = 0; c < NR; c++)
      if (c < nrhs) p.Y[(size_t)i + (size_t)c * p.ldy] = acc[c];
}

/* 64 rows per workgroup; the K block in LDS shared by NR right-hand sides */
template <int NR>
__global__ void __launch_bounds__(TPB) kop_tile_kernel(const KopDesc* __restrict__ d, int count,
                                                       const EntrySource* __restrict__ srcs,
                                                       int nrhs, int c0) {
  constexpr int CPT = NR / 4;  // right-hand sides per thread
  __shared__ double Ks[64][KB + 1];
  __shared__ double Xs[KB][NR];
  const KopDesc p = d[find_tile(d, count, blockIdx.x)];
  const EntrySource src = srcs[p.inst];
  const int n = src.n;
  const int t = threadIdx.x, r = t & 63, g = t >> 6;
  const int i0 = ((int)blockIdx.x - p.tile0) * 64;
  const int i = i0 + r;
  double acc[CPT];
#pragma unroll
  for (int c = 0; c < CPT; c++) acc[c] = 0.0;
  for (int j0 = 0; j0 < n; j0 += KB) {
    __syncthreads();
    // the K block: thread (r, g) evaluates row i, columns g*16 .. g*16+15
#pragma unroll 4
    for (int q = 0; q < KB / 4; q++) {
      const int jj = g * (KB / 4) + q, j = j0 + jj;
      Ks[r][jj] = (i < n && j < n) ? entry_at(src, i, j) : 0.0;
    }
    for (int e = t; e < KB * NR; e += TPB) {
      const int jj = e / NR, c = e - jj * NR;
      Xs[jj][c] = (j0 + jj < n && c0 + c < nrhs)
                      ? p.X[(size_t)(j0 + jj) + (size_t)(c0 + c) * p.ldx]
                      : 0.0;
    }
    __syncthreads();
    const int jn = min(KB, n - j0);
    for (int jj = 0; jj < jn; jj++) {
      const double k = Ks[r][jj];
#pragma unroll
      for (int c = 0; c < CPT; c++) acc[c] += k * Xs[jj][g * CPT + c];
    }
  }
  if (i < n)
#pragma unroll
    for (int c = 0; c < CPT; c++) {
      const int col = c0 + g * CPT + c;
      if (col < nrhs) p.Y[(size_t)i + (size_t)col * p.ldy] = acc[c];
    }
}

}  // namespace

int kop_tiles(int n, int nrhs) {
  const int rows = nrhs <= 4 ? TPB : 64;
  return (n + rows - 1) / rows;
}

void launch_kernel_operator(const KopDesc* d, int count, int ntiles, const EntrySource* srcs,
                            int nrhs, hipStream_t s) {
  if (count == 0 || ntiles == 0 || nrhs <= 0) return;
  if (nrhs == 1) {
    kop_rows_kernel<1><<<ntiles, TPB, 0, s>>>(d, count, srcs, nrhs);
  } else if (nrhs <= 4) {
    kop_rows_kernel<4><<<ntiles, TPB, 0, s>>>(d, count, srcs, nrhs);
  } else if (nrhs <= 16) {
    kop_tile_kernel<16><<<ntiles, TPB, 0, s>>>(d, count, srcs, nrhs, 0);
  } else {
    for (int c0 = 0; c0 < nrhs; c0 += 64) kop_tile_kernel<64><<<ntiles, TPB, 0, s>>>(d, count, srcs, nrhs, c0);
  }
  HODLR_HIP_CHECK(hipGetLastError());
}

}  // namespace hodlr
