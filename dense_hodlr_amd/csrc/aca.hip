/*
 * Level-batched partial-pivot ACA fused with kernel-entry assembly.
 *
 * Restates partialPivACA_LowRankApprox (src/lowRank.cpp:117-279) and
 * chooseNextRowCol (:291-298) for every off-diagonal block of a set of tree
 * levels at once.  One ACA iteration of every active block is three launches:
 *
 *   row phase   residual row r = A(i,:) - W(i,:) V^T, written to V(:,k);
 *               per-workgroup argmax |r|, first unchosen column with
 *               |r| > minPivot, sum r^2 and t = V^T A(i,:) partials
 *   col phase   pick the pivot column j (first max, else the first unchosen
 *               one), residual column c = A(:,j) - V(j,:) W^T, write
 *               w = c / r_j to W(:,k); argmax |c|, first unchosen row,
 *               sum w^2 and u = W^T A(:,j) partials
 *   finalize    Gram columns d = V^T r = t - G_V W(i,:)^T and
 *               e = W^T w = (u - G_W V(j,:)^T) / r_j, the incremental
 *               Frobenius norm (:234-239), epsilon (:241), the next row and
 *               the stopping rule (:149, :246); the last workgroup of a block
 *               to arrive reduces the partials in a fixed order.
 *
 * The residual sums run sequentially over l with separately rounded multiply
 * and add (__dmul_rn/__dadd_rn), the order and rounding of the reference's
 * Eigen loop `sum += W(i,l) * V.col(l)`, so a residual entry is bit-identical
 * to the CPU restatement's.  The Gram identity lets the dot products that the
 * Frobenius update needs come out of the same single pass over the V and W
 * panels that the residuals take, so each iteration streams each panel once.
 */
#include "aca.h"

#include <algorithm>
#include <vector>

namespace hodlr {

namespace {

constexpr int TPB = 256;   // threads per workgroup
constexpr int LCH = 16;    // l-chunk of the transposed partial reduction
constexpr int FIN_L = 256; // Gram entries per finalize workgroup

__device__ inline void argmax_merge(double& bv, int& bi, double v, int i) {
  // larger value wins; on equal values the lower index wins (first max)
  if (v > bv || (v == bv && i >= 0 && (bi < 0 || i < bi))) {
    bv = v;
    bi = i;
  }
}

__device__ inline bool bit_test(const uint32_t* bits, int i) {
  return (bits[i >> 5] >> (i & 31)) & 1u;
}

__device__ inline void bit_set(uint32_t* bits, int i) { atomicOr(&bits[i >> 5], 1u << (i & 31)); }

/* Workgroup reduction of (max |.|, argmax, first-unchosen index, sum of squares). */
struct PivRed {
  double v;
  int i;
  int first;
  double sq;
};

__device__ PivRed wg_reduce_piv(PivRed p, PivRed* sh) {
  const int t = threadIdx.x;
  // wave-level
  for (int off = 32; off > 0; off >>= 1) {
    double ov = __shfl_xor(p.v, off);
    int oi = __shfl_xor(p.i, off);
    int of = __shfl_xor(p.first, off);
    double os = __shfl_xor(p.sq, off);
    argmax_merge(p.v, p.i, ov, oi);
    if (of >= 0 && (p.first < 0 || of < p.first)) p.first = of;
    p.sq += os;
  }
  const int lane = t & 63, wave = t >> 6;
  if (lane == 0) sh[wave] = p;
  __syncthreads();
  if (t == 0) {
    PivRed r = sh[0];
    for (int w = 1; w < TPB / 64; w++) {
      argmax_merge(r.v, r.i, sh[w].v, sh[w].i);
      if (sh[w].first >= 0 && (r.first < 0 || sh[w].first < r.first)) r.first = sh[w].first;
      r.sq += sh[w].sq;
    }
    sh[0] = r;
  }
  __syncthreads();
  PivRed r = sh[0];
  __syncthreads();
  return r;
}

/* Partial slot layout: (cap + 4) doubles per workgroup slot:
 * [0, cap) dot partials, cap: max |.|, cap+1: argmax, cap+2: first, cap+3: sum sq. */
__device__ inline double* slot_ptr(const AcaLevels& L, const AcaBlock& b, int slot) {
  const int cap = L.cap[b.level];
  return L.P[b.level] + (size_t)(b.part_off + slot) * (size_t)(cap + 4);
}

__device__ inline void read_piv(const double* s, int cap, double& v, int& i, int& f) {
  v = s[cap];
  i = (int)s[cap + 1];
  f = (int)s[cap + 2];
}

/* Reduce the row-phase pivot partials of block b; returns the chosen column
 * (>=0), -1 for a row jump (max |r| <= minPivot), -2 for no valid column. */
__device__ int decide_column(const AcaLevels& L, const AcaBlock& b, const uint32_t* colbits,
                             int min_pivot, double* rmax_out) {
  const int cap = L.cap[b.level];
  double bv = -1.0;
  int bi = -1, bf = -1;
  for (int w = 0; w < b.row_nwg; w++) {
    double v;
    int i, f;
    read_piv(slot_ptr(L, b, w), cap, v, i, f);
    argmax_merge(bv, bi, v, i);
    if (f >= 0 && (bf < 0 || f < bf)) bf = f;
  }
  if (rmax_out) *rmax_out = bv;
  if (bv <= (double)min_pivot) return -1;
  if (!bit_test(colbits, bi)) return bi;
  return bf >= 0 ? bf : -2;
}

__global__ void __launch_bounds__(TPB) aca_init_kernel(const AcaBlock* blocks, AcaState* st,
                                                       uint32_t* bits, int nblocks) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblocks) return;
  AcaState s;
  s.k = 0;
  s.cur_row = 0;
  s.status = (blocks[b].m > 0 && blocks[b].n > 0) ? 0 : 1;
  s.cur_col = -1;
  s.frob2 = 0.0;
  s.eps = 1.0;
  s.iters = 0;
  s.pad_[0] = s.pad_[1] = s.pad_[2] = 0;
  st[b] = s;
  bits[blocks[b].bits_row0] |= 1u;  // row 0 is chosen first (lowRank.cpp:159)
}

template <class SRC>
__global__ void __launch_bounds__(TPB)
    aca_row_kernel(AcaLevels L, const AcaBlock* __restrict__ blocks, const int* __restrict__ wg2blk,
                   const AcaState* __restrict__ st, const uint32_t* __restrict__ bits,
                   const SRC* __restrict__ srcs, int jpt, int min_pivot) {
  __shared__ double red[LCH][TPB];
  __shared__ PivRed pr_sh[TPB / 64];
  const int bidx = wg2blk[blockIdx.x];
  const AcaBlock b = blocks[bidx];
  const AcaState s = st[bidx];
  if (s.status != 0) return;
  const int local = blockIdx.x - b.row_wg0;
  const int t = threadIdx.x;
  const int ld = L.ld;
  const int k = s.k;
  const int i = s.cur_row;
  const double* __restrict__ Wp = L.W[b.level] + b.w_row0;
  double* __restrict__ Vp = L.V[b.level] + b.v_row0;
  const SRC src = srcs[b.inst];
  const uint32_t* colbits = bits + b.bits_col0;
  const int cpw = TPB * jpt;
  const int jbase = local * cpw + t;

  constexpr int JMAX = 4;
  double a[JMAX], sacc[JMAX];
  int jj[JMAX];
#pragma unroll
  for (int q = 0; q < JMAX; q++) {
    jj[q] = jbase + q * TPB;
    bool ok = q < jpt && jj[q] < b.n;
    a[q] = ok ? entry_at(src, b.i0 + i, b.j0 + jj[q]) : 0.0;
    if (!ok) jj[q] = -1;
    sacc[q] = 0.0;
  }
  double* slot = slot_ptr(L, b, local);
  for (int l0 = 0; l0 < k; l0 += LCH) {
    const int lc = min(LCH, k - l0);
    for (int ll = 0; ll < lc; ll++) {
      const int l = l0 + ll;
      const double w = Wp[(size_t)i + (size_t)l * ld];
      double p = 0.0;
#pragma unroll
      for (int q = 0; q < JMAX; q++) {
        if (jj[q] >= 0) {
          const double v = Vp[(size_t)jj[q] + (size_t)l * ld];
          sacc[q] = __dadd_rn(sacc[q], __dmul_rn(w, v));
          p += v * a[q];
        }
      }
      red[ll][t] = p;
    }
    __syncthreads();
    {
      const int ll = t >> 4, part = t & 15;
      double sum = 0.0;
      if (ll < lc) {
#pragma unroll
        for (int u = 0; u < 16; u++) sum += red[ll][part * 16 + u];
      }
      for (int off = 8; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
      if (part == 0 && ll < lc) slot[l0 + ll] = sum;
    }
    __syncthreads();
  }
  PivRed p;
  p.v = -1.0;
  p.i = -1;
  p.first = -1;
  p.sq = 0.0;
#pragma unroll
  for (int q = 0; q < JMAX; q++) {
    if (jj[q] >= 0) {
      const double r = __dadd_rn(a[q], -sacc[q]);
      Vp[(size_t)jj[q] + (size_t)k * ld] = r;
      const double ar = fabs(r);
      argmax_merge(p.v, p.i, ar, jj[q]);
      if (!bit_test(colbits, jj[q]) && ar > (double)min_pivot && (p.first < 0 || jj[q] < p.first))
        p.first = jj[q];
      p.sq += r * r;
    }
  }
  PivRed r = wg_reduce_piv(p, pr_sh);
  if (t == 0) {
    const int cap = L.cap[b.level];
    slot[cap] = r.v;
    slot[cap + 1] = (double)r.i;
    slot[cap + 2] = (double)r.first;
    slot[cap + 3] = r.sq;
  }
}

template <class SRC>
__global__ void __launch_bounds__(TPB)
    aca_col_kernel(AcaLevels L, const AcaBlock* __restrict__ blocks, const int* __restrict__ wg2blk,
                   AcaState* __restrict__ st, const uint32_t* __restrict__ bits,
                   const SRC* __restrict__ srcs, int ipt, int min_pivot) {
  __shared__ double red[LCH][TPB];
  __shared__ PivRed pr_sh[TPB / 64];
  const int bidx = wg2blk[blockIdx.x];
  const AcaBlock b = blocks[bidx];
  const AcaState s = st[bidx];
  if (s.status != 0) return;
  const int local = blockIdx.x - b.col_wg0;
  const int t = threadIdx.x;
  const int ld = L.ld;
  const int k = s.k;
  const int j = decide_column(L, b, bits + b.bits_col0, min_pivot, nullptr);
  if (local == 0 && t == 0) st[bidx].cur_col = j;
  if (j < 0) return;
  double* __restrict__ Wp = L.W[b.level] + b.w_row0;
  const double* __restrict__ Vp = L.V[b.level] + b.v_row0;
  const SRC src = srcs[b.inst];
  const uint32_t* rowbits = bits + b.bits_row0;
  const double pivot = 1.0 / Vp[(size_t)j + (size_t)k * ld];  // lowRank.cpp:208
  const int rpw = TPB * ipt;
  const int ibase = local * rpw + t;

  constexpr int IMAX = 4;
  double a[IMAX], sacc[IMAX];
  int ii[IMAX];
#pragma unroll
  for (int q = 0; q < IMAX; q++) {
    ii[q] = ibase + q * TPB;
    bool ok = q < ipt && ii[q] < b.m;
    a[q] = ok ? entry_at(src, b.i0 + ii[q], b.j0 + j) : 0.0;
    if (!ok) ii[q] = -1;
    sacc[q] = 0.0;
  }
  double* slot = slot_ptr(L, b, b.row_nwg + local);
  for (int l0 = 0; l0 < k; l0 += LCH) {
    const int lc = min(LCH, k - l0);
    for (int ll = 0; ll < lc; ll++) {
      const int l = l0 + ll;
      const double v = Vp[(size_t)j + (size_t)l * ld];
      double p = 0.0;
#pragma unroll
      for (int q = 0; q < IMAX; q++) {
        if (ii[q] >= 0) {
          const double w = Wp[(size_t)ii[q] + (size_t)l * ld];
          sacc[q] = __dadd_rn(sacc[q], __dmul_rn(v, w));
          p += w * a[q];
        }
      }
      red[ll][t] = p;
    }
    __syncthreads();
    {
      const int ll = t >> 4, part = t & 15;
      double sum = 0.0;
      if (ll < lc) {
#pragma unroll
        for (int u = 0; u < 16; u++) sum += red[ll][part * 16 + u];
      }
      for (int off = 8; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
      if (part == 0 && ll < lc) slot[l0 + ll] = sum;
    }
    __syncthreads();
  }
  PivRed p;
  p.v = -1.0;
  p.i = -1;
  p.first = -1;
  p.sq = 0.0;
#pragma unroll
  for (int q = 0; q < IMAX; q++) {
    if (ii[q] >= 0) {
      const double c = __dadd_rn(a[q], -sacc[q]);
      const double w = __dmul_rn(pivot, c);  // lowRank.cpp:230
      Wp[(size_t)ii[q] + (size_t)k * ld] = w;
      const double ac = fabs(c);
      argmax_merge(p.v, p.i, ac, ii[q]);
      if (!bit_test(rowbits, ii[q]) && ac > (double)min_pivot && (p.first < 0 || ii[q] < p.first))
        p.first = ii[q];
      p.sq += w * w;
    }
  }
  PivRed r = wg_reduce_piv(p, pr_sh);
  if (t == 0) {
    const int cap = L.cap[b.level];
    slot[cap] = r.v;
    slot[cap + 1] = (double)r.i;
    slot[cap + 2] = (double)r.first;
    slot[cap + 3] = r.sq;
  }
}

__global__ void __launch_bounds__(TPB)
    aca_fin_kernel(AcaLevels L, const AcaBlock* __restrict__ blocks, const int* __restrict__ wg2blk,
                   const int* __restrict__ wg2chunk, const int* __restrict__ nchunk,
                   AcaState* __restrict__ st, uint32_t* __restrict__ bits,
                   double* __restrict__ fin_part, int* __restrict__ counters, double tol,
                   int min_rank, int max_rank, int min_pivot) {
  __shared__ double wrow[FIN_L];
  __shared__ double vrow[FIN_L];
  __shared__ double red[TPB / 64];
  __shared__ int last;
  const int bidx = wg2blk[blockIdx.x];
  const int chunk = wg2chunk[blockIdx.x];
  const AcaBlock b = blocks[bidx];
  const AcaState s = st[bidx];
  if (s.status != 0) return;
  const int t = threadIdx.x;
  const int ld = L.ld;
  const int k = s.k;
  const int i = s.cur_row;
  const int j = s.cur_col;
  const int cap = L.cap[b.level];
  const double* Wp = L.W[b.level] + b.w_row0;
  const double* Vp = L.V[b.level] + b.v_row0;
  double* GV = L.G[b.level] + (size_t)b.lidx * 2 * (size_t)cap * cap;
  double* GW = GV + (size_t)cap * cap;

  double cross = 0.0;
  if (j >= 0) {
    const double pivot = 1.0 / Vp[(size_t)j + (size_t)k * ld];
    const int l = chunk * FIN_L + t;
    double tl = 0.0, ul = 0.0;
    if (l < k) {
      for (int w = 0; w < b.row_nwg; w++) tl += slot_ptr(L, b, w)[l];
      for (int w = 0; w < b.col_nwg; w++) ul += slot_ptr(L, b, b.row_nwg + w)[l];
    }
    // d = t - G_V W(i,:)^T,  e = pivot (u - G_W V(j,:)^T)
    double gd = 0.0, ge = 0.0;
    for (int l0 = 0; l0 < k; l0 += FIN_L) {
      const int lc = min(FIN_L, k - l0);
      __syncthreads();
      if (t < lc) {
        wrow[t] = Wp[(size_t)i + (size_t)(l0 + t) * ld];
        vrow[t] = Vp[(size_t)j + (size_t)(l0 + t) * ld];
      }
      __syncthreads();
      if (l < k) {
        for (int q = 0; q < lc; q++) {
          const size_t o = (size_t)l + (size_t)(l0 + q) * cap;
          gd += GV[o] * wrow[q];
          ge += GW[o] * vrow[q];
        }
      }
    }
    if (l < k) {
      const double d = tl - gd;
      const double e = pivot * (ul - ge);
      GV[(size_t)l + (size_t)k * cap] = d;
      GV[(size_t)k + (size_t)l * cap] = d;
      GW[(size_t)l + (size_t)k * cap] = e;
      GW[(size_t)k + (size_t)l * cap] = e;
      cross = d * e;
    }
  }
  // workgroup sum of the cross-term partial
  for (int off = 32; off > 0; off >>= 1) cross += __shfl_xor(cross, off);
  if ((t & 63) == 0) red[t >> 6] = cross;
  __syncthreads();
  if (t == 0) {
    double c = 0.0;
    for (int w = 0; w < TPB / 64; w++) c += red[w];
    fin_part[(size_t)bidx * MAX_FIN_CHUNKS + chunk] = c;
    __threadfence();
    int old = atomicAdd(&counters[bidx], 1);
    last = (old == nchunk[bidx] - 1);
  }
  __syncthreads();
  if (!last || t != 0) return;
  __threadfence();
  counters[bidx] = 0;
  AcaState ns = s;
  ns.iters = s.iters + 1;
  uint32_t* rowbits = bits + b.bits_row0;
  uint32_t* colbits = bits + b.bits_col0;
  if (j == -1) {
    // max |r| <= minPivot: continue from the first unchosen row (lowRank.cpp:187-195)
    int nr = -1;
    const int nw = (b.m + 31) >> 5;
    for (int w = 0; w < nw && nr < 0; w++) {
      uint32_t x = ~rowbits[w];
      if (w == nw - 1 && (b.m & 31)) x &= (1u << (b.m & 31)) - 1u;
      if (x) nr = w * 32 + __ffs(x) - 1;
    }
    if (nr < 0) {
      ns.status = 1;
    } else {
      ns.cur_row = nr;
      bit_set(rowbits, nr);
    }
    st[bidx] = ns;
    return;
  }
  if (j == -2) {  // no admissible column (:200-202)
    ns.status = 1;
    st[bidx] = ns;
    return;
  }
  bit_set(colbits, j);  // :206
  // next row (:218-227)
  double cv = -1.0;
  int ci = -1, cf = -1;
  for (int w = 0; w < b.col_nwg; w++) {
    double v;
    int ii, f;
    read_piv(slot_ptr(L, b, b.row_nwg + w), cap, v, ii, f);
    argmax_merge(cv, ci, v, ii);
    if (f >= 0 && (cf < 0 || f < cf)) cf = f;
  }
  int next = (ci >= 0 && !bit_test(rowbits, ci)) ? ci : cf;
  if (next < 0) {  // term k is discarded (:226-227)
    ns.status = 1;
    st[bidx] = ns;
    return;
  }
  double crossSum = 0.0;
  for (int c = 0; c < nchunk[bidx]; c++) crossSum += fin_part[(size_t)bidx * MAX_FIN_CHUNKS + c];
  double nv2 = 0.0, nw2 = 0.0;
  for (int w = 0; w < b.row_nwg; w++) nv2 += slot_ptr(L, b, w)[cap + 3];
  for (int w = 0; w < b.col_nwg; w++) nw2 += slot_ptr(L, b, b.row_nwg + w)[cap + 3];
  GV[(size_t)k + (size_t)k * cap] = nv2;
  GW[(size_t)k + (size_t)k * cap] = nw2;
  ns.frob2 = s.frob2 + 2.0 * crossSum + nw2 * nv2;  // :238
  ns.eps = sqrt(nw2) * sqrt(nv2) / sqrt(ns.frob2);  // :241
  ns.k = k + 1;
  ns.cur_row = next;
  bit_set(rowbits, next);
  const int upper = min(b.m, b.n);
  const bool cont = ((ns.eps > tol) || (ns.k < min_rank)) && (ns.k < upper);
  if (!cont || ns.k == max_rank)
    ns.status = 1;
  else if (ns.k >= cap)
    ns.status = 2;
  st[bidx] = ns;
}

/* Post-processing of the returned factors (lowRank.cpp:250-274):
 * rank 0 -> one zero column; rank >= min(m,n)-1 -> the dense block with an
 * identity factor.  One workgroup per block. */
template <class SRC>
__global__ void __launch_bounds__(TPB)
    aca_post_kernel(AcaLevels L, const AcaBlock* __restrict__ blocks, AcaState* __restrict__ st,
                    const SRC* __restrict__ srcs, int* __restrict__ ranks) {
  const int bidx = blockIdx.x;
  const AcaBlock b = blocks[bidx];
  const int k = st[bidx].k;
  const int ld = L.ld;
  double* Wp = L.W[b.level] + b.w_row0;
  double* Vp = L.V[b.level] + b.v_row0;
  const int upper = min(b.m, b.n);
  if (k == 0) {
    for (int i = threadIdx.x; i < b.m; i += TPB) Wp[i] = 0.0;
    for (int jj = threadIdx.x; jj < b.n; jj += TPB) Vp[jj] = 0.0;
    if (threadIdx.x == 0) ranks[bidx] = 1;
    return;
  }
  if (k >= upper - 1) {
    const SRC src = srcs[b.inst];
    if (b.n <= b.m) {  // W = block (m x n), V = I_n
      for (size_t e = threadIdx.x; e < (size_t)b.m * b.n; e += TPB) {
        const int ii = (int)(e % b.m), jj = (int)(e / b.m);
        Wp[(size_t)ii + (size_t)jj * ld] = entry_at(src, b.i0 + ii, b.j0 + jj);
      }
      for (size_t e = threadIdx.x; e < (size_t)b.n * b.n; e += TPB) {
        const int r = (int)(e % b.n), c = (int)(e / b.n);
        Vp[(size_t)r + (size_t)c * ld] = (r == c) ? 1.0 : 0.0;
      }
      if (threadIdx.x == 0) ranks[bidx] = b.n;
    } else {  // W = I_m, V = block^T (n x m)
      for (size_t e = threadIdx.x; e < (size_t)b.m * b.m; e += TPB) {
        const int r = (int)(e % b.m), c = (int)(e / b.m);
        Wp[(size_t)r + (size_t)c * ld] = (r == c) ? 1.0 : 0.0;
      }
      for (size_t e = threadIdx.x; e < (size_t)b.m * b.n; e += TPB) {
        const int ii = (int)(e % b.m), jj = (int)(e / b.m);
        Vp[(size_t)jj + (size_t)ii * ld] = entry_at(src, b.i0 + ii, b.j0 + jj);
      }
      if (threadIdx.x == 0) ranks[bidx] = b.m;
    }
    return;
  }
  if (threadIdx.x == 0) ranks[bidx] = k;
}

}  // namespace

void aca_launch_init(const AcaBlock* d_blocks, AcaState* d_state, uint32_t* d_bits, int nblocks,
                     hipStream_t s) {
  aca_init_kernel<<<(nblocks + TPB - 1) / TPB, TPB, 0, s>>>(d_blocks, d_state, d_bits, nblocks);
  HODLR_HIP_CHECK(hipGetLastError());
}

void aca_launch_iteration(const AcaLevels& L, const AcaLaunch& a, const EntrySource* d_srcs,
                          hipStream_t s) {
  aca_row_kernel<EntrySource><<<a.n_row_wg, TPB, 0, s>>>(L, a.blocks, a.wg2blk_row, a.state,
                                                        a.bits, d_srcs, a.jpt, a.min_pivot);
  HODLR_HIP_CHECK(hipGetLastError());
  aca_col_kernel<EntrySource><<<a.n_col_wg, TPB, 0, s>>>(L, a.blocks, a.wg2blk_col, a.state,
                                                        a.bits, d_srcs, a.ipt, a.min_pivot);
  HODLR_HIP_CHECK(hipGetLastError());
  aca_fin_kernel<<<a.n_fin_wg, TPB, 0, s>>>(L, a.blocks, a.wg2blk_fin, a.wg2chunk_fin,
                                             a.nchunk, a.state, a.bits, a.fin_part, a.counters,
                                             a.tol, a.min_rank, a.max_rank, a.min_pivot);
  HODLR_HIP_CHECK(hipGetLastError());
}

void aca_launch_post(const AcaLevels& L, const AcaBlock* d_blocks, AcaState* d_state,
                     const EntrySource* d_srcs, int* d_ranks, int nblocks, hipStream_t s) {
  aca_post_kernel<EntrySource><<<nblocks, TPB, 0, s>>>(L, d_blocks, d_state, d_srcs, d_ranks);
  HODLR_HIP_CHECK(hipGetLastError());
}

}  // namespace hodlr
