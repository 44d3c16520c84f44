/*
 * Batched dense kernels: one launch covers every matrix of a tree level (or
 * every leaf), one workgroup per matrix or per right-hand-side chunk.
 */
#include "dense.h"

#include <algorithm>

namespace hodlr {

namespace {

constexpr int TPB = 256;

__global__ void __launch_bounds__(TPB)
    assemble_kernel(const AssembleDesc* __restrict__ d, const EntrySource* __restrict__ srcs) {
  const AssembleDesc a = d[blockIdx.y];
  const EntrySource src = srcs[a.inst];
  const size_t total = (size_t)a.m * a.n;
  for (size_t e = (size_t)blockIdx.x * TPB + threadIdx.x; e < total; e += (size_t)gridDim.x * TPB) {
    const int i = (int)(e % a.m), j = (int)(e / a.m);
    const double v = entry_at(src, a.i0 + i, a.j0 + j);
    if (a.trans)
      a.out[(size_t)j + (size_t)i * a.ldo] = v;
    else
      a.out[(size_t)i + (size_t)j * a.ldo] = v;
  }
}

/* entries of low-rank pieces, the product summed sequentially over l */
__global__ void __launch_bounds__(TPB) lr_fill_kernel(const LrFillDesc* __restrict__ d) {
  const LrFillDesc a = d[blockIdx.y];
  const size_t total = (size_t)a.m * a.n;
  for (size_t e = (size_t)blockIdx.x * TPB + threadIdx.x; e < total; e += (size_t)gridDim.x * TPB) {
    const int i = (int)(e % a.m), j = (int)(e / a.m);
    double acc = 0.0;
    for (int l = 0; l < a.k; l++) acc += a.U[(size_t)i + (size_t)l * a.ldu] * a.V[(size_t)j + (size_t)l * a.ldv];
    a.out[(size_t)i + (size_t)j * a.ldo] = acc;
  }
}

__global__ void __launch_bounds__(TPB) identity_kernel(const CopyDesc* __restrict__ d) {
  const CopyDesc c = d[blockIdx.y];
  const size_t tot = (size_t)c.m * c.n;
  for (size_t e = (size_t)blockIdx.x * TPB + threadIdx.x; e < tot; e += (size_t)gridDim.x * TPB) {
    const int i = (int)(e % c.m), j = (int)(e / c.m);
    c.dst[(size_t)i + (size_t)j * c.ldd] = (i == j) ? 1.0 : 0.0;
  }
}

__global__ void __launch_bounds__(TPB) copy_kernel(const CopyDesc* __restrict__ d) {
  const CopyDesc c = d[blockIdx.y];
  const size_t tot = (size_t)c.m * c.n;
  for (size_t e = (size_t)blockIdx.x * TPB + threadIdx.x; e < tot; e += (size_t)gridDim.x * TPB) {
    const int i = (int)(e % c.m), j = (int)(e / c.m);
    c.dst[(size_t)i + (size_t)j * c.ldd] = c.src[(size_t)i + (size_t)j * c.lds];
  }
}

/* Per matrix: sum log|u_ii| and prod u_ii, summed in index order by one thread
 * after a fixed-shape tree reduction (deterministic). */
__global__ void __launch_bounds__(TPB)
    logdet_kernel(const LuDesc* __restrict__ d, double* out_log, double* out_prod) {
  __shared__ double sl[TPB];
  __shared__ double sp[TPB];
  const LuDesc m = d[blockIdx.x];
  const int t = threadIdx.x;
  double lg = 0.0, pr = 1.0;
  for (int i = t; i < m.n; i += TPB) {
    const double u = m.A[(size_t)i + (size_t)i * m.ld];
    lg += log(fabs(u));
    pr *= u;
  }
  sl[t] = lg;
  sp[t] = pr;
  __syncthreads();
  for (int s = TPB / 2; s > 0; s >>= 1) {
    if (t < s) {
      sl[t] += sl[t + s];
      sp[t] *= sp[t + s];
    }
    __syncthreads();
  }
  if (t == 0) {
    out_log[blockIdx.x] = sl[0];
    out_prod[blockIdx.x] = sp[0];
  }
}

/* Build the node's K x K matrix and right-hand side from its Gram products.
 *   recSM (perturbI.cpp:29-37): C = [[I_kt, V_t^T U_b^SM], [V_b^T U_t^SM, I_kb]],
 *                               Y = [V_t^T X_bot ; V_b^T X_top]
 *   recLU (HODLR_Matrix.cpp:663-673): S = [[I_kb, V_b^T U_t^SM], [V_t^T U_b^SM, I_kt]],
 *                               Y = [V_b^T X_top ; V_t^T X_bot]            */
template <class TZ>
__global__ void __launch_bounds__(TPB) capacitance_kernel(const CapDescT<TZ>* __restrict__ d) {
  const CapDescT<TZ> c = d[blockIdx.y];
  const int K = c.kt + c.kb;
  const int ka = c.lu_order ? c.kb : c.kt;  // rows of the first block
  const size_t totC = c.C ? (size_t)K * K : 0;  // solve phase: right-hand side only
  for (size_t e = (size_t)blockIdx.x * TPB + threadIdx.x; e < totC; e += (size_t)gridDim.x * TPB) {
    const int i = (int)(e % K), j = (int)(e / K);
    double v;
    if (i < ka && j < ka) {
      v = (i == j) ? 1.0 : 0.0;
    } else if (i >= ka && j >= ka) {
      v = (i == j) ? 1.0 : 0.0;
    } else if (i < ka) {  // upper-right block
      const int jj = j - ka;
      v = c.lu_order ? (double)c.Zt[(size_t)i + (size_t)(c.own0 + jj) * c.kb]
                     : (double)c.Zb[(size_t)i + (size_t)(c.own0 + jj) * c.kt];
    } else {  // lower-left block
      const int ii = i - ka;
      v = c.lu_order ? (double)c.Zb[(size_t)ii + (size_t)(c.own0 + j) * c.kt]
                     : (double)c.Zt[(size_t)ii + (size_t)(c.own0 + j) * c.kb];
    }
    c.C[e] = v;
  }
  const size_t totY = (size_t)K * c.nanc;
  for (size_t e = (size_t)blockIdx.x * TPB + threadIdx.x; e < totY; e += (size_t)gridDim.x * TPB) {
    const int i = (int)(e % K), j = (int)(e / K);
    double v;
    if (i < ka)
      v = c.lu_order ? (double)c.Zt[(size_t)i + (size_t)j * c.kb]
                     : (double)c.Zb[(size_t)i + (size_t)j * c.kt];
    else
      v = c.lu_order ? (double)c.Zb[(size_t)(i - ka) + (size_t)j * c.kt]
                     : (double)c.Zt[(size_t)(i - ka) + (size_t)j * c.kb];
    c.Y[e] = v;
  }
}

/* Panel step of the blocked right-looking LU: factor columns [p0, p0+pw) of
 * rows [p0, n) with partial pivoting (first maximum), swapping whole rows so
 * the trailing columns follow (LAPACK getrf order).  One workgroup per
 * matrix; matrices with n <= p0 are done. */
__global__ void __launch_bounds__(TPB) lu_panel_kernel(const LuDesc* __restrict__ d, int p0,
                                                       int nb) {
  __shared__ double sv[TPB / 64];
  __shared__ int si[TPB / 64];
  __shared__ int piv_sh;
  const LuDesc m = d[blockIdx.x];
  const int n = m.n, ld = m.ld, t = threadIdx.x;
  if (n <= p0) return;
  double* A = m.A;
  const int pe = min(n, p0 + nb);
  for (int c = p0; c < pe; c++) {
    double bv = -1.0;
    int bi = n;
    for (int i = c + t; i < n; i += TPB) {
      const double v = fabs(A[(size_t)i + (size_t)c * ld]);
      if (v > bv) {
        bv = v;
        bi = i;
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_xor(bv, off);
      const int oi = __shfl_xor(bi, off);
      if (ov > bv || (ov == bv && oi < bi)) {
        bv = ov;
        bi = oi;
      }
    }
    if ((t & 63) == 0) {
      sv[t >> 6] = bv;
      si[t >> 6] = bi;
    }
    __syncthreads();
    if (t == 0) {
      double v = sv[0];
      int ix = si[0];
      for (int w = 1; w < TPB / 64; w++)
        if (sv[w] > v || (sv[w] == v && si[w] < ix)) {
          v = sv[w];
          ix = si[w];
        }
      // a column of NaNs has no maximum: keep row c (never an index >= n,
      // which would swap with memory past the matrix)
      if (ix >= n) ix = c;
      piv_sh = ix;
      m.piv[c] = ix;
    }
    __syncthreads();
    const int p = piv_sh;
    if (p != c)
      for (int j = t; j < n; j += TPB) {
        const double x = A[(size_t)c + (size_t)j * ld];
        A[(size_t)c + (size_t)j * ld] = A[(size_t)p + (size_t)j * ld];
        A[(size_t)p + (size_t)j * ld] = x;
      }
    __syncthreads();
    const double pv = A[(size_t)c + (size_t)c * ld];
    if (pv != 0.0) {
      const double inv = 1.0 / pv;
      for (int i = c + 1 + t; i < n; i += TPB) A[(size_t)i + (size_t)c * ld] *= inv;
    }
    __syncthreads();
    const int r = n - c - 1, w = pe - c - 1;  // update the rest of the panel only
    const size_t tot = (size_t)r * w;
    for (size_t e = t; e < tot; e += TPB) {
      const int i = c + 1 + (int)(e % r), j = c + 1 + (int)(e / r);
      A[(size_t)i + (size_t)j * ld] -= A[(size_t)i + (size_t)c * ld] * A[(size_t)c + (size_t)j * ld];
    }
    __syncthreads();
  }
}

/* (v, i) <- the wave's first maximum (largest v, smallest i among equal v),
 * in every lane.  Within rows of 16 lanes by DPP moves (pairs swapped,
 * pairs of pairs swapped, half-row mirror, row mirror: after each step a
 * lane holds the best of a group twice as large), then the four rows'
 * results through readlane.  The order of the comparisons does not matter:
 * "first maximum" is a total order on distinct indices. */
template <int CTRL>
__device__ __forceinline__ void first_max_step(double& v, int& i) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
  const int oi = __builtin_amdgcn_update_dpp(0, i, CTRL, 0xF, 0xF, false);
  const double ov =
      __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
  if (ov > v || (ov == v && oi < i)) {
    v = ov;
    i = oi;
  }
}

__device__ __forceinline__ double read_lane_f64(double v, int lane) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, lane);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ void wave_first_max(double& v, int& i) {
  first_max_step<0xB1>(v, i);   // quad_perm [1, 0, 3, 2]
  first_max_step<0x4E>(v, i);   // quad_perm [2, 3, 0, 1]
  first_max_step<0x141>(v, i);  // row_half_mirror
  first_max_step<0x140>(v, i);  // row_mirror
  double wv = read_lane_f64(v, 0);
  int wi = __builtin_amdgcn_readlane(i, 0);
#pragma unroll
  for (int r = 16; r < 64; r += 16) {
    const double ov = read_lane_f64(v, r);
    const int oi = __builtin_amdgcn_readlane(i, r);
    if (ov > wv || (ov == wv && oi < wi)) {
      wv = ov;
      wi = oi;
    }
  }
  v = wv;
  i = wi;
}

/* The same panel step with the panel held in registers: thread t owns rows
 * p0 + t + NT q (q < RPT) of the panel columns [p0, p0 + pw), pw <= PW.
 * Per column: first-max pivot search over the owned rows (wave shuffles,
 * then LDS), exchange of the two rows through LDS, broadcast of the pivot
 * row, scaling and rank-1 update in registers.  The row interchanges of the
 * columns outside the panel are recorded for lu_step_kernel (a row of a
 * column-major matrix is n separate cache lines: one workgroup swapping
 * them column by column took half the panel step).  Used when the largest matrix of the batch has at most
 * NT * RPT rows. */
template <int NT, int RPT, int PW>
__global__ void __launch_bounds__(NT) lu_panel_reg_kernel(const LuDesc* __restrict__ d, int p0,
                                                          int nb) {
  constexpr int NW = NT / 64;
  // two barriers per column: LDS buffers alternate by column parity, every
  // thread finishes the pivot search from the waves' partials itself, and
  // the pivot row and the row it replaces are published together
  __shared__ double sv[2][NW];
  __shared__ int si[2][NW];
  __shared__ double prow[2][PW];  // old row p: the pivot row after the exchange
  __shared__ double crow[2][PW];  // old row cr: moves to row p
  __shared__ int piv_sh[PW];
  __shared__ int perm[NT * RPT];
  const LuDesc m = d[blockIdx.x];
  const int n = m.n, ld = m.ld, t = threadIdx.x;
  if (n <= p0) return;
  const gptr<double> A = to_global(m.A);
  const int pw = min(min(nb, PW), n - p0);
  // Rows as vectors: the column c of the loop (uniform, not unrolled) is
  // read and written by register indexing (one index-mode move per value)
  // and the rank-1 update runs over the whole row against the pivot row
  // with its entries at and left of c zeroed, so no per-element selects.
  // (rows of more than 16 values as several vectors: a larger vector
  // indexed at run time goes to scratch memory)
  constexpr int W = PW > 16 ? 16 : PW, NP = PW / W;
  static_assert(NP <= 2, "at most two vectors per row");
  typedef double RowVec __attribute__((ext_vector_type(W)));
  RowVec R[RPT][NP];
#pragma unroll
  for (int q = 0; q < RPT; q++) {
    const int r = p0 + t + NT * q;
#pragma unroll
    for (int j = 0; j < PW; j++)
      R[q][j / W][j % W] = (r < n && j < pw) ? A[(size_t)r + (size_t)(p0 + j) * ld] : 0.0;
  }
  for (int c = 0; c < pw; c++) {
    const int cr = p0 + c, buf = c & 1;
    double rc[RPT];  // this thread's entries of column c
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      rc[q] = R[q][0][c & (W - 1)];
      if (NP > 1 && c >= W) rc[q] = R[q][NP - 1][c & (W - 1)];
    }
    // first maximum of |A(i, cr)|, i >= cr
    double bv = -1.0;
    int bi = n;
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      const int r = p0 + t + NT * q;
      if (r >= cr && r < n) {
        const double v = fabs(rc[q]);
        if (v > bv) {
          bv = v;
          bi = r;
        }
      }
    }
    wave_first_max(bv, bi);
    if ((t & 63) == 0) {
      sv[buf][t >> 6] = bv;
      si[buf][t >> 6] = bi;
    }
    __syncthreads();
    double v = sv[buf][0];
    int p = si[buf][0];
    for (int w = 1; w < NW; w++)
      if (sv[buf][w] > v || (sv[buf][w] == v && si[buf][w] < p)) {
        v = sv[buf][w];
        p = si[buf][w];
      }
    if (p >= n) p = cr;  // a column of NaNs has no maximum: keep the row
    if (t == 0) piv_sh[c] = p;  // read after the loop, behind this column's second barrier
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      const int r = p0 + t + NT * q;
      if (r == p)
#pragma unroll
        for (int j = 0; j < PW; j++) prow[buf][j] = R[q][j / W][j % W];
      if (r == cr)
#pragma unroll
        for (int j = 0; j < PW; j++) crow[buf][j] = R[q][j / W][j % W];
    }
    __syncthreads();
    // exchange rows cr and p of the panel
    if (p != cr) {
#pragma unroll
      for (int q = 0; q < RPT; q++) {
        const int r = p0 + t + NT * q;
        if (r == cr) {
#pragma unroll
          for (int j = 0; j < PW; j++) R[q][j / W][j % W] = prow[buf][j];
        } else if (r == p) {
#pragma unroll
          for (int j = 0; j < PW; j++) R[q][j / W][j % W] = crow[buf][j];
          rc[q] = crow[buf][c];
        }
      }
    }
    // rows below the pivot: multiplier l = A(i, cr) / pivot (kept unscaled
    // for a zero pivot), A(i, j) -= l A(cr, j) for j > c, then A(i, cr) = l.
    // The update runs over the whole row and subtracts l * 0 at j <= c,
    // which leaves those entries as they are (for a finite l; an entry -0
    // may turn +0).
    const double pv = prow[buf][c];
    const double inv = pv != 0.0 ? 1.0 / pv : 1.0;
    double pm[PW];
#pragma unroll
    for (int j = 0; j < PW; j++) pm[j] = j > c ? prow[buf][j] : 0.0;
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      const int r = p0 + t + NT * q;
      if (r > cr && r < n) {
        const double l = pv != 0.0 ? rc[q] * inv : rc[q];
#pragma unroll
        for (int j = 0; j < PW; j++) R[q][j / W][j % W] = fma(-l, pm[j], R[q][j / W][j % W]);
        if (NP > 1 && c >= W)
          R[q][NP - 1][c & (W - 1)] = l;
        else
          R[q][0][c & (W - 1)] = l;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < RPT; q++) {
    const int r = p0 + t + NT * q;
    if (r < n) {
#pragma unroll
      for (int j = 0; j < PW; j++)
        if (j < pw) A[(size_t)r + (size_t)(p0 + j) * ld] = R[q][j / W][j % W];
    }
  }
  // The interchanges of the columns outside the panel are applied by
  // lu_step_kernel, many workgroups wide: here they are composed into one
  // permutation (row p0 + o then holds the old row p0 + perm[o]), and the
  // rows it moves are recorded: the panel rows and the pivot rows, as
  // destinations with their sources (a pivot row may repeat: same source).
  for (int o = t; o < n - p0; o += NT) perm[o] = o;
  __syncthreads();
  if (t == 0)
    for (int c = 0; c < pw; c++) {
      const int q = piv_sh[c] - p0;
      if (q != c) {
        const int x = perm[c];
        perm[c] = perm[q];
        perm[q] = x;
      }
    }
  __syncthreads();
  if (t < pw) to_global(m.piv)[p0 + t] = piv_sh[t];
  if (t < 2 * pw) {
    const int dr = t < pw ? p0 + t : piv_sh[t - pw];
    to_global(m.swp)[t] = dr;
    to_global(m.swp)[LU_SWP_SRC + t] = p0 + perm[dr - p0];
  }
}

/* The blocked LU's work outside the panel [p0, p0 + pw), 64 columns per
 * workgroup of 256 threads: the panel's interchanges (the recorded 2 pw
 * moves of each column, spread over all four waves: all loads, a barrier,
 * all stores), then for the columns right of the panel U12 = L11^-1 A12 by
 * forward substitution in the first wave, one column per thread, operation
 * for operation as trsm_kernel.  PWM is the widest panel: 32 for the
 * register panels (the only ones with recorded interchanges; the moves are
 * staged in registers and the LDS holds 25 KB, so six workgroups fit on a
 * CU), 64 for the global-memory panel (no interchanges here). */
constexpr int LU_STEP_NT = 256;

template <int PWM>
__global__ void __launch_bounds__(LU_STEP_NT) lu_step_kernel(const LuStepDesc* __restrict__ d,
                                                              int count, int swap) {
  __shared__ double T[PWM][PWM + 1];
  __shared__ double X[PWM][65];
  __shared__ int sd[2 * 32], ss[2 * 32];
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].chunk0 <= (int)blockIdx.x)
      lo = mid;
    else
      hi = mid - 1;
  }
  const LuStepDesc p = d[lo];
  const gptr<double> pA = to_global(p.A);
  const int t = threadIdx.x, pw = p.pw, p0 = p.p0;
  const int ch = (int)blockIdx.x - p.chunk0;
  const bool left = ch < p.nleft;
  const int j0 = left ? ch * 64 : p0 + pw + (ch - p.nleft) * 64;  // first column of the chunk
  const int jend = left ? p0 : p.n;
  // (recorded only by the register panels: pw <= 32)
  const int nm = (PWM == 32 && swap) ? 2 * pw : 0;
  for (int u = t; u < nm; u += LU_STEP_NT) {
    sd[u] = to_global(p.swp)[u];
    ss[u] = to_global(p.swp)[LU_SWP_SRC + u];
  }
  if (!left)
    for (int e = t; e < pw * pw; e += LU_STEP_NT) {
      const int r = e % pw, c = e / pw;
      T[r][c] = pA[(size_t)(p0 + r) + (size_t)(p0 + c) * p.ld];
    }
  __syncthreads();
  if (PWM == 32 && nm > 0) {
    constexpr int MQ = 64 * 64 / LU_STEP_NT;
    double v[MQ];
#pragma unroll
    for (int q = 0; q < MQ; q++) {
      const int e = t + LU_STEP_NT * q, u = e >> 6, c = e & 63;
      const bool ok = u < nm && j0 + c < jend && ss[u] != sd[u];
      v[q] = ok ? pA[(size_t)ss[u] + (size_t)(j0 + c) * p.ld] : 0.0;
    }
    // every source value is in registers before any row is overwritten
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MQ; q++) {
      const int e = t + LU_STEP_NT * q, u = e >> 6, c = e & 63;
      if (u < nm && j0 + c < jend && ss[u] != sd[u]) pA[(size_t)sd[u] + (size_t)(j0 + c) * p.ld] = v[q];
    }
    __syncthreads();  // the substitution below reads rows the stores wrote
  }
  if (left || t >= 64) return;
  const int j = j0 + t;
  const bool ok = j < jend;
  const gptr<double> col = pA + (size_t)j * p.ld;
  for (int r = 0; r < pw; r++) X[r][t] = ok ? col[p0 + r] : 0.0;
  for (int r = 1; r < pw; r++) {
    double v = X[r][t];
    for (int q = 0; q < r; q++) v -= T[r][q] * X[q][t];
    X[r][t] = v;
  }
  if (ok)
    for (int r = 0; r < pw; r++) col[p0 + r] = X[r][t];
}

/* Triangular solve of a pw x pw diagonal block (pw <= 64) on a chunk of 64
 * right-hand-side columns: unit lower (forward) or upper (backward).  One
 * wave per chunk, one column per lane, the block and the columns in LDS. */
__global__ void __launch_bounds__(64) trsm_kernel(const TrsmDesc* __restrict__ d, int count) {
  __shared__ double T[64][65];
  __shared__ double X[64][65];
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].chunk0 <= (int)blockIdx.x)
      lo = mid;
    else
      hi = mid - 1;
  }
  const TrsmDesc p = d[lo];
  const int c0 = ((int)blockIdx.x - p.chunk0) * 64;
  const int t = threadIdx.x, pw = p.pw;
  for (int e = t; e < pw * pw; e += 64) {
    const int r = e % pw, c = e / pw;
    T[r][c] = p.T[(size_t)r + (size_t)c * p.ldt];
  }
  const int col = c0 + t;
  const bool ok = col < p.ncols;
  for (int r = 0; r < pw; r++) X[r][t] = ok ? p.B[(size_t)r + (size_t)col * p.ldb] : 0.0;
  __syncthreads();
  if (!p.upper) {
    for (int r = 1; r < pw; r++) {
      double v = X[r][t];
      for (int q = 0; q < r; q++) v -= T[r][q] * X[q][t];
      X[r][t] = v;
    }
  } else {
    for (int r = pw - 1; r >= 0; r--) {
      double v = X[r][t];
      for (int q = r + 1; q < pw; q++) v -= T[r][q] * X[q][t];
      X[r][t] = v / T[r][r];
    }
  }
  if (ok)
    for (int r = 0; r < pw; r++) p.B[(size_t)r + (size_t)col * p.ldb] = X[r][t];
}

/* Row interchanges of an LU applied to right-hand sides, one column per thread. */
__global__ void __launch_bounds__(TPB) permute_kernel(const PermDesc* __restrict__ d, int count) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].chunk0 <= (int)blockIdx.x)
      lo = mid;
    else
      hi = mid - 1;
  }
  const PermDesc p = d[lo];
  const int col = ((int)blockIdx.x - p.chunk0) * TPB + threadIdx.x;
  if (col >= p.ncols) return;
  double* b = p.B + (size_t)col * p.ldb;
  for (int k = 0; k < p.n; k++) {
    const int q = p.piv[k];
    if (q != k) {
      const double x = b[k];
      b[k] = b[q];
      b[q] = x;
    }
  }
}

template <class S, class D>
__global__ void __launch_bounds__(TPB) convert_kernel(const ConvDesc<S, D>* __restrict__ d) {
  const ConvDesc<S, D> c = d[blockIdx.y];
  const size_t tot = (size_t)c.m * c.n;
  for (size_t e = (size_t)blockIdx.x * TPB + threadIdx.x; e < tot; e += (size_t)gridDim.x * TPB) {
    const int i = (int)(e % c.m), j = (int)(e / c.m);
    c.dst[(size_t)i + (size_t)j * c.ldd] = (D)c.src[(size_t)i + (size_t)j * c.lds];
  }
}

/* Max-norms of u_c and v_c (no overflow for any finite input), then the
 * rebalanced FP32 copies. */
__global__ void __launch_bounds__(TPB) balance_kernel(const BalanceDesc* __restrict__ d) {
  const BalanceDesc b = d[blockIdx.y];
  const int c = blockIdx.x;
  if (c >= b.k) return;
  const double* u = b.U + (size_t)c * b.ldu;
  const double* v = b.V + (size_t)c * b.ldv;
  double mu = 0.0, mv = 0.0;
  for (int i = threadIdx.x; i < b.mu; i += TPB) mu = fmax(mu, fabs(u[i]));
  for (int i = threadIdx.x; i < b.mv; i += TPB) mv = fmax(mv, fabs(v[i]));
  __shared__ double su[TPB], sv[TPB];
  su[threadIdx.x] = mu;
  sv[threadIdx.x] = mv;
  __syncthreads();
  for (int w = TPB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      su[threadIdx.x] = fmax(su[threadIdx.x], su[threadIdx.x + w]);
      sv[threadIdx.x] = fmax(sv[threadIdx.x], sv[threadIdx.x + w]);
    }
    __syncthreads();
  }
  mu = su[0];
  mv = sv[0];
  const double s = (mu > 0.0 && mv > 0.0) ? sqrt(mv / mu) : 1.0;
  const double is = 1.0 / s;
  float* u32 = b.U32 + (size_t)c * b.ldu32;
  float* v32 = b.V32 + (size_t)c * b.ldv32;
  for (int i = threadIdx.x; i < b.mu; i += TPB) u32[i] = (float)(u[i] * s);
  for (int i = threadIdx.x; i < b.mv; i += TPB) v32[i] = (float)(v[i] * is);
}

/* squared norms of rows [row0[s], row0[s+1]) of column c: out[s + c * nseg] */
__global__ void __launch_bounds__(TPB)
    segnorm2_kernel(const double* __restrict__ X, int ld, const int* __restrict__ row0, int nseg,
                    double* __restrict__ out) {
  __shared__ double sh[TPB];
  const int sg = blockIdx.y, c = blockIdx.x, sp = blockIdx.z, nsp = gridDim.z;
  const double* x = X + (size_t)c * ld;
  // split sp of nsp of the segment's rows; four loads in flight per thread
  const int r0 = row0[sg], len = row0[sg + 1] - r0;
  const int part = (len + nsp - 1) / nsp;
  const int lo = r0 + min(len, sp * part), hi = r0 + min(len, (sp + 1) * part);
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int i = lo + threadIdx.x;
  for (; i + 3 * TPB < hi; i += 4 * TPB) {
    const double v0 = x[i], v1 = x[i + TPB], v2 = x[i + 2 * TPB], v3 = x[i + 3 * TPB];
    a0 += v0 * v0;
    a1 += v1 * v1;
    a2 += v2 * v2;
    a3 += v3 * v3;
  }
  for (; i < hi; i += TPB) a0 += x[i] * x[i];
  const double acc = (a0 + a1) + (a2 + a3);
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = TPB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[sp + (size_t)nsp * (sg + (size_t)c * nseg)] = sh[0];
}

__global__ void __launch_bounds__(TPB)
    segscale_kernel(double* __restrict__ X, int ld, const int* __restrict__ row0, int nseg,
                    const double* __restrict__ sc) {
  const int sg = blockIdx.y, c = blockIdx.x;
  const double f = sc[sg + (size_t)c * nseg];
  double* x = X + (size_t)c * ld;
  for (int i = row0[sg] + threadIdx.x; i < row0[sg + 1]; i += TPB) x[i] *= f;
}

/* rows [row0[s], row0[s+1]) of column c copied from X to Y where
 * mask[s + c * nseg] is set; split z of the grid's z parts of the rows */
__global__ void __launch_bounds__(TPB)
    segcopy_kernel(const double* __restrict__ X, double* __restrict__ Y, int ld,
                   const int* __restrict__ row0, int nseg, const int* __restrict__ mask) {
  const int sg = blockIdx.y, c = blockIdx.x, sp = blockIdx.z, nsp = gridDim.z;
  if (!mask[sg + (size_t)c * nseg]) return;
  const int r0 = row0[sg], len = row0[sg + 1] - r0;
  const int part = (len + nsp - 1) / nsp;
  const int lo = r0 + min(len, sp * part), hi = r0 + min(len, (sp + 1) * part);
  const double* x = X + (size_t)c * ld;
  double* y = Y + (size_t)c * ld;
  for (int i = lo + threadIdx.x; i < hi; i += TPB) y[i] = x[i];
}

__global__ void __launch_bounds__(TPB)
    colnorm2_kernel(const double* __restrict__ X, int n, int ld, double* __restrict__ out) {
  __shared__ double sh[TPB];
  const double* x = X + (size_t)blockIdx.x * ld;
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += TPB) s += x[i] * x[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = TPB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(TPB) add_f32_kernel(double* __restrict__ X, int ldx,
                                                      const float* __restrict__ D, int ldd, int n,
                                                      int ncols) {
  const size_t tot = (size_t)n * ncols;
  for (size_t e = (size_t)blockIdx.x * TPB + threadIdx.x; e < tot; e += (size_t)gridDim.x * TPB) {
    const int i = (int)(e % n), j = (int)(e / n);
    X[(size_t)i + (size_t)j * ldx] += (double)D[(size_t)i + (size_t)j * ldd];
  }
}

__global__ void __launch_bounds__(TPB) sub_kernel(const double* __restrict__ B, int ldb,
                                                  const double* __restrict__ Y, int ldy,
                                                  double* __restrict__ R, int ldr, int n, int ncols) {
  const size_t tot = (size_t)n * ncols;
  for (size_t e = (size_t)blockIdx.x * TPB + threadIdx.x; e < tot; e += (size_t)gridDim.x * TPB) {
    const int i = (int)(e % n), j = (int)(e / n);
    R[(size_t)i + (size_t)j * ldr] = B[(size_t)i + (size_t)j * ldb] - Y[(size_t)i + (size_t)j * ldy];
  }
}

/* out[i + c*nvec] = V_i(:,c) . w(:,c) for basis vectors V_i = V + i*vstride. */
__global__ void __launch_bounds__(TPB)
    multi_dot_kernel(const double* __restrict__ V, size_t vstride, int nvec,
                     const double* __restrict__ w, int n, int ld, double* __restrict__ out) {
  __shared__ double sh[TPB];
  const int c = blockIdx.x, i = blockIdx.y;
  const double* v = V + (size_t)i * vstride + (size_t)c * ld;
  const double* x = w + (size_t)c * ld;
  double s = 0.0;
  for (int r = threadIdx.x; r < n; r += TPB) s += v[r] * x[r];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int h = TPB / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[i + (size_t)c * nvec] = sh[0];
}

/* w(:,c) = (assign ? 0 : w(:,c)) + sign * sum_i V_i(:,c) coef(i,c) */
__global__ void __launch_bounds__(TPB)
    multi_axpy_kernel(const double* __restrict__ V, size_t vstride, int nvec,
                      const double* __restrict__ coef, double* __restrict__ w, int n, int ld,
                      int ncols, double sign, int assign) {
  const size_t tot = (size_t)n * ncols;
  for (size_t e = (size_t)blockIdx.x * TPB + threadIdx.x; e < tot; e += (size_t)gridDim.x * TPB) {
    const int r = (int)(e % n), c = (int)(e / n);
    double s = 0.0;
    for (int i = 0; i < nvec; i++)
      s += V[(size_t)i * vstride + r + (size_t)c * ld] * coef[i + (size_t)c * nvec];
    double* o = w + r + (size_t)c * ld;
    *o = (assign ? 0.0 : *o) + sign * s;
  }
}

/* x <- fl(x + d) and d <- the exact step x_new - x_old rounded once:
 * TwoSum gives x_old + d = x_new + e exactly, so x_new - x_old = d - e. */
__global__ void __launch_bounds__(TPB) step_update_kernel(double* __restrict__ x,
                                                          double* __restrict__ d, int ld, int n,
                                                          int ncols) {
#pragma clang fp contract(off)
  const size_t tot = (size_t)n * ncols;
  for (size_t e = (size_t)blockIdx.x * TPB + threadIdx.x; e < tot; e += (size_t)gridDim.x * TPB) {
    const size_t o = (size_t)(e % n) + (size_t)(e / n) * ld;
    const double a = x[o], b = d[o];
    const double s = a + b;
    const double bb = s - a;
    const double err = (a - (s - bb)) + (b - bb);
    x[o] = s;
    d[o] = b - err;
  }
}

__global__ void __launch_bounds__(TPB) col_scale_kernel(double* __restrict__ X, int ld, int n,
                                                        int ncols, const double* __restrict__ s) {
  const size_t tot = (size_t)n * ncols;
  for (size_t e = (size_t)blockIdx.x * TPB + threadIdx.x; e < tot; e += (size_t)gridDim.x * TPB) {
    const int r = (int)(e % n), c = (int)(e / n);
    X[r + (size_t)c * ld] *= s[c];
  }
}

template <class T>
__global__ void __launch_bounds__(TPB) nonfinite_kernel(const T* __restrict__ x, size_t n,
                                                        unsigned long long* out) {
  unsigned long long c = 0;
  for (size_t e = (size_t)blockIdx.x * TPB + threadIdx.x; e < n; e += (size_t)gridDim.x * TPB)
    if (!isfinite((double)x[e])) c++;
  if (c) atomicAdd(out, c);
}

int grid_x(size_t elems, int ndesc) {
  size_t g = (elems + TPB - 1) / TPB;
  size_t cap = ndesc > 64 ? 16 : 1024;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return (int)g;
}

/* grid.y is limited to 65535: descriptors are launched in chunks */
/* Whole LU solve A X = B (row interchanges, unit lower, upper) of small
 * systems in one launch: one workgroup per problem and CB columns of B.
 * Thread t keeps rows t + 256 r (r < R) of its columns in registers for the
 * whole solve; each substitution step the owner of the pivot row publishes
 * x_l through LDS (double-buffered by step parity, one barrier per step) and
 * every thread updates its rows with column l of the factor.  The factor's
 * columns are loaded D steps ahead into a register ring (the step loop is
 * unrolled by D, so the ring is indexed statically): a step then waits for
 * a load issued D steps earlier, not for one issued in the step before.
 * Fewer columns per workgroup for larger systems (more workgroups for the
 * few large capacitance systems of the upper tree levels). */
constexpr int solve_rows(int maxn) { return maxn <= 256 ? 1 : maxn <= 512 ? 2 : maxn <= 1024 ? 4 : 8; }
constexpr int solve_cols(int R) { return R == 1 ? 8 : R == 8 ? 2 : 4; }
constexpr int SOLVE_D = 8;

template <int R>
__global__ void __launch_bounds__(TPB) lu_solve_small_kernel(const SolveDesc* __restrict__ d,
                                                             int count) {
  constexpr int CB = solve_cols(R), D = SOLVE_D;
  __shared__ int perm[R * TPB];
  __shared__ int pv[R * TPB];
  __shared__ double xs[2][CB];
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].chunk0 <= (int)blockIdx.x)
      lo = mid;
    else
      hi = mid - 1;
  }
  const SolveDesc p = d[lo];
  const gptr<double> pB = to_global(p.B);
  const int n = p.n, t = threadIdx.x;
  const int c0 = ((int)blockIdx.x - p.chunk0) * CB;
  const int nc = min(CB, p.ncols - c0);
  // the interchanges piv[0..n), applied in order, as one permutation (read
  // by all threads, replayed from LDS by one)
  for (int i = t; i < n; i += TPB) {
    perm[i] = i;
    pv[i] = to_global(p.piv)[i];
  }
  __syncthreads();
  if (t == 0)
    for (int k = 0; k < n; k++) {
      const int q = pv[k];
      if (q != k) {
        const int x = perm[k];
        perm[k] = perm[q];
        perm[q] = x;
      }
    }
  __syncthreads();
  double x[R][CB];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int i = t + TPB * r;
#pragma unroll
    for (int c = 0; c < CB; c++)
      x[r][c] = (i < n && c < nc) ? pB[(size_t)perm[i] + (size_t)(c0 + c) * p.ldb] : 0.0;
  }
  const gptr<const double> LU = to_global(p.LU);
  const size_t ld = (size_t)p.ld_lu;
  double fr[D][R];
  // forward: unit lower triangle; fr[s] holds column l (rows below the
  // diagonal) for the steps l = s (mod D)
#pragma unroll
  for (int s = 0; s < D; s++)
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int i = t + TPB * r;
      fr[s][r] = (s < n && i > s && i < n) ? LU[(size_t)i + (size_t)s * ld] : 0.0;
    }
  for (int l0 = 0; l0 < n; l0 += D) {
#pragma unroll
    for (int s = 0; s < D; s++) {
      const int l = l0 + s;
      if (l < n) {
        const int own = l / TPB;
        if (t == l % TPB) {
#pragma unroll
          for (int r = 0; r < R; r++)
            if (r == own)
#pragma unroll
              for (int c = 0; c < CB; c++) xs[l & 1][c] = x[r][c];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; r++) {
          const int i = t + TPB * r;
          if (i > l && i < n)
#pragma unroll
            for (int c = 0; c < CB; c++) x[r][c] -= fr[s][r] * xs[l & 1][c];
        }
        const int ln = l + D;
#pragma unroll
        for (int r = 0; r < R; r++) {
          const int i = t + TPB * r;
          fr[s][r] = (ln < n && i > ln && i < n) ? LU[(size_t)i + (size_t)ln * ld] : 0.0;
        }
      }
    }
  }
  __syncthreads();
  // backward: upper triangle with its diagonal; fr[s] holds column l (rows
  // on and above the diagonal) for the steps l = n - 1 - s (mod D)
#pragma unroll
  for (int s = 0; s < D; s++) {
    const int l = n - 1 - s;
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int i = t + TPB * r;
      fr[s][r] = (l >= 0 && i <= l) ? LU[(size_t)i + (size_t)l * ld] : 0.0;
    }
  }
  for (int l0 = n - 1; l0 >= 0; l0 -= D) {
#pragma unroll
    for (int s = 0; s < D; s++) {
      const int l = l0 - s;
      if (l >= 0) {
        const int own = l / TPB;
        if (t == l % TPB) {
#pragma unroll
          for (int r = 0; r < R; r++)
            if (r == own)
#pragma unroll
              for (int c = 0; c < CB; c++) {
                x[r][c] = x[r][c] / fr[s][r];
                xs[l & 1][c] = x[r][c];
              }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; r++) {
          const int i = t + TPB * r;
          if (i < l)
#pragma unroll
            for (int c = 0; c < CB; c++) x[r][c] -= fr[s][r] * xs[l & 1][c];
        }
        const int ln = l - D;
#pragma unroll
        for (int r = 0; r < R; r++) {
          const int i = t + TPB * r;
          fr[s][r] = (ln >= 0 && i <= ln) ? LU[(size_t)i + (size_t)ln * ld] : 0.0;
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int i = t + TPB * r;
    if (i < n)
#pragma unroll
      for (int c = 0; c < CB; c++)
        if (c < nc) pB[(size_t)i + (size_t)(c0 + c) * p.ldb] = x[r][c];
  }
}

/* The same solves for systems of at most 64 R rows (R <= 8), one wave per
 * system and CB columns of B, four independent waves per workgroup: lane i
 * keeps rows i + 64 r of its columns in registers, and the value of each
 * step's pivot row is broadcast from its lane with readlane, so a step needs
 * neither LDS nor a barrier.  The steps run in blocks of D (a multiple of D
 * never crosses a 64-row group), which keeps the row group of the pivot row
 * and the slot of the prefetch ring compile-time constants.  The same
 * arithmetic as lu_solve_small_kernel, operation for operation; the ring
 * is shorter (D = 4) for four or eight rows per lane, which keeps two or
 * three waves per SIMD resident. */
__device__ __forceinline__ double wave_bcast(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

/* Workgroups are dispatched to the 8 XCDs round-robin; this renumbering
 * gives each XCD one contiguous range of workgroups, so the units of one
 * system (consecutive numbers) share that XCD's L2 copy of its factor.
 * A bijection of [0, gridDim.x). */
__device__ __forceinline__ int xcd_block() {
  const int n = (int)gridDim.x, i = (int)blockIdx.x;
  const int q = n / 8, r = n % 8, x = i % 8, p = i / 8;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + p;
}

template <int R, int CB, int D>
__global__ void __launch_bounds__(TPB) lu_solve_wave_kernel(const SolveDesc* __restrict__ d,
                                                            int count, int nunits) {
  static_assert(64 % D == 0, "step blocks stay inside a row group");
  const int lane = threadIdx.x & 63;
  const int unit = xcd_block() * (TPB / 64) + (int)(threadIdx.x >> 6);
  if (unit >= nunits) return;  // (whole waves; nothing below synchronizes waves)
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].chunk0 <= unit)
      lo = mid;
    else
      hi = mid - 1;
  }
  const SolveDesc p = d[lo];
  const gptr<double> pB = to_global(p.B);
  const int n = p.n;
  const int c0 = (unit - p.chunk0) * CB;
  const int nc = min(CB, p.ncols - c0);
  // the interchanges piv[0..n), applied in order, as one permutation: lane i
  // holds entries i + 64 r
  int pm[R], pvr[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int i = lane + 64 * r;
    pm[r] = i;
    pvr[r] = i < n ? to_global(p.piv)[i] : i;
  }
#pragma unroll
  for (int g = 0; g < R; g++)
    for (int k = 64 * g; k < min(n, 64 * g + 64); k++) {
      const int q = __builtin_amdgcn_readlane(pvr[g], k - 64 * g);
      if (q != k && (unsigned)q < (unsigned)n) {
        const int a = __builtin_amdgcn_readlane(pm[g], k - 64 * g);
        int b = a;
#pragma unroll
        for (int r = 0; r < R; r++)
          if (r == (q >> 6)) b = __builtin_amdgcn_readlane(pm[r], q & 63);
#pragma unroll
        for (int r = 0; r < R; r++) {
          const int i = lane + 64 * r;
          if (i == k) pm[r] = b;
          if (i == q) pm[r] = a;
        }
      }
    }
  double x[R][CB];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int i = lane + 64 * r;
#pragma unroll
    for (int c = 0; c < CB; c++)
      x[r][c] = (i < n && c < nc) ? pB[(size_t)pm[r] + (size_t)(c0 + c) * p.ldb] : 0.0;
  }
  const gptr<const double> LU = to_global(p.LU);
  const size_t ld = (size_t)p.ld_lu;
  double fr[D][R];
  // forward: unit lower triangle; fr[s] holds column l (rows below the
  // diagonal) for the steps l = s (mod D)
#pragma unroll
  for (int s = 0; s < D; s++)
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int i = lane + 64 * r;
      fr[s][r] = (s < n && i > s && i < n) ? LU[(size_t)i + (size_t)s * ld] : 0.0;
    }
#pragma unroll
  for (int g = 0; g < R; g++)
    for (int l0 = 64 * g; l0 < min(n, 64 * g + 64); l0 += D) {
#pragma unroll
      for (int s = 0; s < D; s++) {
        const int l = l0 + s;
        if (l < n) {
          double xl[CB];
#pragma unroll
          for (int c = 0; c < CB; c++) xl[c] = wave_bcast(x[g][c], l - 64 * g);
#pragma unroll
          for (int r = g; r < R; r++) {  // rows below l lie in groups g and up
            const int i = lane + 64 * r;
            if (i > l && i < n)
#pragma unroll
              for (int c = 0; c < CB; c++) x[r][c] -= fr[s][r] * xl[c];
          }
          const int ln = l + D;
#pragma unroll
          for (int r = 0; r < R; r++) {
            const int i = lane + 64 * r;
            fr[s][r] = (ln < n && i > ln && i < n) ? LU[(size_t)i + (size_t)ln * ld] : 0.0;
          }
        }
      }
    }
  // backward: upper triangle with its diagonal, from the top step block
  // (l0 = D - 1 mod D) down; fr[s] holds column l for the steps with
  // l = D - 1 - s (mod D).  The steps l >= n of the top block only pass
  // their slot on to column l - D.
  const int top = (n + D - 1) / D * D - 1;
#pragma unroll
  for (int s = 0; s < D; s++) {
    const int l = top - s;
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int i = lane + 64 * r;
      fr[s][r] = (l >= 0 && l < n && i <= l) ? LU[(size_t)i + (size_t)l * ld] : 0.0;
    }
  }
#pragma unroll
  for (int g = R - 1; g >= 0; g--)
    for (int l0 = min(top, 64 * g + 63); l0 >= 64 * g; l0 -= D) {
#pragma unroll
      for (int s = 0; s < D; s++) {
        const int l = l0 - s;
        if (l < n) {
          const int own = l - 64 * g;
          const double dg = wave_bcast(fr[s][g], own);
          // the owner's CB values to lanes 0..CB-1, one division for all
          double y = wave_bcast(x[g][0], own);
#pragma unroll
          for (int c = 1; c < CB; c++) {
            const double v = wave_bcast(x[g][c], own);
            if (lane == c) y = v;
          }
          y = y / dg;
          double xl[CB];
#pragma unroll
          for (int c = 0; c < CB; c++) xl[c] = wave_bcast(y, c);
          if (lane == own)
#pragma unroll
            for (int c = 0; c < CB; c++) x[g][c] = xl[c];
#pragma unroll
          for (int r = 0; r <= g; r++) {  // rows above l lie in groups g and below
            const int i = lane + 64 * r;
            if (i < l)
#pragma unroll
              for (int c = 0; c < CB; c++) x[r][c] -= fr[s][r] * xl[c];
          }
        }
        const int ln = l - D;
#pragma unroll
        for (int r = 0; r < R; r++) {
          const int i = lane + 64 * r;
          fr[s][r] = (ln >= 0 && ln < n && i <= ln) ? LU[(size_t)i + (size_t)ln * ld] : 0.0;
        }
      }
    }
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int i = lane + 64 * r;
    if (i < n)
#pragma unroll
      for (int c = 0; c < CB; c++)
        if (c < nc) pB[(size_t)i + (size_t)(c0 + c) * p.ldb] = x[r][c];
  }
}

/* Solves of many right-hand sides (the factorization's leaf steps, hundreds
 * of columns) for systems of at most 512 rows.  One workgroup takes four
 * consecutive column chunks of one system, one per wave, and the waves share
 * the factor: its columns are staged through LDS, LU_SHARED_PIECE at a time,
 * and read from there, so each factor column is loaded once per four chunks
 * instead of once per chunk (per chunk the one-wave kernel streams the whole
 * factor).  Per element the operations and their order are those of
 * lu_solve_wave_kernel (the interchanges as one permutation, the forward
 * steps in increasing order, the backward steps with one division by the
 * diagonal each), so results are bit-identical to it. */
constexpr int LU_SHARED_PIECE = 16;

template <int R, int CB>
__global__ void __launch_bounds__(TPB) lu_solve_shared_kernel(const SolveDesc* __restrict__ d,
                                                              int count) {
  constexpr int PS = LU_SHARED_PIECE;
  constexpr int LDR = 64 * R;
  __shared__ double Ls[PS][LDR];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].chunk0 <= (int)blockIdx.x)
      lo = mid;
    else
      hi = mid - 1;
  }
  const SolveDesc p = d[lo];
  const gptr<double> pB = to_global(p.B);
  const int n = p.n;
  const int c0 = ((int)blockIdx.x - p.chunk0) * (4 * CB) + w * CB;
  const int nc = min(CB, p.ncols - c0);  // (<= 0: the wave only takes part in the barriers)
  int pm[R], pvr[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int i = lane + 64 * r;
    pm[r] = i;
    pvr[r] = i < n ? to_global(p.piv)[i] : i;
  }
#pragma unroll
  for (int g = 0; g < R; g++)
    for (int k = 64 * g; k < min(n, 64 * g + 64); k++) {
      const int q = __builtin_amdgcn_readlane(pvr[g], k - 64 * g);
      if (q != k && (unsigned)q < (unsigned)n) {
        const int a = __builtin_amdgcn_readlane(pm[g], k - 64 * g);
        int b = a;
#pragma unroll
        for (int r = 0; r < R; r++)
          if (r == (q >> 6)) b = __builtin_amdgcn_readlane(pm[r], q & 63);
#pragma unroll
        for (int r = 0; r < R; r++) {
          const int i = lane + 64 * r;
          if (i == k) pm[r] = b;
          if (i == q) pm[r] = a;
        }
      }
    }
  double x[R][CB];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int i = lane + 64 * r;
#pragma unroll
    for (int c = 0; c < CB; c++)
      x[r][c] = (i < n && c < nc) ? pB[(size_t)pm[r] + (size_t)(c0 + c) * p.ldb] : 0.0;
  }
  const gptr<const double> LU = to_global(p.LU);
  const size_t ld = (size_t)p.ld_lu;
  // forward: unit lower triangle, the columns l of row group g in pieces
#pragma unroll
  for (int g = 0; g < R; g++)
    for (int l0 = 64 * g; l0 < min(n, 64 * g + 64); l0 += PS) {
      __syncthreads();
      for (int e = t; e < PS * LDR; e += TPB) {
        const int j = e / LDR, i = e - j * LDR, l = l0 + j;
        Ls[j][i] = (l < n && i > l && i < n) ? LU[(size_t)i + (size_t)l * ld] : 0.0;
      }
      __syncthreads();
      for (int s = 0; s < PS; s++) {
        const int l = l0 + s;
        if (l >= n) break;
        double xl[CB];
#pragma unroll
        for (int c = 0; c < CB; c++) xl[c] = wave_bcast(x[g][c], l - 64 * g);
#pragma unroll
        for (int r = g; r < R; r++) {
          const int i = lane + 64 * r;
          if (i > l && i < n) {
            const double f = Ls[s][i];
#pragma unroll
            for (int c = 0; c < CB; c++) x[r][c] -= f * xl[c];
          }
        }
      }
    }
  // backward: upper triangle with its diagonal, from the last column down
#pragma unroll
  for (int g = R - 1; g >= 0; g--)
    for (int lh = min(n - 1, 64 * g + 63); lh >= 64 * g; lh -= PS) {
      __syncthreads();
      for (int e = t; e < PS * LDR; e += TPB) {
        const int j = e / LDR, i = e - j * LDR, l = lh - j;
        Ls[j][i] = (l >= 64 * g && i <= l) ? LU[(size_t)i + (size_t)l * ld] : 0.0;
      }
      __syncthreads();
      for (int s = 0; s < PS; s++) {
        const int l = lh - s;
        if (l < 64 * g) break;
        const int own = l - 64 * g;
        const double dg = Ls[s][l];
        double xl[CB];
#pragma unroll
        for (int c = 0; c < CB; c++) xl[c] = wave_bcast(x[g][c], own) / dg;
        if (lane == own)
#pragma unroll
          for (int c = 0; c < CB; c++) x[g][c] = xl[c];
#pragma unroll
        for (int r = 0; r <= g; r++) {
          const int i = lane + 64 * r;
          if (i < l) {
            const double f = Ls[s][i];
#pragma unroll
            for (int c = 0; c < CB; c++) x[r][c] -= f * xl[c];
          }
        }
      }
    }
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int i = lane + 64 * r;
    if (i < n)
#pragma unroll
      for (int c = 0; c < CB; c++)
        if (c < nc) pB[(size_t)i + (size_t)(c0 + c) * p.ldb] = x[r][c];
  }
}

/* The same solves for systems of 512-2048 rows, in blocks of 64 rows: the
 * wave that holds a block's rows solves the block's diagonal part alone
 * (staged in LDS, the step's value broadcast by readlane, no barrier), then
 * the whole workgroup subtracts the block's 64 solved values from the rows
 * below it (forward) or above it (backward): two barriers per block instead
 * of one per row.  Every row still receives its updates one at a time, in
 * the order of the step-by-step substitution, so results are bit-identical
 * to lu_solve_small_kernel. */
template <int R>
__global__ void __launch_bounds__(TPB) lu_solve_blk_kernel(const SolveDesc* __restrict__ d,
                                                           int count) {
  constexpr int CB = solve_cols(R);
  __shared__ int perm[R * TPB];
  __shared__ int pv[R * TPB];
  __shared__ double blk[64][65];  // diagonal block: blk[i][j] = A(b0 + i, b0 + j)
  __shared__ double xs[64][CB];   // the block's solved values
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].chunk0 <= (int)blockIdx.x)
      lo = mid;
    else
      hi = mid - 1;
  }
  const SolveDesc p = d[lo];
  const gptr<double> pB = to_global(p.B);
  const int n = p.n, t = threadIdx.x;
  const int c0 = ((int)blockIdx.x - p.chunk0) * CB;
  const int nc = min(CB, p.ncols - c0);
  // the interchanges piv[0..n), applied in order, as one permutation (read
  // by all threads, replayed from LDS by one)
  for (int i = t; i < n; i += TPB) {
    perm[i] = i;
    pv[i] = to_global(p.piv)[i];
  }
  __syncthreads();
  if (t == 0)
    for (int k = 0; k < n; k++) {
      const int q = pv[k];
      if (q != k) {
        const int x = perm[k];
        perm[k] = perm[q];
        perm[q] = x;
      }
    }
  __syncthreads();
  double x[R][CB];
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int i = t + TPB * r;
#pragma unroll
    for (int c = 0; c < CB; c++)
      x[r][c] = (i < n && c < nc) ? pB[(size_t)perm[i] + (size_t)(c0 + c) * p.ldb] : 0.0;
  }
  const gptr<const double> LU = to_global(p.LU);
  const size_t ld = (size_t)p.ld_lu;
  const int lane = t & 63, wave = t >> 6;
  const int nblk = (n + 63) / 64;
  auto stage = [&](int b0, int bw) {
    for (int e = t; e < 64 * 64; e += TPB) {
      const int i = e & 63, j = e >> 6;
      blk[i][j] = (i < bw && j < bw) ? LU[(size_t)(b0 + i) + (size_t)(b0 + j) * ld] : 0.0;
    }
  };
  // forward: unit lower triangle
  for (int bi = 0; bi < nblk; bi++) {
    const int b0 = bi * 64, bw = min(64, n - b0);
    stage(b0, bw);
    __syncthreads();
    if (wave == (bi & 3)) {  // rows b0 + lane, register slot b0 / 256
#pragma unroll
      for (int q = 0; q < R; q++)
        if (q == (b0 >> 8)) {
          for (int l = 0; l < bw; l++) {
            double xl[CB];
#pragma unroll
            for (int c = 0; c < CB; c++) xl[c] = wave_bcast(x[q][c], l);
            if (lane > l && lane < bw) {
              const double a = blk[lane][l];
#pragma unroll
              for (int c = 0; c < CB; c++) x[q][c] -= a * xl[c];
            }
          }
          if (lane < bw)
#pragma unroll
            for (int c = 0; c < CB; c++) xs[lane][c] = x[q][c];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < R; q++) {
      const int i = t + TPB * q;
      if (i >= b0 + bw && i < n) {
        const gptr<const double> col = LU + (size_t)i + (size_t)b0 * ld;
#pragma unroll 8
        for (int j = 0; j < bw; j++) {
          const double a = col[(size_t)j * ld];
#pragma unroll
          for (int c = 0; c < CB; c++) x[q][c] -= a * xs[j][c];
        }
      }
    }
    __syncthreads();  // blk and xs are rewritten for the next block
  }
  // backward: upper triangle with its diagonal
  for (int bi = nblk - 1; bi >= 0; bi--) {
    const int b0 = bi * 64, bw = min(64, n - b0);
    stage(b0, bw);
    __syncthreads();
    if (wave == (bi & 3)) {
#pragma unroll
      for (int q = 0; q < R; q++)
        if (q == (b0 >> 8)) {
          for (int l = bw - 1; l >= 0; l--) {
            if (lane == l) {
              const double dg = blk[l][l];
#pragma unroll
              for (int c = 0; c < CB; c++) x[q][c] = x[q][c] / dg;
            }
            double xl[CB];
#pragma unroll
            for (int c = 0; c < CB; c++) xl[c] = wave_bcast(x[q][c], l);
            if (lane < l) {
              const double a = blk[lane][l];
#pragma unroll
              for (int c = 0; c < CB; c++) x[q][c] -= a * xl[c];
            }
          }
          if (lane < bw)
#pragma unroll
            for (int c = 0; c < CB; c++) xs[lane][c] = x[q][c];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < R; q++) {
      const int i = t + TPB * q;
      if (i < b0) {
        const gptr<const double> col = LU + (size_t)i + (size_t)b0 * ld;
#pragma unroll 8
        for (int j = bw - 1; j >= 0; j--) {
          const double a = col[(size_t)j * ld];
#pragma unroll
          for (int c = 0; c < CB; c++) x[q][c] -= a * xs[j][c];
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int i = t + TPB * r;
    if (i < n)
#pragma unroll
      for (int c = 0; c < CB; c++)
        if (c < nc) pB[(size_t)i + (size_t)(c0 + c) * p.ldb] = x[r][c];
  }
}

constexpr int wave_rows(int maxn) { return maxn <= 64 ? 1 : maxn <= 128 ? 2 : maxn <= 256 ? 4 : 8; }
// columns per wave: fewer for more rows per lane (registers: R x CB values
// and the D x R ring; beyond 8 rows per lane the compiler no longer keeps
// the unrolled arrays in registers)
constexpr int wave_cols(int R) { return R <= 4 ? 8 : 4; }

constexpr int MAX_GRID_Y = 65535;

/* LU solves on the matrix cores (the factorization's leaf steps: systems of
 * at most 16 RT rows, hundreds of columns).  A workgroup of four waves takes
 * one system and 64 columns of B, 16 per wave.  A wave keeps X = P B as RT
 * tiles of 16 x 16 in the result layout of v_mfma_f64_16x16x4f64 (lane l,
 * register i: row (l >> 4) + 4 i of the tile, column l & 15), which is also
 * that instruction's B-operand layout for the tile's rows 4 i .. 4 i + 3: a
 * solved tile multiplies the factor's off-diagonal blocks as it stands.
 * Forward (unit L) and backward (U) by tiles: the diagonal tile by
 * substitution, row j's values sent to the lanes of their column by
 * ds_bpermute; then every tile below (above) gets X_u -= F_ut X_t as four
 * matrix-core products.  The factor's block column of step t (its diagonal
 * tile and the tiles below it, or above it going backward) is staged in LDS
 * by the whole workgroup in the A-operand layout (lane l: row l & 15, column
 * l >> 4 of a 16 x 4 slice), and the next step's block column is loaded into
 * registers while the current one is used: each factor element is read from
 * memory once per 64 columns, off the substitution's critical path.  The row
 * interchanges come as one permutation (SolveDesc::perm: row i of P B is row
 * perm[i] of B).  Consecutive workgroups (the column chunks of one system)
 * run on one XCD, whose L2 then holds the system's factor. */
typedef double mf4d __attribute__((ext_vector_type(4)));

template <int RT>
__global__ void __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(2, 2)))
    lu_solve_mfma_kernel(const SolveDesc* __restrict__ d, int count) {
  __shared__ double Pa[RT][4][64];        // block column: tile u, slice kk, A-operand lane
  __shared__ double Ds[TPB / 64][16][17];  // each wave's transpose buffer for B
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int blk = xcd_block();
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].chunk0 <= blk)
      lo = mid;
    else
      hi = mid - 1;
  }
  const SolveDesc p = d[lo];
  const int n = p.n, nt = (n + 15) >> 4;
  const int q = lane >> 4, c = lane & 15;
  const size_t ld = (size_t)p.ld_lu;
  const gptr<const double> F = to_global(p.LU);
  const gptr<double> pB = to_global(p.B);
  double(*D)[17] = Ds[w];
  // this thread's share of a block column: slice t >> 6, lane t & 63 of
  // every tile: factor row 16 (t0 + u) + (t & 15), column 16 T + 4 (t >> 6)
  // + ((t >> 4) & 3); rows and columns beyond n read as the identity
  const int pr = t & 15, pc = 4 * (t >> 6) + ((t >> 4) & 3);
  double pf[RT];
  auto load_col = [&](int T, int t0, int cnt) {
    const int gc = 16 * T + pc;
#pragma unroll
    for (int u = 0; u < RT; u++) {
      const int gr = 16 * (t0 + u) + pr;
      pf[u] = (u < cnt) ? ((gr < n && gc < n) ? F[(size_t)gr + (size_t)gc * ld] : (gr == gc ? 1.0 : 0.0))
                        : 0.0;
    }
  };
  auto store_col = [&](int cnt, bool reverse) {
#pragma unroll
    for (int u = 0; u < RT; u++)
      if (u < cnt) Pa[reverse ? cnt - 1 - u : u][t >> 6][t & 63] = pf[u];
  };
  // P B tile by tile through the wave's LDS block: lane l reads row l & 15 of
  // the tile in columns 4 (l >> 4) + i (16 consecutive rows of a column per
  // quarter wave), and takes back its result-layout values
  const int c0 = (blk - p.chunk0) * 64 + 16 * w;
  const int lr = lane & 15, lc = (lane >> 4) * 4;
  mf4d x[RT];
#pragma unroll
  for (int tt = 0; tt < RT; tt++) {
    const int r = 16 * tt + lr;
    const int src = r < n ? to_global(p.perm)[r] : 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int cc = c0 + lc + i;
      D[lr][lc + i] = (r < n && cc < p.ncols) ? pB[(size_t)src + (size_t)cc * p.ldb] : 0.0;
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int i = 0; i < 4; i++) x[tt][i] = D[q + 4 * i][c];
    __builtin_amdgcn_wave_barrier();
  }
  // Forward, tile T = 0 .. RT-1: the tile being solved is always x[0] and
  // x[u] holds tile T + u; after each step the tiles rotate down by one (the
  // solved tile to x[RT-1]), so after RT steps every tile is back in place.
  // (Fixed register positions keep the loop body small: fully unrolled over
  // the tiles the code outgrows the instruction cache.)  Pa[u] holds the
  // factor tile (T + u, T).
  load_col(0, 0, nt);
  for (int T = 0; T < RT; T++) {
    if (T < nt) {
      __syncthreads();  // every wave is done with the previous block column
      store_col(nt - T, false);
      __syncthreads();
      if (T + 1 < nt) load_col(T + 1, T + 1, nt - T - 1);
#pragma unroll
      for (int j = 0; j < 15; j++) {
        const double xj = __shfl(x[0][j >> 2], ((j & 3) << 4) | c, 64);
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int r = q + 4 * i;
          if (r > j) x[0][i] -= Pa[0][j >> 2][((j & 3) << 4) | r] * xj;
        }
      }
#pragma unroll
      for (int u = 1; u < RT; u++) {
        if (T + u < nt) {
          double a[4];
#pragma unroll
          for (int kk = 0; kk < 4; kk++) a[kk] = -Pa[u][kk][lane];
#pragma unroll
          for (int kk = 0; kk < 4; kk++)
            x[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kk], x[0][kk], x[u], 0, 0, 0);
        }
      }
    }
    const mf4d s0 = x[0];
#pragma unroll
    for (int u = 0; u < RT - 1; u++) x[u] = x[u + 1];
    x[RT - 1] = s0;
  }
  // Backward, tile T = RT-1 .. 0: the tile being solved is always x[RT-1]
  // and x[RT-1-v] holds tile T - v; the tiles rotate up by one per step.
  // Pa[v] holds the factor tile (T - v, T).
  if (nt > 0) load_col(nt - 1, 0, nt);
#pragma unroll 1
  for (int T = RT - 1; T >= 0; T--) {
    if (T < nt) {
      __syncthreads();
      store_col(T + 1, true);
      __syncthreads();
      if (T > 0) load_col(T - 1, 0, T);
#pragma unroll
      for (int j = 15; j >= 0; j--) {
        const double xj = __shfl(x[RT - 1][j >> 2], ((j & 3) << 4) | c, 64) /
                          Pa[0][j >> 2][((j & 3) << 4) | j];
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int r = q + 4 * i;
          if (r == j)
            x[RT - 1][i] = xj;
          else if (r < j)
            x[RT - 1][i] -= Pa[0][j >> 2][((j & 3) << 4) | r] * xj;
        }
      }
#pragma unroll
      for (int v = 1; v < RT; v++) {
        if (v <= T) {
          double a[4];
#pragma unroll
          for (int kk = 0; kk < 4; kk++) a[kk] = -Pa[v][kk][lane];
#pragma unroll
          for (int kk = 0; kk < 4; kk++)
            x[RT - 1 - v] =
                __builtin_amdgcn_mfma_f64_16x16x4f64(a[kk], x[RT - 1][kk], x[RT - 1 - v], 0, 0, 0);
        }
      }
    }
    const mf4d s1 = x[RT - 1];
#pragma unroll
    for (int u = RT - 1; u > 0; u--) x[u] = x[u - 1];
    x[0] = s1;
  }
  // back through LDS: quarter waves write 16 consecutive rows of a column
#pragma unroll
  for (int tt = 0; tt < RT; tt++) {
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int i = 0; i < 4; i++) D[q + 4 * i][c] = x[tt][i];
    __builtin_amdgcn_wave_barrier();
    const int r = 16 * tt + lr;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int cc = c0 + lc + i;
      if (r < n && cc < p.ncols) pB[(size_t)r + (size_t)cc * p.ldb] = D[lr][lc + i];
    }
  }
}

/* The interchanges piv[0..n) of each matrix, applied in order to the
 * identity, as one permutation (perm[i]: the row that ends in row i); one
 * wave per matrix, the swaps done in order by lane 0 in LDS. */
__global__ void __launch_bounds__(64) lu_perm_kernel(const PermBuild* __restrict__ d) {
  __shared__ int pm[LU_PERM_MAX];
  const PermBuild b = d[blockIdx.x];
  const int lane = threadIdx.x;
  for (int i = lane; i < b.n; i += 64) pm[i] = i;
  __syncthreads();
  if (lane == 0)
    for (int k = 0; k < b.n; k++) {
      const int q = b.piv[k];
      if (q != k && (unsigned)q < (unsigned)b.n) {
        const int t = pm[k];
        pm[k] = pm[q];
        pm[q] = t;
      }
    }
  __syncthreads();
  for (int i = lane; i < b.n; i += 64) b.perm[i] = pm[i];
}

}  // namespace

void launch_assemble(const AssembleDesc* d, int count, int max_elems, const EntrySource* srcs,
                     hipStream_t s) {
  if (count == 0) return;
  for (int c0 = 0; c0 < count; c0 += MAX_GRID_Y) {
    const int c = std::min(MAX_GRID_Y, count - c0);
    dim3 g(grid_x((size_t)max_elems, count), c);
    assemble_kernel<<<g, TPB, 0, s>>>(d + c0, srcs);
    HODLR_HIP_CHECK(hipGetLastError());
  }
}

int lu_panel_width(int maxn, int nb) {
  if (maxn <= 1024) return std::min(nb, 32);
  if (maxn <= 2048) return std::min(nb, 16);
  if (maxn <= 4096) return std::min(nb, 8);
  return nb;
}

void launch_lu_panel(const LuDesc* d, int count, int p0, int nb, int maxn, hipStream_t s) {
  if (count == 0) return;
  if (maxn <= 256 && nb <= 32)
    lu_panel_reg_kernel<256, 1, 32><<<count, 256, 0, s>>>(d, p0, nb);
  else if (maxn <= 512 && nb <= 32)
    lu_panel_reg_kernel<512, 1, 32><<<count, 512, 0, s>>>(d, p0, nb);
  else if (maxn <= 1024 && nb <= 32)
    lu_panel_reg_kernel<512, 2, 32><<<count, 512, 0, s>>>(d, p0, nb);
  else if (maxn <= 2048 && nb <= 16)
    lu_panel_reg_kernel<512, 4, 16><<<count, 512, 0, s>>>(d, p0, nb);
  else if (maxn <= 4096 && nb <= 8)
    lu_panel_reg_kernel<512, 8, 8><<<count, 512, 0, s>>>(d, p0, nb);
  else
    lu_panel_kernel<<<count, TPB, 0, s>>>(d, p0, nb);
  HODLR_HIP_CHECK(hipGetLastError());
}

bool lu_panel_defers_swaps(int maxn, int nb) {
  return (maxn <= 1024 && nb <= 32) || (maxn <= 2048 && nb <= 16) || (maxn <= 4096 && nb <= 8);
}

void launch_lu_step(const LuStepDesc* d, int count, int nwg, bool swap, int nb, hipStream_t s) {
  if (count == 0 || nwg == 0) return;
  if (nb <= 32)
    lu_step_kernel<32><<<nwg, LU_STEP_NT, 0, s>>>(d, count, swap ? 1 : 0);
  else
    lu_step_kernel<64><<<nwg, LU_STEP_NT, 0, s>>>(d, count, 0);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_trsm(const TrsmDesc* d, int count, int nwg, hipStream_t s) {
  if (count == 0 || nwg == 0) return;
  trsm_kernel<<<nwg, 64, 0, s>>>(d, count);
  HODLR_HIP_CHECK(hipGetLastError());
}

int lu_solve_mfma_wgs(int ncols) { return (ncols + 63) / 64; }

bool launch_lu_solve_mfma(const SolveDesc* d, int count, int nwg, int maxn, hipStream_t s) {
  if (maxn > LU_SOLVE_MFMA_MAX) return false;
  if (count == 0 || nwg == 0) return true;
  if (maxn <= 64)
    lu_solve_mfma_kernel<4><<<nwg, TPB, 0, s>>>(d, count);
  else if (maxn <= 128)
    lu_solve_mfma_kernel<8><<<nwg, TPB, 0, s>>>(d, count);
  else
    lu_solve_mfma_kernel<16><<<nwg, TPB, 0, s>>>(d, count);
  HODLR_HIP_CHECK(hipGetLastError());
  return true;
}

void launch_lu_perm(const PermBuild* d, int count, hipStream_t s) {
  if (count == 0) return;
  lu_perm_kernel<<<count, 64, 0, s>>>(d);
  HODLR_HIP_CHECK(hipGetLastError());
}

int lu_solve_shared_wgs(int ncols, int maxn) {
  const int cb = 4 * wave_cols(wave_rows(maxn));
  return (ncols + cb - 1) / cb;
}

bool launch_lu_solve_shared(const SolveDesc* d, int count, int nwg, int maxn, hipStream_t s) {
  if (maxn > LU_SOLVE_WAVE_MAX) return false;
  if (count == 0 || nwg == 0) return true;
  switch (wave_rows(maxn)) {
    case 1: lu_solve_shared_kernel<1, 8><<<nwg, TPB, 0, s>>>(d, count); break;
    case 2: lu_solve_shared_kernel<2, 8><<<nwg, TPB, 0, s>>>(d, count); break;
    case 4: lu_solve_shared_kernel<4, 8><<<nwg, TPB, 0, s>>>(d, count); break;
    default: lu_solve_shared_kernel<8, 4><<<nwg, TPB, 0, s>>>(d, count); break;
  }
  HODLR_HIP_CHECK(hipGetLastError());
  return true;
}

int lu_solve_small_wgs(int ncols, int maxn, bool wave) {
  const int cb =
      wave && maxn <= LU_SOLVE_WAVE_MAX ? wave_cols(wave_rows(maxn)) : solve_cols(solve_rows(maxn));
  return (ncols + cb - 1) / cb;
}

int lu_solve_wide_units(int ncols) { return (ncols + 15) / 16; }

bool launch_lu_solve_wide(const SolveDesc* d, int count, int nunits, int maxn, hipStream_t s) {
  if (maxn > LU_SOLVE_WIDE_MAX) return false;
  if (count == 0 || nunits == 0) return true;
  const int grid = (nunits + TPB / 64 - 1) / (TPB / 64);
  switch (wave_rows(maxn)) {
    case 1: lu_solve_wave_kernel<1, 16, 8><<<grid, TPB, 0, s>>>(d, count, nunits); break;
    case 2: lu_solve_wave_kernel<2, 16, 8><<<grid, TPB, 0, s>>>(d, count, nunits); break;
    default: lu_solve_wave_kernel<4, 16, 4><<<grid, TPB, 0, s>>>(d, count, nunits); break;
  }
  HODLR_HIP_CHECK(hipGetLastError());
  return true;
}

bool launch_lu_solve_small(const SolveDesc* d, int count, int nwg, int maxn, bool wave,
                           bool blocked, hipStream_t s) {
  if (maxn > LU_SOLVE_SMALL_MAX) return false;
  if (count == 0 || nwg == 0) return true;
  if (wave && maxn <= LU_SOLVE_WAVE_MAX) {
    const int grid = (nwg + TPB / 64 - 1) / (TPB / 64);
    switch (wave_rows(maxn)) {
      case 1: lu_solve_wave_kernel<1, 8, 8><<<grid, TPB, 0, s>>>(d, count, nwg); break;
      case 2: lu_solve_wave_kernel<2, 8, 8><<<grid, TPB, 0, s>>>(d, count, nwg); break;
      case 4: lu_solve_wave_kernel<4, 8, 4><<<grid, TPB, 0, s>>>(d, count, nwg); break;
      default: lu_solve_wave_kernel<8, 4, 4><<<grid, TPB, 0, s>>>(d, count, nwg); break;
    }
    HODLR_HIP_CHECK(hipGetLastError());
    return true;
  }
  if (blocked && maxn > 512) {
    if (solve_rows(maxn) == 4)
      lu_solve_blk_kernel<4><<<nwg, TPB, 0, s>>>(d, count);
    else
      lu_solve_blk_kernel<8><<<nwg, TPB, 0, s>>>(d, count);
    HODLR_HIP_CHECK(hipGetLastError());
    return true;
  }
  switch (solve_rows(maxn)) {
    case 1: lu_solve_small_kernel<1><<<nwg, TPB, 0, s>>>(d, count); break;
    case 2: lu_solve_small_kernel<2><<<nwg, TPB, 0, s>>>(d, count); break;
    case 4: lu_solve_small_kernel<4><<<nwg, TPB, 0, s>>>(d, count); break;
    default: lu_solve_small_kernel<8><<<nwg, TPB, 0, s>>>(d, count); break;
  }
  HODLR_HIP_CHECK(hipGetLastError());
  return true;
}

void launch_permute(const PermDesc* d, int count, int nwg, hipStream_t s) {
  if (count == 0 || nwg == 0) return;
  permute_kernel<<<nwg, TPB, 0, s>>>(d, count);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_set_identity(const CopyDesc* d, int count, hipStream_t s) {
  if (count == 0) return;
  for (int c0 = 0; c0 < count; c0 += MAX_GRID_Y) {
    dim3 g(16, std::min(MAX_GRID_Y, count - c0));
    identity_kernel<<<g, TPB, 0, s>>>(d + c0);
    HODLR_HIP_CHECK(hipGetLastError());
  }
}

void launch_lr_fill(const LrFillDesc* d, int count, size_t max_elems, hipStream_t s) {
  if (count == 0) return;
  for (int c0 = 0; c0 < count; c0 += MAX_GRID_Y) {
    const int c = std::min(MAX_GRID_Y, count - c0);
    dim3 g(grid_x(max_elems, count), c);
    lr_fill_kernel<<<g, TPB, 0, s>>>(d + c0);
    HODLR_HIP_CHECK(hipGetLastError());
  }
}

void launch_copy(const CopyDesc* d, int count, hipStream_t s) {
  if (count == 0) return;
  for (int c0 = 0; c0 < count; c0 += MAX_GRID_Y) {
    dim3 g(16, std::min(MAX_GRID_Y, count - c0));
    copy_kernel<<<g, TPB, 0, s>>>(d + c0);
    HODLR_HIP_CHECK(hipGetLastError());
  }
}

void launch_logdet(const LuDesc* d, int count, double* out_log, double* out_prod, hipStream_t s) {
  if (count == 0) return;
  logdet_kernel<<<count, TPB, 0, s>>>(d, out_log, out_prod);
  HODLR_HIP_CHECK(hipGetLastError());
}

template <class TZ>
void launch_capacitance(const CapDescT<TZ>* d, int count, hipStream_t s) {
  if (count == 0) return;
  for (int c0 = 0; c0 < count; c0 += MAX_GRID_Y) {
    dim3 g(16, std::min(MAX_GRID_Y, count - c0));
    capacitance_kernel<TZ><<<g, TPB, 0, s>>>(d + c0);
    HODLR_HIP_CHECK(hipGetLastError());
  }
}
template void launch_capacitance<double>(const CapDesc*, int, hipStream_t);
template void launch_capacitance<float>(const CapDescF*, int, hipStream_t);

template <class S, class D>
void launch_convert(const ConvDesc<S, D>* d, int count, hipStream_t s) {
  if (count == 0) return;
  for (int c0 = 0; c0 < count; c0 += MAX_GRID_Y) {
    dim3 g(count > 64 ? 16 : 512, std::min(MAX_GRID_Y, count - c0));
    convert_kernel<S, D><<<g, TPB, 0, s>>>(d + c0);
    HODLR_HIP_CHECK(hipGetLastError());
  }
}
template void launch_convert<double, float>(const ConvDesc<double, float>*, int, hipStream_t);
template void launch_convert<float, double>(const ConvDesc<float, double>*, int, hipStream_t);
template void launch_convert<float, float>(const ConvDesc<float, float>*, int, hipStream_t);

void launch_balance(const BalanceDesc* d, int count, int kmax, hipStream_t s) {
  if (count == 0 || kmax == 0) return;
  for (int c0 = 0; c0 < count; c0 += MAX_GRID_Y) {
    dim3 g(kmax, std::min(MAX_GRID_Y, count - c0));
    balance_kernel<<<g, TPB, 0, s>>>(d + c0);
    HODLR_HIP_CHECK(hipGetLastError());
  }
}

template <class T>
void launch_nonfinite(const T* x, size_t n, unsigned long long* out, hipStream_t s) {
  nonfinite_kernel<T><<<1024, TPB, 0, s>>>(x, n, out);
  HODLR_HIP_CHECK(hipGetLastError());
}
template void launch_nonfinite<float>(const float*, size_t, unsigned long long*, hipStream_t);
template void launch_nonfinite<double>(const double*, size_t, unsigned long long*, hipStream_t);

void launch_segnorm2(const double* X, int ld, const int* row0, int nseg, int ncols, int nsplit,
                     double* out, hipStream_t s) {
  if (ncols <= 0 || nseg <= 0) return;
  if (nsplit < 1 || nsplit > 64) throw std::invalid_argument("segnorm2: nsplit out of range");
  dim3 g(ncols, nseg, nsplit);
  segnorm2_kernel<<<g, TPB, 0, s>>>(X, ld, row0, nseg, out);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_segscale(double* X, int ld, const int* row0, int nseg, int ncols, const double* sc,
                     hipStream_t s) {
  if (ncols <= 0 || nseg <= 0) return;
  dim3 g(ncols, nseg);
  segscale_kernel<<<g, TPB, 0, s>>>(X, ld, row0, nseg, sc);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_segcopy(const double* X, double* Y, int ld, const int* row0, int nseg, int ncols,
                    int nsplit, const int* mask, hipStream_t s) {
  if (ncols <= 0 || nseg <= 0) return;
  if (nsplit < 1 || nsplit > 64) throw std::invalid_argument("segcopy: nsplit out of range");
  dim3 g(ncols, nseg, nsplit);
  segcopy_kernel<<<g, TPB, 0, s>>>(X, Y, ld, row0, nseg, mask);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_colnorm2(const double* X, int n, int ld, int ncols, double* out, hipStream_t s) {
  if (ncols <= 0) return;
  colnorm2_kernel<<<ncols, TPB, 0, s>>>(X, n, ld, out);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_add_f32(double* X, int ldx, const float* D, int ldd, int n, int ncols, hipStream_t s) {
  add_f32_kernel<<<2048, TPB, 0, s>>>(X, ldx, D, ldd, n, ncols);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_multi_dot(const double* V, size_t vstride, int nvec, const double* w, int n, int ld,
                      int ncols, double* out, hipStream_t s) {
  if (nvec <= 0 || ncols <= 0) return;
  dim3 g(ncols, nvec);
  multi_dot_kernel<<<g, TPB, 0, s>>>(V, vstride, nvec, w, n, ld, out);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_multi_axpy(const double* V, size_t vstride, int nvec, const double* coef, double* w,
                       int n, int ld, int ncols, double sign, int assign, hipStream_t s) {
  multi_axpy_kernel<<<2048, TPB, 0, s>>>(V, vstride, nvec, coef, w, n, ld, ncols, sign, assign);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_step_update(double* x, double* d, int ld, int n, int ncols, hipStream_t s) {
  if ((size_t)n * ncols == 0) return;
  step_update_kernel<<<2048, TPB, 0, s>>>(x, d, ld, n, ncols);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_col_scale(double* X, int ld, int n, int ncols, const double* sc, hipStream_t s) {
  col_scale_kernel<<<2048, TPB, 0, s>>>(X, ld, n, ncols, sc);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_sub(const double* B, int ldb, const double* Y, int ldy, double* R, int ldr, int n,
                int ncols, hipStream_t s) {
  sub_kernel<<<2048, TPB, 0, s>>>(B, ldb, Y, ldy, R, ldr, n, ncols);
  HODLR_HIP_CHECK(hipGetLastError());
}

}  // namespace hodlr
