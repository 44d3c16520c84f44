/*
 * Batched dense kernels: one launch covers every matrix of a tree level (or
 * every leaf), one workgroup per matrix or per right-hand-side chunk.
 */
#include "dense.h"

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
    a.out[(size_t)i + (size_t)j * a.ldo] = entry_at(src, a.i0 + i, a.j0 + j);
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
__global__ void __launch_bounds__(TPB) capacitance_kernel(const CapDesc* __restrict__ d) {
  const CapDesc c = d[blockIdx.y];
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
      v = c.lu_order ? c.Zt[(size_t)i + (size_t)(c.own0 + jj) * c.kb]
                     : c.Zb[(size_t)i + (size_t)(c.own0 + jj) * c.kt];
    } else {  // lower-left block
      const int ii = i - ka;
      v = c.lu_order ? c.Zb[(size_t)ii + (size_t)(c.own0 + j) * c.kt]
                     : c.Zt[(size_t)ii + (size_t)(c.own0 + j) * c.kb];
    }
    c.C[e] = v;
  }
  const size_t totY = (size_t)K * c.nanc;
  for (size_t e = (size_t)blockIdx.x * TPB + threadIdx.x; e < totY; e += (size_t)gridDim.x * TPB) {
    const int i = (int)(e % K), j = (int)(e / K);
    double v;
    if (i < ka)
      v = c.lu_order ? c.Zt[(size_t)i + (size_t)j * c.kb] : c.Zb[(size_t)i + (size_t)j * c.kt];
    else
      v = c.lu_order ? c.Zb[(size_t)(i - ka) + (size_t)j * c.kt]
                     : c.Zt[(size_t)(i - ka) + (size_t)j * c.kb];
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

int grid_x(size_t elems, int ndesc) {
  size_t g = (elems + TPB - 1) / TPB;
  size_t cap = ndesc > 64 ? 16 : 1024;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return (int)g;
}

}  // namespace

void launch_assemble(const AssembleDesc* d, int count, int max_elems, const EntrySource* srcs,
                     hipStream_t s) {
  if (count == 0) return;
  dim3 g(grid_x((size_t)max_elems, count), count);
  assemble_kernel<<<g, TPB, 0, s>>>(d, srcs);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_lu_panel(const LuDesc* d, int count, int p0, int nb, hipStream_t s) {
  if (count == 0) return;
  lu_panel_kernel<<<count, TPB, 0, s>>>(d, p0, nb);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_trsm(const TrsmDesc* d, int count, int nwg, hipStream_t s) {
  if (count == 0 || nwg == 0) return;
  trsm_kernel<<<nwg, 64, 0, s>>>(d, count);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_permute(const PermDesc* d, int count, int nwg, hipStream_t s) {
  if (count == 0 || nwg == 0) return;
  permute_kernel<<<nwg, TPB, 0, s>>>(d, count);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_set_identity(const CopyDesc* d, int count, hipStream_t s) {
  if (count == 0) return;
  dim3 g(16, count);
  identity_kernel<<<g, TPB, 0, s>>>(d);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_copy(const CopyDesc* d, int count, hipStream_t s) {
  if (count == 0) return;
  dim3 g(16, count);
  copy_kernel<<<g, TPB, 0, s>>>(d);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_logdet(const LuDesc* d, int count, double* out_log, double* out_prod, hipStream_t s) {
  if (count == 0) return;
  logdet_kernel<<<count, TPB, 0, s>>>(d, out_log, out_prod);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_capacitance(const CapDesc* d, int count, hipStream_t s) {
  if (count == 0) return;
  dim3 g(16, count);
  capacitance_kernel<<<g, TPB, 0, s>>>(d);
  HODLR_HIP_CHECK(hipGetLastError());
}

}  // namespace hodlr
