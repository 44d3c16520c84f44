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

/* Right-looking partial-pivot LU, one workgroup per matrix.  The pivot is
 * the first row of maximum |a_ik| (LAPACK getrf / Eigen PartialPivLU). */
__global__ void __launch_bounds__(TPB) lu_kernel(const LuDesc* __restrict__ d) {
  __shared__ double sv[TPB / 64];
  __shared__ int si[TPB / 64];
  __shared__ int piv_sh;
  const LuDesc m = d[blockIdx.x];
  double* A = m.A;
  const int n = m.n, ld = m.ld, t = threadIdx.x;
  for (int k = 0; k < n; k++) {
    double bv = -1.0;
    int bi = n;
    for (int i = k + t; i < n; i += TPB) {
      const double v = fabs(A[(size_t)i + (size_t)k * ld]);
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
      m.piv[k] = ix;
    }
    __syncthreads();
    const int p = piv_sh;
    if (p != k)
      for (int j = t; j < n; j += TPB) {
        const double x = A[(size_t)k + (size_t)j * ld];
        A[(size_t)k + (size_t)j * ld] = A[(size_t)p + (size_t)j * ld];
        A[(size_t)p + (size_t)j * ld] = x;
      }
    __syncthreads();
    const double pv = A[(size_t)k + (size_t)k * ld];
    if (pv != 0.0) {
      const double inv = 1.0 / pv;
      for (int i = k + 1 + t; i < n; i += TPB) A[(size_t)i + (size_t)k * ld] *= inv;
    }
    __syncthreads();
    const int r = n - k - 1;
    const size_t tot = (size_t)r * r;
    for (size_t e = t; e < tot; e += TPB) {
      const int i = k + 1 + (int)(e % r), j = k + 1 + (int)(e / r);
      A[(size_t)i + (size_t)j * ld] -= A[(size_t)i + (size_t)k * ld] * A[(size_t)k + (size_t)j * ld];
    }
    __syncthreads();
  }
}

/* In-place LU solve of a chunk of ch right-hand-side columns, staged in LDS.
 * The workgroup -> (problem, chunk) mapping is a binary search on chunk0. */
__global__ void __launch_bounds__(TPB)
    lu_solve_kernel(const SolveDesc* __restrict__ d, int count, int ch) {
  extern __shared__ double x[];
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].chunk0 <= (int)blockIdx.x)
      lo = mid;
    else
      hi = mid - 1;
  }
  const SolveDesc p = d[lo];
  const int c0 = ((int)blockIdx.x - p.chunk0) * ch;
  const int nc = min(ch, p.ncols - c0);
  if (nc <= 0) return;
  const int n = p.n, t = threadIdx.x;
  const double* LU = p.LU;
  const int ldl = p.ld_lu;
  for (int e = t; e < n * nc; e += TPB) {
    const int i = e % n, c = e / n;
    x[i + c * n] = p.B[(size_t)i + (size_t)(c0 + c) * p.ldb];
  }
  __syncthreads();
  if (t < nc)
    for (int k = 0; k < n; k++) {
      const int q = p.piv[k];
      if (q != k) {
        const double tmp = x[k + t * n];
        x[k + t * n] = x[q + t * n];
        x[q + t * n] = tmp;
      }
    }
  __syncthreads();
  for (int k = 0; k < n - 1; k++) {  // unit lower triangle
    const int r = n - k - 1;
    for (int e = t; e < r * nc; e += TPB) {
      const int i = k + 1 + e % r, c = e / r;
      x[i + c * n] -= LU[(size_t)i + (size_t)k * ldl] * x[k + c * n];
    }
    __syncthreads();
  }
  for (int k = n - 1; k >= 0; k--) {  // upper triangle
    if (t < nc) x[k + t * n] /= LU[(size_t)k + (size_t)k * ldl];
    __syncthreads();
    for (int e = t; e < k * nc; e += TPB) {
      const int i = e % k, c = e / k;
      x[i + c * n] -= LU[(size_t)i + (size_t)k * ldl] * x[k + c * n];
    }
    __syncthreads();
  }
  for (int e = t; e < n * nc; e += TPB) {
    const int i = e % n, c = e / n;
    p.B[(size_t)i + (size_t)(c0 + c) * p.ldb] = x[i + c * n];
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

void launch_lu(const LuDesc* d, int count, hipStream_t s) {
  if (count == 0) return;
  lu_kernel<<<count, TPB, 0, s>>>(d);
  HODLR_HIP_CHECK(hipGetLastError());
}

void launch_lu_solve(const SolveDesc* d, int count, int nwg, int ch, int max_n, hipStream_t s) {
  if (nwg == 0 || count == 0) return;
  const size_t lds = (size_t)max_n * ch * sizeof(double);
  if (lds > 64 * 1024) throw HipError("lu_solve: chunk does not fit in LDS");
  lu_solve_kernel<<<nwg, TPB, lds, s>>>(d, count, ch);
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
