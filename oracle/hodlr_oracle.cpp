/*
 * ORACLE -- test infrastructure only.  CPU FP64 restatement of the reference
 * HODLR path (amiraa127/Dense_HODLR): default tree, partial-pivot ACA, SVD
 * low-rank approximation, recSM factor/solve (perturbI Woodbury nodes),
 * recLU factor/solve and the log-determinant.  Only tests/, the smoke check
 * and the CPU-baseline leg of bench.py load this library; the product path
 * never calls it.
 *
 * Each routine cites the reference file:line it restates.  Matrices are
 * column-major FP64 (the reference uses Eigen::MatrixXd, also column-major).
 * Built with -ffp-contract=off so the ACA residual arithmetic is performed as
 * separate IEEE multiplies and adds in the same order as the reference's
 * Eigen expressions (no FMA), which is also how the GPU kernels evaluate it.
 *
 * Parity pins (tests/test_oracle.py): tree rules by brute-force enumeration;
 * ACA by exact-rank inputs, the dense fallback and hand-traced pivots; SVD by
 * numpy golden singular values; LU by scipy getrf golden vectors; recSM and
 * recLU against dense LU on exact-HODLR and kernel matrices; log-det against
 * the dense LU log-det of the assembled HODLR operator.
 */
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#include "../dense_hodlr_amd/csrc/kernel_fn.h"

namespace {

struct Mat {
  int r = 0, c = 0;
  std::vector<double> a;
  Mat() {}
  Mat(int rr, int cc) : r(rr), c(cc), a((size_t)rr * cc, 0.0) {}
  double& operator()(int i, int j) { return a[(size_t)i + (size_t)j * r]; }
  double operator()(int i, int j) const { return a[(size_t)i + (size_t)j * r]; }
  double* col(int j) { return a.data() + (size_t)j * r; }
  const double* col(int j) const { return a.data() + (size_t)j * r; }
};

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

/* ------------------------------------------------------------------ */
/* Entry sources: typed kernel (kernel.cpp:19-25,37-39) or dense copy   */
/* (HODLR_Matrix.cpp:59-66).                                            */
/* ------------------------------------------------------------------ */
struct Source {
  bool dense = false;
  int n = 0;
  std::vector<double> A;  // dense copy, n x n column-major
  std::vector<double> xs, ys, zs;
  HodlrKernelParams kp{};
  double operator()(int i, int j) const {
    if (dense) return A[(size_t)i + (size_t)j * n];
    return hodlr_kernel_entry(kp, i, j);
  }
  Mat block(int i0, int j0, int m, int nn) const {
    Mat B(m, nn);
    for (int j = 0; j < nn; j++)
      for (int i = 0; i < m; i++) B(i, j) = (*this)(i0 + i, j0 + j);
    return B;
  }
};

/* ------------------------------------------------------------------ */
/* Dense LU with partial pivoting: the first row of maximum |a_ik| is   */
/* the pivot (Eigen PartialPivLU / LAPACK getrf semantics).             */
/* ------------------------------------------------------------------ */
struct LU {
  Mat lu;
  std::vector<int> perm;  // row interchange at step k: swap k <-> perm[k]
  void factor(const Mat& A) {
    lu = A;
    int n = lu.r;
    perm.assign(n, 0);
    for (int k = 0; k < n; k++) {
      int p = k;
      double best = std::fabs(lu(k, k));
      for (int i = k + 1; i < n; i++) {
        double v = std::fabs(lu(i, k));
        if (v > best) { best = v; p = i; }
      }
      perm[k] = p;
      if (p != k)
        for (int j = 0; j < n; j++) std::swap(lu(k, j), lu(p, j));
      double piv = lu(k, k);
      if (piv != 0.0) {
        double inv = 1.0 / piv;
        for (int i = k + 1; i < n; i++) lu(i, k) *= inv;
      }
      for (int j = k + 1; j < n; j++) {
        double ukj = lu(k, j);
        if (ukj == 0.0) continue;
        double* cj = lu.col(j);
        const double* ck = lu.col(k);
        for (int i = k + 1; i < n; i++) cj[i] -= ck[i] * ukj;
      }
    }
  }
  /* X <- A^{-1} X in place */
  void solve_inplace(Mat& X) const {
    int n = lu.r;
    for (int c = 0; c < X.c; c++) {
      double* x = X.col(c);
      for (int k = 0; k < n; k++)
        if (perm[k] != k) std::swap(x[k], x[perm[k]]);
      for (int k = 0; k < n; k++) {
        double xk = x[k];
        if (xk == 0.0) continue;
        const double* ck = lu.col(k);
        for (int i = k + 1; i < n; i++) x[i] -= ck[i] * xk;
      }
      for (int k = n - 1; k >= 0; k--) {
        x[k] /= lu(k, k);
        double xk = x[k];
        if (xk == 0.0) continue;
        const double* ck = lu.col(k);
        for (int i = 0; i < k; i++) x[i] -= ck[i] * xk;
      }
    }
  }
  Mat solve(const Mat& B) const {
    Mat X = B;
    solve_inplace(X);
    return X;
  }
  double log_abs_det() const {
    double s = 0;
    for (int i = 0; i < lu.r; i++) s += std::log(std::fabs(lu(i, i)));
    return s;
  }
  double det_no_sign() const {  // HODLR_Matrix.cpp:1684-1688 (no permutation sign)
    double s = 1;
    for (int i = 0; i < lu.r; i++) s *= lu(i, i);
    return s;
  }
};

/* C = A * B (plain loops, j-k-i order) */
Mat matmul(const Mat& A, const Mat& B) {
  Mat C(A.r, B.c);
  for (int j = 0; j < B.c; j++) {
    double* cj = C.col(j);
    for (int k = 0; k < A.c; k++) {
      double b = B(k, j);
      if (b == 0.0) continue;
      const double* ak = A.col(k);
      for (int i = 0; i < A.r; i++) cj[i] += ak[i] * b;
    }
  }
  return C;
}
/* C = A^T * B */
Mat matmul_tn(const Mat& A, const Mat& B) {
  Mat C(A.c, B.c);
  for (int j = 0; j < B.c; j++) {
    const double* bj = B.col(j);
    for (int i = 0; i < A.c; i++) {
      const double* ai = A.col(i);
      double s = 0;
      for (int k = 0; k < A.r; k++) s += ai[k] * bj[k];
      C(i, j) = s;
    }
  }
  return C;
}
Mat rows_of(const Mat& A, int r0, int nr) {
  Mat B(nr, A.c);
  for (int j = 0; j < A.c; j++)
    for (int i = 0; i < nr; i++) B(i, j) = A(r0 + i, j);
  return B;
}
void set_rows(Mat& A, int r0, const Mat& B) {
  for (int j = 0; j < B.c; j++)
    for (int i = 0; i < B.r; i++) A(r0 + i, j) = B(i, j);
}

/* ------------------------------------------------------------------ */
/* Partial-pivot ACA: lowRank.cpp:117-279, chooseNextRowCol :291-298.   */
/* ------------------------------------------------------------------ */
int choose_next_row_col(const std::vector<char>& chosen, const std::vector<double>& absv,
                        int minPivot) {
  // lowRank.cpp:291-298: the FIRST unchosen index whose |value| exceeds minPivot
  // (and the pivot floor, HODLR_PIVOT_FLOOR)
  for (size_t idx = 0; idx < absv.size(); idx++)
    if (!chosen[idx] && absv[idx] > minPivot && absv[idx] >= HODLR_PIVOT_FLOOR) return (int)idx;
  return -1;
}

struct AcaTrace {
  std::vector<int> rows, cols;
  int iterations = 0;
};

template <class Src>
double aca_partial_pivot(const Src& A, int i0, int j0, int m, int n, double tol,
                         int minRank, int maxRank, int minPivot, Mat& W, Mat& V,
                         int& rank, AcaTrace* tr) {
  if (maxRank == 0) {  // :120-125
    W = Mat(m, 1);
    V = Mat(n, 1);
    rank = 1;
    return 1;
  }
  const int upper = std::min(m, n);  // :127
  int capW = 2;
  Mat tW(m, capW), tV(n, capW);
  std::vector<char> chosenRows(m, 0), chosenCols(n, 0);
  std::vector<double> rrow(n), rcol(m), absr(n), absc(m);
  double frobNormSq = 0, frobNorm = 0, epsilon = 1;
  int currRow = 0, currCol = 0, nextRow = 0, k = 0;

  while (((epsilon > tol) || (k < minRank)) && (k < upper)) {  // :149
    if (k == capW - 1) {  // grow storage (:152-157)
      int nc = 2 * capW;
      Mat nW(m, nc), nV(n, nc);
      std::copy(tW.a.begin(), tW.a.end(), nW.a.begin());
      std::copy(tV.a.begin(), tV.a.end(), nV.a.begin());
      tW = std::move(nW);
      tV = std::move(nV);
      capW = nc;
    }
    chosenRows[currRow] = 1;  // :159
    if (tr) tr->iterations++;
    // residual row (:161-167): sum_l W(i,l) V(:,l), then row - sum
    const int gi = i0 + currRow;
    std::vector<double> sum(n, 0.0);
    for (int l = 0; l < k; l++) {
      const double w = tW(currRow, l);
      const double* vl = tV.col(l);
      for (int j = 0; j < n; j++) sum[j] = sum[j] + w * vl[j];
    }
    for (int j = 0; j < n; j++) rrow[j] = A(gi, j0 + j) - sum[j];
    // argmax |r|, first max (:170-173)
    int maxInd = 0;
    double maxValue = -1;
    for (int j = 0; j < n; j++) {
      absr[j] = std::fabs(rrow[j]);
      if (absr[j] > maxValue) { maxValue = absr[j]; maxInd = j; }
    }
    if (maxValue <= minPivot || maxValue < HODLR_PIVOT_FLOOR) {  // :187-195: jump to the first unchosen row
      auto it = std::find(chosenRows.begin(), chosenRows.end(), 0);
      if (it == chosenRows.end()) break;
      currRow = (int)(it - chosenRows.begin());
      continue;
    }
    if (!chosenCols[maxInd]) {  // :197-203
      currCol = maxInd;
    } else {
      currCol = choose_next_row_col(chosenCols, absr, minPivot);
      if (currCol == -1) break;
    }
    chosenCols[currCol] = 1;  // :206
    const double pivot = 1 / rrow[currCol];  // :208
    const int gj = j0 + currCol;
    std::vector<double> sumc(m, 0.0);
    for (int l = 0; l < k; l++) {  // :211-215
      const double v = tV(currCol, l);
      const double* wl = tW.col(l);
      for (int i = 0; i < m; i++) sumc[i] = sumc[i] + v * wl[i];
    }
    for (int i = 0; i < m; i++) rcol[i] = A(i0 + i, gj) - sumc[i];
    maxInd = 0;
    maxValue = -1;
    for (int i = 0; i < m; i++) {  // :218-219
      absc[i] = std::fabs(rcol[i]);
      if (absc[i] > maxValue) { maxValue = absc[i]; maxInd = i; }
    }
    if (!chosenRows[maxInd])  // :221-225
      nextRow = maxInd;
    else
      nextRow = choose_next_row_col(chosenRows, absc, minPivot);
    if (nextRow == -1) break;  // :226-227 (term k is discarded)
    double* wk = tW.col(k);
    double* vk = tV.col(k);
    for (int i = 0; i < m; i++) wk[i] = pivot * rcol[i];  // :230
    for (int j = 0; j < n; j++) vk[j] = rrow[j];          // :231
    // incremental Frobenius norm (:234-239)
    double sumNorm = 0;
    for (int l = 0; l < k; l++) {
      const double* wl = tW.col(l);
      const double* vl = tV.col(l);
      double dw = 0, dv = 0;
      for (int i = 0; i < m; i++) dw += wk[i] * wl[i];
      for (int j = 0; j < n; j++) dv += vl[j] * vk[j];
      sumNorm += dw * dv;
    }
    double nw = 0, nv = 0;
    for (int i = 0; i < m; i++) nw += wk[i] * wk[i];
    for (int j = 0; j < n; j++) nv += vk[j] * vk[j];
    frobNormSq += 2 * sumNorm + nw * nv;
    frobNorm = std::sqrt(frobNormSq);
    epsilon = std::sqrt(nw) * std::sqrt(nv) / frobNorm;  // :241
    if (tr) {
      tr->rows.push_back(currRow);
      tr->cols.push_back(currCol);
    }
    currRow = nextRow;  // :244
    k++;
    if (k == maxRank) break;  // :246-247
  }
  rank = k;
  if (k == 0) {  // :251-257
    W = Mat(m, 1);
    V = Mat(n, 1);
    rank = 1;
    return epsilon;
  }
  if (k >= upper - 1) {  // :260-274: dense block with an identity factor
    Mat B(m, n);
    for (int j = 0; j < n; j++)
      for (int i = 0; i < m; i++) B(i, j) = A(i0 + i, j0 + j);
    if (n <= m) {
      W = B;
      V = Mat(n, n);
      for (int i = 0; i < n; i++) V(i, i) = 1;
      rank = n;
    } else {
      W = Mat(m, m);
      for (int i = 0; i < m; i++) W(i, i) = 1;
      V = Mat(n, m);
      for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) V(j, i) = B(i, j);
      rank = m;
    }
    return epsilon;
  }
  W = Mat(m, k);
  V = Mat(n, k);
  std::copy(tW.a.begin(), tW.a.begin() + (size_t)m * k, W.a.begin());
  std::copy(tV.a.begin(), tV.a.begin() + (size_t)n * k, V.a.begin());
  return epsilon;
}

/* ------------------------------------------------------------------ */
/* SVD low-rank approximation: lowRank.cpp:520-569 (rank = number of    */
/* leading singular values >= tol, an absolute threshold; stored as     */
/* U*Sigma and V, HODLR_Matrix.cpp:361-379).  The SVD itself is the     */
/* textbook one-sided Jacobi (Hestenes) method, swept to convergence.   */
/* ------------------------------------------------------------------ */
void jacobi_svd(const Mat& Ain, std::vector<double>& sigma, Mat& US, Mat& Vout) {
  bool tr = Ain.r < Ain.c;
  Mat A;
  if (tr) {
    A = Mat(Ain.c, Ain.r);
    for (int j = 0; j < Ain.c; j++)
      for (int i = 0; i < Ain.r; i++) A(j, i) = Ain(i, j);
  } else {
    A = Ain;
  }
  const int m = A.r, n = A.c;
  Mat V(n, n);
  for (int i = 0; i < n; i++) V(i, i) = 1;
  const double eps = 2.220446049250313e-16;
  // couplings below (eps * largest initial column norm)^2 are negligible at
  // the accuracy the singular values are used at (absolute, relative to the
  // largest one), like Eigen's JacobiSVD threshold relative to its largest
  // entry; without it, rotations among columns of rounding noise never stop
  double cmax2 = 0;
  for (int j = 0; j < n; j++) {
    double t = 0;
    for (int i = 0; i < m; i++) t += A(i, j) * A(i, j);
    cmax2 = std::max(cmax2, t);
  }
  const double floor2 = eps * eps * cmax2;
  for (int sweep = 0; sweep < 60; sweep++) {
    bool rotated = false;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        double* ap = A.col(p);
        double* aq = A.col(q);
        double alpha = 0, beta = 0, gamma = 0;
        for (int i = 0; i < m; i++) {
          alpha += ap[i] * ap[i];
          beta += aq[i] * aq[i];
          gamma += ap[i] * aq[i];
        }
        if (gamma == 0.0 || std::fabs(gamma) <= eps * std::sqrt(alpha * beta) ||
            std::fabs(gamma) <= floor2)
          continue;
        rotated = true;
        double zeta = (beta - alpha) / (2 * gamma);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1 + zeta * zeta));
        double c = 1 / std::sqrt(1 + t * t), s = c * t;
        for (int i = 0; i < m; i++) {
          double x = ap[i], y = aq[i];
          ap[i] = c * x - s * y;
          aq[i] = s * x + c * y;
        }
        double* vp = V.col(p);
        double* vq = V.col(q);
        for (int i = 0; i < n; i++) {
          double x = vp[i], y = vq[i];
          vp[i] = c * x - s * y;
          vq[i] = s * x + c * y;
        }
      }
    if (!rotated) break;
  }
  std::vector<double> s(n);
  for (int j = 0; j < n; j++) {
    double t = 0;
    for (int i = 0; i < m; i++) t += A(i, j) * A(i, j);
    s[j] = std::sqrt(t);
  }
  std::vector<int> ord(n);
  for (int j = 0; j < n; j++) ord[j] = j;
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return s[a] > s[b]; });
  sigma.resize(n);
  Mat AS(m, n), VS(n, n);
  for (int j = 0; j < n; j++) {
    sigma[j] = s[ord[j]];
    std::copy(A.col(ord[j]), A.col(ord[j]) + m, AS.col(j));
    std::copy(V.col(ord[j]), V.col(ord[j]) + n, VS.col(j));
  }
  if (!tr) {  // A = (A V) V^T: U*Sigma = AS, V = VS
    US = AS;
    Vout = VS;
  } else {  // A^T = AS VS^T  =>  A = VS AS^T; U*Sigma = VS*diag(sigma), V = AS / sigma
    US = Mat(Ain.r, n);
    Vout = Mat(Ain.c, n);
    for (int j = 0; j < n; j++) {
      for (int i = 0; i < Ain.r; i++) US(i, j) = VS(i, j) * sigma[j];
      double inv = sigma[j] > 0 ? 1.0 / sigma[j] : 0.0;
      for (int i = 0; i < Ain.c; i++) Vout(i, j) = AS(i, j) * inv;
    }
  }
}

int svd_lowrank(const Mat& B, double tol, int minRank, Mat& W, Mat& V,
                std::vector<double>* sig_out) {
  std::vector<double> sigma;
  Mat US, VV;
  jacobi_svd(B, sigma, US, VV);
  if (sig_out) *sig_out = sigma;
  int ns = (int)sigma.size();
  int rank = ns;
  for (int i = 0; i < ns; i++)
    if (sigma[i] < tol) { rank = i; break; }  // lowRank.cpp:529-533
  if (rank < minRank) rank = minRank;
  if (rank == 0) {  // :549-561
    W = Mat(B.r, 1);
    V = Mat(B.c, 1);
    return 1;
  }
  W = Mat(B.r, rank);
  V = Mat(B.c, rank);
  std::copy(US.a.begin(), US.a.begin() + (size_t)B.r * rank, W.a.begin());
  std::copy(VV.a.begin(), VV.a.begin() + (size_t)B.c * rank, V.a.begin());
  return rank;
}

/* ------------------------------------------------------------------ */
/* perturbI: perturbI.cpp:15-83.  U = blockdiag(U_t^SM, U_b^SM),        */
/* VT = [[0, V_t^T], [V_b^T, 0]], C = I + VT U, solve(R) = R - U C^-1 VT R */
/* ------------------------------------------------------------------ */
struct PerturbI {
  Mat U, VT;
  LU eq;
  bool built = false;
  void build(const Mat& tU, const Mat& tV, const Mat& bU, const Mat& bV) {
    int K = tU.c + bU.c;
    int s = tU.r + bU.r;
    U = Mat(s, K);
    VT = Mat(K, s);
    for (int j = 0; j < tU.c; j++)
      for (int i = 0; i < tU.r; i++) U(i, j) = tU(i, j);
    for (int j = 0; j < bU.c; j++)
      for (int i = 0; i < bU.r; i++) U(tU.r + i, tU.c + j) = bU(i, j);
    for (int j = 0; j < tV.c; j++)  // topRightCorner = topV^T
      for (int i = 0; i < tV.r; i++) VT(j, tU.r + i) = tV(i, j);
    for (int j = 0; j < bV.c; j++)  // bottomLeftCorner = bottV^T
      for (int i = 0; i < bV.r; i++) VT(tV.c + j, i) = bV(i, j);
    Mat C = matmul(VT, U);
    for (int i = 0; i < K; i++) C(i, i) += 1;
    eq.factor(C);
    built = true;
  }
  Mat solve(const Mat& R) const {
    Mat y = matmul(VT, R);
    eq.solve_inplace(y);
    Mat t = matmul(U, y);
    Mat out = R;
    for (size_t i = 0; i < out.a.size(); i++) out.a[i] -= t.a[i];
    return out;
  }
};

/* ------------------------------------------------------------------ */
/* HODLR tree node (HODLR_Tree.hpp:20-60) and default tree              */
/* (HODLR_Tree.cpp:33-132).                                             */
/* ------------------------------------------------------------------ */
struct Node {
  Node* left = nullptr;
  Node* right = nullptr;
  int level = 0, min_i = 0, max_i = 0, split = -1;
  bool leaf = false;
  int kt = 0, kb = 0;
  int lr = -1;                               // LR_Method of the node (-1: the matrix's)
  int minR[2] = {-1, -1}, maxR[2] = {-1, -1};  // topOffDiag_/bottOffDiag_ min/maxRank
  Mat tU, tV, bU, bV;  // top: rows [min,split] x cols [split+1,max]
  Mat tU_SM, bU_SM;
  Mat leafMat;
  LU leafLU;
  Mat leafInv;  // explicit inverse (the leaf-inverse variant only)
  PerturbI pI;
  Mat topSolnLR, bottSolnLR;  // recLU: A11^-1 U_t, A22^-1 U_b
  LU schur;
};

struct Hodlr {
  Source src;
  int n = 0, leaf = 30, numLevels = 0;
  double tol = 1e-6;  // HODLR_Matrix.cpp:7
  int minPivot = 0;   // int parameter in lowRank.hpp:25
  int method = 0;     // 0 = partialPiv_ACA (HODLR_Tree.cpp:5), 1 = SVD
  std::vector<std::unique_ptr<Node>> pool;
  Node* root = nullptr;
  std::vector<std::vector<Node*>> levels;  // non-leaf nodes per level, left to right
  std::vector<Node*> leaves;
  bool compressed = false, smFactored = false, luFactored = false;
  /* Variant for measurement, not the reference: leaf solves multiply by an
   * explicitly formed inverse (LU solves of the identity), as the GPU
   * engine's leaf steps do. */
  bool leafInverse = false;
  void leafSolve(Node* c, Mat& rhs) {
    if (!leafInverse) {
      c->leafLU.solve_inplace(rhs);
      return;
    }
    if (c->leafInv.r == 0) {
      const int s = c->leafLU.lu.r;
      c->leafInv = Mat(s, s);
      for (int i = 0; i < s; i++) c->leafInv(i, i) = 1;
      c->leafLU.solve_inplace(c->leafInv);
    }
    rhs = matmul(c->leafInv, rhs);
  }
  double tCompress = 0, tSmFactor = 0, tSmSolve = 0, tLuFactor = 0, tLuSolve = 0;

  Node* newNode() {
    pool.emplace_back(new Node());
    return pool.back().get();
  }
  void addLevel(Node* nd) {
    if ((int)levels.size() <= nd->level) levels.resize(nd->level + 1);
    levels[nd->level].push_back(nd);
  }
  void build(Node* r) {  // HODLR_Tree.cpp:59-132
    int topSize = r->split - r->min_i + 1;
    int botSize = r->max_i - r->split;
    if (r->level > numLevels - 1) numLevels = r->level + 1;
    Node* L = newNode();
    Node* R = newNode();
    L->min_i = r->min_i;
    L->max_i = r->min_i + topSize - 1;
    L->level = r->level + 1;
    r->left = L;
    if (topSize > leaf) {
      L->split = r->min_i + topSize / 2 - 1;  // :86-87
      addLevel(L);
      build(L);
    } else {
      L->leaf = true;
      leaves.push_back(L);
      if (L->level > numLevels - 1) numLevels = L->level + 1;
    }
    R->min_i = r->min_i + topSize;
    R->max_i = r->max_i;
    R->level = r->level + 1;
    r->right = R;
    if (botSize > leaf) {
      R->split = r->max_i - botSize / 2;  // :117-118
      addLevel(R);
      build(R);
    } else {
      R->leaf = true;
      leaves.push_back(R);
      if (R->level > numLevels - 1) numLevels = R->level + 1;
    }
  }
  void createTree() {  // HODLR_Tree.cpp:33-57
    root = newNode();
    root->min_i = 0;
    root->max_i = n - 1;
    root->split = n / 2 - 1;
    root->level = 0;
    numLevels = 1;
    if (n < leaf) {  // root leaf rule uses '<' (:49)
      root->leaf = true;
      leaves.push_back(root);
    } else {
      addLevel(root);
      build(root);
    }
  }
  /* userTree_To_HODLRTree (HODLR_Tree.cpp:207-321); user nodes are 8 ints:
   * split, top/bottom min rank, top/bottom max rank, method, left, right */
  void fromUser(const int* un, int ui, int level, int lo, int hi, Node* nd) {
    nd->level = level;
    nd->min_i = lo;
    nd->max_i = hi;
    if (level > numLevels - 1) numLevels = level + 1;
    const int* u = un + 8 * ui;
    const int size = hi - lo + 1;
    if (size <= leaf && u[0] == -1) {  // :212-231
      nd->leaf = true;
      leaves.push_back(nd);
      return;
    }
    nd->split = u[0];  // :233-246
    nd->minR[0] = u[1];
    nd->minR[1] = u[2];
    nd->maxR[0] = u[3];
    nd->maxR[1] = u[4];
    nd->lr = u[5];
    addLevel(nd);
    Node* L = newNode();
    Node* R = newNode();
    nd->left = L;
    nd->right = R;
    if (u[6] == -1 && u[7] == -1) {  // default children (:259-315)
      const int ls = u[0] - lo + 1, rs = hi - u[0];
      L->level = level + 1;
      L->min_i = lo;
      L->max_i = u[0];
      if (ls <= leaf) {
        L->leaf = true;
        leaves.push_back(L);
        if (L->level > numLevels - 1) numLevels = L->level + 1;
      } else {
        L->split = lo + ls / 2 - 1;
        addLevel(L);
        build(L);
      }
      R->level = level + 1;
      R->min_i = u[0] + 1;
      R->max_i = hi;
      if (rs <= leaf) {
        R->leaf = true;
        leaves.push_back(R);
        if (R->level > numLevels - 1) numLevels = R->level + 1;
      } else {
        R->split = u[0] + rs / 2;  // (the reference's rule here, not max - rs/2)
        addLevel(R);
        build(R);
      }
    } else {  // :316-319
      fromUser(un, u[6], level + 1, lo, u[0], L);
      fromUser(un, u[7], level + 1, u[0] + 1, hi, R);
    }
  }
  void createFromUser(const int* un, int count) {  // createFromUsrTree (:323-331)
    if (count == 0) {
      createTree();
      return;
    }
    root = newNode();
    numLevels = 1;
    fromUser(un, 0, 0, 0, n - 1, root);
  }
  void resetAttributes() {  // HODLR_Matrix.cpp:1232-1246
    compressed = smFactored = luFactored = false;
    tCompress = tSmFactor = tSmSolve = tLuFactor = tLuSolve = 0;
  }

  /* storeLRinTree: HODLR_Matrix.cpp:257-430 (preorder) */
  void store(Node* r) {
    if (!r) return;
    if (r->leaf) {
      int s = r->max_i - r->min_i + 1;
      r->leafMat = src.block(r->min_i, r->min_i, s, s);
      return;
    }
    int mT = r->split - r->min_i + 1, nT = r->max_i - r->split;
    int mB = r->max_i - r->split, nB = r->split - r->min_i + 1;
    const int m = r->lr >= 0 ? r->lr : method;  // the node's LR_Method (:330-380)
    if (m == 0) {
      aca_partial_pivot(src, r->min_i, r->split + 1, mT, nT, tol, r->minR[0], r->maxR[0],
                        minPivot, r->tU, r->tV, r->kt, nullptr);
      aca_partial_pivot(src, r->split + 1, r->min_i, mB, nB, tol, r->minR[1], r->maxR[1],
                        minPivot, r->bU, r->bV, r->kb, nullptr);
    } else {
      r->kt = svd_lowrank(src.block(r->min_i, r->split + 1, mT, nT), tol, r->minR[0], r->tU,
                          r->tV, nullptr);
      r->kb = svd_lowrank(src.block(r->split + 1, r->min_i, mB, nB), tol, r->minR[1], r->bU,
                          r->bV, nullptr);
    }
    store(r->left);
    store(r->right);
  }
  void compress() {
    if (compressed) return;
    double t0 = now_s();
    store(root);
    tCompress = now_s() - t0;
    compressed = true;
  }

  /* recSM_Factorize(node, left, right, desLevel): HODLR_Matrix.cpp:537-592 */
  void smFactorLevel(Node* r, std::vector<Node*>& lc, std::vector<Node*>& rc, int d) {
    if (r->left->level == d) {
      if (r->left->leaf) r->left->leafLU.factor(r->left->leafMat);
      lc.push_back(r->left);
    } else if (!r->left->leaf) {
      std::vector<Node*> l, rr;
      smFactorLevel(r->left, l, rr, d);
      l.insert(l.end(), rr.begin(), rr.end());
      lc = l;
    }
    if (r->right->level == d) {
      if (r->right->leaf) r->right->leafLU.factor(r->right->leafMat);
      rc.push_back(r->right);
    } else if (!r->right->leaf) {
      std::vector<Node*> l, rr;
      smFactorLevel(r->right, l, rr, d);
      rr.insert(rr.end(), l.begin(), l.end());
      rc = rr;
    }
    if (r->tU_SM.r == 0) r->tU_SM = r->tU;  // :561-565
    if (r->bU_SM.r == 0) r->bU_SM = r->bU;
    for (Node* c : lc) {  // :567-578
      int cs = c->max_i - c->min_i + 1, off = c->min_i - r->min_i;
      Mat rhs = rows_of(r->tU_SM, off, cs);
      if (c->leaf) {
        leafSolve(c, rhs);
      } else {
        if (!c->pI.built) c->pI.build(c->tU_SM, c->tV, c->bU_SM, c->bV);
        rhs = c->pI.solve(rhs);
      }
      set_rows(r->tU_SM, off, rhs);
    }
    for (Node* c : rc) {  // :580-591
      int cs = c->max_i - c->min_i + 1, off = c->min_i - r->split - 1;
      Mat rhs = rows_of(r->bU_SM, off, cs);
      if (c->leaf) {
        leafSolve(c, rhs);
      } else {
        if (!c->pI.built) c->pI.build(c->tU_SM, c->tV, c->bU_SM, c->bV);
        rhs = c->pI.solve(rhs);
      }
      set_rows(r->bU_SM, off, rhs);
    }
  }
  void smFactor() {  // :520-534
    if (smFactored) return;
    double t0 = now_s();
    for (int i = numLevels - 1; i > 0; i--) {
      std::vector<Node*> l, r;
      smFactorLevel(root, l, r, i);
    }
    if (!root->leaf) root->pI.build(root->tU_SM, root->tV, root->bU_SM, root->bV);
    // n < leaf: the reference never factors the root leaf (the loop above is
    // empty) and solves with an unfactored PartialPivLU; it is factored here.
    if (root->leaf) root->leafLU.factor(root->leafMat);
    tSmFactor = now_s() - t0;
    smFactored = true;
  }
  void smSolve(Node* r, Mat& X) {  // :594-610
    int s = r->max_i - r->min_i + 1;
    if (r->leaf) {
      Mat b = rows_of(X, r->min_i, s);
      leafSolve(r, b);
      set_rows(X, r->min_i, b);
      return;
    }
    smSolve(r->left, X);
    smSolve(r->right, X);
    Mat b = rows_of(X, r->min_i, s);
    set_rows(X, r->min_i, r->pI.solve(b));
  }

  /* recLU_Factorize(rhs, node): HODLR_Matrix.cpp:612-692 */
  Mat luFactor(const Mat& rhs, Node* r) {
    if (r->leaf) {
      r->leafLU.factor(r->leafMat);
      return r->leafLU.solve(rhs);
    }
    int ts = r->split - r->min_i + 1, bs = r->max_i - r->split, pc = rhs.c;
    Mat tr(ts, pc + r->tU.c);
    for (int j = 0; j < pc; j++)
      for (int i = 0; i < ts; i++) tr(i, j) = rhs(i, j);
    for (int j = 0; j < r->tU.c; j++)
      for (int i = 0; i < ts; i++) tr(i, pc + j) = r->tU(i, j);
    Mat tsol = luFactor(tr, r->left);
    Mat tp(ts, pc);
    r->topSolnLR = Mat(ts, r->tU.c);
    for (int j = 0; j < pc; j++)
      for (int i = 0; i < ts; i++) tp(i, j) = tsol(i, j);
    for (int j = 0; j < r->tU.c; j++)
      for (int i = 0; i < ts; i++) r->topSolnLR(i, j) = tsol(i, pc + j);
    Mat br(bs, pc + r->bU.c);
    for (int j = 0; j < pc; j++)
      for (int i = 0; i < bs; i++) br(i, j) = rhs(ts + i, j);
    for (int j = 0; j < r->bU.c; j++)
      for (int i = 0; i < bs; i++) br(i, pc + j) = r->bU(i, j);
    Mat bsol = luFactor(br, r->right);
    Mat bp(bs, pc);
    r->bottSolnLR = Mat(bs, r->bU.c);
    for (int j = 0; j < pc; j++)
      for (int i = 0; i < bs; i++) bp(i, j) = bsol(i, j);
    for (int j = 0; j < r->bU.c; j++)
      for (int i = 0; i < bs; i++) r->bottSolnLR(i, j) = bsol(i, pc + j);
    // Schur matrix S = [[I_kb, V_b^T X_t], [V_t^T X_b, I_kt]] (:661-675)
    int kt = r->kt, kb = r->kb, sd = kt + kb;
    Mat S(sd, sd);
    for (int i = 0; i < kb; i++) S(i, i) = 1;
    for (int i = 0; i < kt; i++) S(kb + i, kb + i) = 1;
    Mat tr12 = matmul_tn(r->bV, r->topSolnLR);   // kb x kt
    Mat bl21 = matmul_tn(r->tV, r->bottSolnLR);  // kt x kb
    for (int j = 0; j < kt; j++)
      for (int i = 0; i < kb; i++) S(i, kb + j) = tr12(i, j);
    for (int j = 0; j < kb; j++)
      for (int i = 0; i < kt; i++) S(kb + i, j) = bl21(i, j);
    r->schur.factor(S);
    return luApply(r, tp, bp);
  }
  Mat luApply(Node* r, const Mat& tsol, const Mat& bsol) {  // :677-687, :726-750
    int kt = r->kt, kb = r->kb, pc = tsol.c;
    Mat y(kt + kb, pc);
    Mat y1 = matmul_tn(r->bV, tsol);  // kb x pc
    Mat y2 = matmul_tn(r->tV, bsol);  // kt x pc
    for (int j = 0; j < pc; j++) {
      for (int i = 0; i < kb; i++) y(i, j) = y1(i, j);
      for (int i = 0; i < kt; i++) y(kb + i, j) = y2(i, j);
    }
    r->schur.solve_inplace(y);
    Mat yb(kt, pc), yt(kb, pc);  // bottomRows(kt), topRows(kb)
    for (int j = 0; j < pc; j++) {
      for (int i = 0; i < kt; i++) yb(i, j) = y(kb + i, j);
      for (int i = 0; i < kb; i++) yt(i, j) = y(i, j);
    }
    Mat t1 = matmul(r->topSolnLR, yb), t2 = matmul(r->bottSolnLR, yt);
    Mat out(tsol.r + bsol.r, pc);
    for (int j = 0; j < pc; j++) {
      for (int i = 0; i < tsol.r; i++) out(i, j) = tsol(i, j) - t1(i, j);
      for (int i = 0; i < bsol.r; i++) out(tsol.r + i, j) = bsol(i, j) - t2(i, j);
    }
    return out;
  }
  Mat luSolve(const Mat& rhs, Node* r) {  // :694-755
    if (r->leaf) return r->leafLU.solve(rhs);
    int ts = r->split - r->min_i + 1, bs = r->max_i - r->split;
    Mat tsol = luSolve(rows_of(rhs, 0, ts), r->left);
    Mat bsol = luSolve(rows_of(rhs, ts, bs), r->right);
    return luApply(r, tsol, bsol);
  }
  void luFactorTop() {  // :505-518 (factor with a dummy linspace(-2,2) RHS)
    if (luFactored) return;
    double t0 = now_s();
    Mat dummy(n, 1);
    for (int i = 0; i < n; i++) dummy(i, 0) = n > 1 ? -2.0 + 4.0 * i / (n - 1) : -2.0;
    luFactor(dummy, root);
    tLuFactor = now_s() - t0;
    luFactored = true;
  }

  /* log-determinant: HODLR_Matrix.cpp:1671-1697, perturbI.cpp:61-83 */
  void det(Node* r, double& la, double& dp) {
    if (r->leaf) {
      la += r->leafLU.log_abs_det();
      dp *= r->leafLU.det_no_sign();
      return;
    }
    la += r->pI.eq.log_abs_det();
    dp *= r->pI.eq.det_no_sign();
    det(r->left, la, dp);
    det(r->right, la, dp);
  }

  /* R = B - A_H X with every product and sum carried as a double-double
   * pair (TwoProd by fma, TwoSum; the accumulation of Ogita, Rump and Oishi,
   * "Accurate sum and dot product", SIAM J. Sci. Comput. 26 (2005),
   * Algorithm Dot2), rounded once at the end: as accurate as the residual
   * computed in twice the working precision.  No reference counterpart: the
   * refinement check of the build's solves (the reference evaluates
   * ||A x - b|| in FP64 for printing only, HODLR_Matrix.cpp:835-836). */
  static void two_sum(double a, double b, double& s, double& e) {
    s = a + b;
    const double z = s - a;
    e = (a - (s - z)) + (b - z);
  }
  static void dd_madd(double& hi, double& lo, double a, double b) {
    const double p = a * b;
    const double pe = std::fma(a, b, -p);
    double s, e;
    two_sum(hi, p, s, e);
    hi = s;
    lo += e + pe;
  }
  void matvec_dd(Node* r, const Mat& X, Mat& Yh, Mat& Yl) {
    if (r->leaf) {
      int s = r->max_i - r->min_i + 1;
      for (int c = 0; c < X.c; c++)
        for (int i = 0; i < s; i++) {
          double h = Yh(r->min_i + i, c), l = Yl(r->min_i + i, c);
          for (int j = 0; j < s; j++) dd_madd(h, l, r->leafMat(i, j), X(r->min_i + j, c));
          Yh(r->min_i + i, c) = h;
          Yl(r->min_i + i, c) = l;
        }
      return;
    }
    const int ts = r->split - r->min_i + 1, bs = r->max_i - r->split;
    auto apply = [&](const Mat& U, const Mat& V, int xrow0, int nx, int yrow0, int ny) {
      const int k = V.c;
      for (int c = 0; c < X.c; c++) {
        std::vector<double> th(k, 0.0), tl(k, 0.0);
        for (int q = 0; q < k; q++)
          for (int i = 0; i < nx; i++) dd_madd(th[q], tl[q], V(i, q), X(xrow0 + i, c));
        for (int i = 0; i < ny; i++) {
          double h = Yh(yrow0 + i, c), l = Yl(yrow0 + i, c);
          for (int q = 0; q < k; q++) {
            dd_madd(h, l, U(i, q), th[q]);
            dd_madd(h, l, U(i, q), tl[q]);
          }
          Yh(yrow0 + i, c) = h;
          Yl(yrow0 + i, c) = l;
        }
      }
    };
    apply(r->tU, r->tV, r->split + 1, bs, r->min_i, ts);
    apply(r->bU, r->bV, r->min_i, ts, r->split + 1, bs);
    matvec_dd(r->left, X, Yh, Yl);
    matvec_dd(r->right, X, Yh, Yl);
  }
  Mat residual_dd(const Mat& B, const Mat& X) {
    Mat Yh(n, X.c), Yl(n, X.c), R(n, X.c);
    matvec_dd(root, X, Yh, Yl);
    for (size_t e = 0; e < R.a.size(); e++) {
      double s, t;
      two_sum(B.a[e], -Yh.a[e], s, t);
      R.a[e] = s + (t - Yl.a[e]);
    }
    return R;
  }

  /* A_H * X from leaves and U V^T (the operator of fill_Block, :1404-1480) */
  void matvec(Node* r, const Mat& X, Mat& Y) {
    if (r->leaf) {
      int s = r->max_i - r->min_i + 1;
      for (int c = 0; c < X.c; c++)
        for (int j = 0; j < s; j++) {
          double xv = X(r->min_i + j, c);
          for (int i = 0; i < s; i++) Y(r->min_i + i, c) += r->leafMat(i, j) * xv;
        }
      return;
    }
    int ts = r->split - r->min_i + 1, bs = r->max_i - r->split;
    Mat xt = rows_of(X, r->min_i, ts), xb = rows_of(X, r->split + 1, bs);
    Mat yt = matmul(r->tU, matmul_tn(r->tV, xb));
    Mat yb = matmul(r->bU, matmul_tn(r->bV, xt));
    for (int c = 0; c < X.c; c++) {
      for (int i = 0; i < ts; i++) Y(r->min_i + i, c) += yt(i, c);
      for (int i = 0; i < bs; i++) Y(r->split + 1 + i, c) += yb(i, c);
    }
    matvec(r->left, X, Y);
    matvec(r->right, X, Y);
  }
};

}  // namespace

/* ==================================================================== */
/* C interface (ctypes)                                                  */
/* ==================================================================== */
extern "C" {

double orc_exp(double x) { return hodlr_exp(x); }

static void fill_kernel(Source& s, int n, int family, int dim, const double* pts, double ell,
                        double diag, int diag_mode) {
  s.dense = false;
  s.n = n;
  s.xs.assign(n, 0.0);
  s.ys.assign(n, 0.0);
  s.zs.assign(n, 0.0);
  for (int i = 0; i < n; i++) {
    s.xs[i] = pts[(size_t)i * dim];
    if (dim > 1) s.ys[i] = pts[(size_t)i * dim + 1];
    if (dim > 2) s.zs[i] = pts[(size_t)i * dim + 2];
  }
  hodlr_kernel_params_init(&s.kp, family, dim, ell, diag, diag_mode, s.xs.data(),
                           s.ys.data(), s.zs.data());
}

void orc_kernel_block(int n, int family, int dim, const double* pts, double ell, double diag,
                      int diag_mode, int i0, int j0, int m, int nn, double* out) {
  Source s;
  fill_kernel(s, n, family, dim, pts, ell, diag, diag_mode);
  for (int j = 0; j < nn; j++)
    for (int i = 0; i < m; i++) out[(size_t)i + (size_t)j * m] = s(i0 + i, j0 + j);
}

void* orc_create_kernel(int n, int family, int dim, const double* pts, double ell, double diag,
                        int diag_mode, int leaf) {
  Hodlr* h = new Hodlr();
  fill_kernel(h->src, n, family, dim, pts, ell, diag, diag_mode);
  h->n = n;
  h->leaf = leaf;
  h->createTree();
  return h;
}

void* orc_create_kernel_tree(int n, int family, int dim, const double* pts, double ell, double diag,
                             int diag_mode, int leaf, const int* nodes, int count) {
  Hodlr* h = new Hodlr();
  fill_kernel(h->src, n, family, dim, pts, ell, diag, diag_mode);
  h->n = n;
  h->leaf = leaf;
  h->createFromUser(nodes, count);
  return h;
}

void* orc_create_dense_tree(int n, const double* A, int lda, int leaf, const int* nodes,
                            int count) {
  Hodlr* h = new Hodlr();
  h->src.dense = true;
  h->src.n = n;
  h->src.A.resize((size_t)n * n);
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) h->src.A[(size_t)i + (size_t)j * n] = A[(size_t)i + (size_t)j * lda];
  h->n = n;
  h->leaf = leaf;
  h->createFromUser(nodes, count);
  return h;
}

void* orc_create_dense(int n, const double* A, int lda, int leaf) {
  Hodlr* h = new Hodlr();
  h->src.dense = true;
  h->src.n = n;
  h->src.A.resize((size_t)n * n);
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) h->src.A[(size_t)i + (size_t)j * n] = A[(size_t)i + (size_t)j * lda];
  h->n = n;
  h->leaf = leaf;
  h->createTree();
  return h;
}

void orc_destroy(void* p) { delete (Hodlr*)p; }

void orc_set_tolerance(void* p, double tol) {
  Hodlr* h = (Hodlr*)p;
  if (h->tol != tol) {
    h->tol = tol;
    h->resetAttributes();
    for (auto& nd : h->pool) {
      nd->tU_SM = Mat();
      nd->bU_SM = Mat();
      nd->pI = PerturbI();
    }
  }
}

void orc_set_min_pivot(void* p, double mp) {
  Hodlr* h = (Hodlr*)p;
  int v = (int)mp;  // the ACA parameter is an int (lowRank.hpp:25)
  if (h->minPivot != v) {
    h->minPivot = v;
    h->resetAttributes();
    for (auto& nd : h->pool) {
      nd->tU_SM = Mat();
      nd->bU_SM = Mat();
      nd->pI = PerturbI();
    }
  }
}

void orc_set_lr_method(void* p, int m) {
  Hodlr* h = (Hodlr*)p;
  if (h->method != m) {  // HODLR_Matrix.cpp:1277-1286, HODLR_Tree.cpp:377-390
    h->method = m;
    h->resetAttributes();
    for (auto& nd : h->pool) {
      nd->lr = -1;
      nd->tU_SM = Mat();
      nd->bU_SM = Mat();
      nd->pI = PerturbI();
    }
  }
}

int orc_num_levels(void* p) { return ((Hodlr*)p)->numLevels; }

int orc_num_nodes(void* p) {
  Hodlr* h = (Hodlr*)p;
  int c = 0;
  for (auto& l : h->levels) c += (int)l.size();
  return c;
}

int orc_num_leaves(void* p) { return (int)((Hodlr*)p)->leaves.size(); }

void orc_compress(void* p) { ((Hodlr*)p)->compress(); }

/* Per non-leaf node, level by level, left to right. */
void orc_get_nodes(void* p, int* level, int* min_i, int* split, int* max_i, int* kt, int* kb) {
  Hodlr* h = (Hodlr*)p;
  int c = 0;
  for (auto& l : h->levels)
    for (Node* nd : l) {
      level[c] = nd->level;
      min_i[c] = nd->min_i;
      split[c] = nd->split;
      max_i[c] = nd->max_i;
      kt[c] = nd->kt;
      kb[c] = nd->kb;
      c++;
    }
}

/* Leaves in left-to-right order: level, min, max. */
void orc_get_leaves(void* p, int* level, int* min_i, int* max_i) {
  Hodlr* h = (Hodlr*)p;
  std::vector<Node*> lv = h->leaves;
  std::sort(lv.begin(), lv.end(), [](Node* a, Node* b) { return a->min_i < b->min_i; });
  for (size_t c = 0; c < lv.size(); c++) {
    level[c] = lv[c]->level;
    min_i[c] = lv[c]->min_i;
    max_i[c] = lv[c]->max_i;
  }
}

/* Low-rank factors of non-leaf node idx (level order). which: 0 tU,1 tV,2 bU,3 bV. */
int orc_get_factor(void* p, int idx, int which, double* out) {
  Hodlr* h = (Hodlr*)p;
  int c = 0;
  for (auto& l : h->levels)
    for (Node* nd : l) {
      if (c == idx) {
        const Mat& M = which == 0 ? nd->tU : which == 1 ? nd->tV : which == 2 ? nd->bU : nd->bV;
        std::copy(M.a.begin(), M.a.end(), out);
        return M.c;
      }
      c++;
    }
  return -1;
}

void orc_recsm_solve(void* p, const double* B, int nrhs, double* X) {
  Hodlr* h = (Hodlr*)p;
  h->compress();
  h->smFactor();
  Mat S(h->n, nrhs);
  std::copy(B, B + (size_t)h->n * nrhs, S.a.begin());
  double t0 = now_s();
  h->smSolve(h->root, S);
  h->tSmSolve = now_s() - t0;
  std::copy(S.a.begin(), S.a.end(), X);
}

void orc_reclu_compute(void* p) {
  Hodlr* h = (Hodlr*)p;
  h->compress();
  h->luFactorTop();
}

void orc_reclu_solve(void* p, const double* B, int nrhs, double* X) {
  Hodlr* h = (Hodlr*)p;
  h->compress();
  h->luFactorTop();
  Mat S(h->n, nrhs);
  std::copy(B, B + (size_t)h->n * nrhs, S.a.begin());
  double t0 = now_s();
  Mat R = h->luSolve(S, h->root);
  h->tLuSolve = now_s() - t0;
  std::copy(R.a.begin(), R.a.end(), X);
}

double orc_logabsdet(void* p, double* det_out) {
  Hodlr* h = (Hodlr*)p;
  h->compress();
  h->smFactor();
  double la = 0, dp = 1;
  h->det(h->root, la, dp);
  if (det_out) *det_out = dp;
  return la;
}

void orc_matvec(void* p, const double* X, int nrhs, double* Y) {
  Hodlr* h = (Hodlr*)p;
  h->compress();
  Mat XM(h->n, nrhs), YM(h->n, nrhs);
  std::copy(X, X + (size_t)h->n * nrhs, XM.a.begin());
  h->matvec(h->root, XM, YM);
  std::copy(YM.a.begin(), YM.a.end(), Y);
}

/* Measurement variant (not the reference): leaf solves by explicit inverses. */
void orc_set_leaf_inverse(void* p, int on) {
  Hodlr* h = (Hodlr*)p;
  h->leafInverse = on != 0;
  h->smFactored = false;
  for (auto& nd : h->pool) {
    nd->tU_SM = Mat();
    nd->bU_SM = Mat();
    nd->pI = PerturbI();
    nd->leafInv = Mat();
  }
}

/* R = B - A_H X in double-double arithmetic, rounded once. */
void orc_residual_dd(void* p, const double* B, const double* X, int nrhs, double* R) {
  Hodlr* h = (Hodlr*)p;
  h->compress();
  Mat BM(h->n, nrhs), XM(h->n, nrhs);
  std::copy(B, B + (size_t)h->n * nrhs, BM.a.begin());
  std::copy(X, X + (size_t)h->n * nrhs, XM.a.begin());
  Mat RM = h->residual_dd(BM, XM);
  std::copy(RM.a.begin(), RM.a.end(), R);
}

/* Iterative refinement of the recSM solution: x = recSM(b), then while the
 * largest column residual ||b - A_H x|| / ||b|| (double-double) exceeds
 * target and fewer than max_steps corrections were made, x += recSM(r).
 * With the residual in extended precision the forward error converges to
 * the order of the unit roundoff whenever the solver contracts (Higham,
 * "Accuracy and Stability of Numerical Algorithms", 2nd ed., section 12.1).
 * Returns the number of corrections; res_out gets the last residual. */
int orc_refined_solve(void* p, const double* B, int nrhs, double* X, double target, int max_steps,
                      double* res_out) {
  Hodlr* h = (Hodlr*)p;
  h->compress();
  h->smFactor();
  const int n = h->n;
  Mat BM(n, nrhs), XM(n, nrhs);
  std::copy(B, B + (size_t)n * nrhs, BM.a.begin());
  XM = BM;
  h->smSolve(h->root, XM);
  std::vector<double> bn(nrhs, 0.0);
  for (int c = 0; c < nrhs; c++) {
    double s = 0;
    for (int i = 0; i < n; i++) s += BM(i, c) * BM(i, c);
    bn[c] = s > 0 ? std::sqrt(s) : 1.0;
  }
  int steps = 0;
  double res = 0;
  for (;;) {
    Mat R = h->residual_dd(BM, XM);
    res = 0;
    for (int c = 0; c < nrhs; c++) {
      double s = 0;
      for (int i = 0; i < n; i++) s += R(i, c) * R(i, c);
      res = std::max(res, std::sqrt(s) / bn[c]);
    }
    if (res <= target || steps >= max_steps) break;
    h->smSolve(h->root, R);
    for (size_t e = 0; e < XM.a.size(); e++) XM.a[e] += R.a[e];
    steps++;
  }
  std::copy(XM.a.begin(), XM.a.end(), X);
  if (res_out) *res_out = res;
  return steps;
}

/* iterative_Solve (HODLR_Matrix.cpp:1138-1213) with IML++'s GMRES
 * (IML/include/gmres.h:80-151) restated for one right-hand side at a time:
 * tolerance init_tol and method set first, x0 = M^-1 b with M the recLU
 * (direct 1) or recSM (0) solve, restart 32, modified Gram-Schmidt, Givens
 * rotations, stopping when |s_(i+1)| / ||M^-1 b|| < tol inside a cycle or
 * ||M^-1 (b - A x)|| / ||M^-1 b|| < tol after one; A is the exact operator
 * (the dense matrix or the kernel's entries).  Returns the iterations used
 * (of the last column) or -1 when a column did not converge; res_out gets
 * the largest final preconditioned relative residual. */
int orc_iterative_solve(void* p, const double* B, int nrhs, double* X, int max_iter, double tol,
                        double init_tol, int method, int direct, double* res_out) {
  Hodlr* h = (Hodlr*)p;
  orc_set_tolerance(p, init_tol);
  orc_set_lr_method(p, method);
  h->compress();
  if (direct == 1)
    h->luFactorTop();
  else
    h->smFactor();
  const int n = h->n, m = 32;
  auto Msolve = [&](const std::vector<double>& v) {
    Mat R(n, 1);
    std::copy(v.begin(), v.end(), R.a.begin());
    if (direct == 1) {
      R = h->luSolve(R, h->root);
    } else {
      h->smSolve(h->root, R);
    }
    return R.a;
  };
  auto Aapply = [&](const std::vector<double>& v) {
    std::vector<double> y(n, 0.0);
    for (int j = 0; j < n; j++) {
      const double vj = v[j];
      for (int i = 0; i < n; i++) y[i] += h->src(i, j) * vj;
    }
    return y;
  };
  auto norm = [&](const std::vector<double>& v) {
    double s2 = 0;
    for (double e : v) s2 += e * e;
    return std::sqrt(s2);
  };
  int status = 0, last_iter = 0;
  double worst = 0;
  for (int c = 0; c < nrhs; c++) {
    std::vector<double> b(B + (size_t)c * n, B + (size_t)(c + 1) * n);
    std::vector<double> x = Msolve(b);  // HODLR_Matrix.cpp:1175-1186
    auto presid = [&]() {
      std::vector<double> Ax = Aapply(x), r(n);
      for (int i = 0; i < n; i++) r[i] = b[i] - Ax[i];
      return Msolve(r);
    };
    double normb = norm(Msolve(b));  // gmres.h:89
    std::vector<double> r = presid();
    double beta = norm(r), resid;
    if (normb == 0.0) normb = 1;
    int j = 1, it = 0;
    bool conv = false;
    if ((resid = norm(r) / normb) <= tol) conv = true;
    std::vector<std::vector<double>> v(m + 1);
    std::vector<double> H((size_t)(m + 1) * m), s(m + 1), cs(m + 1), sn(m + 1);
    auto Hm = [&](int r_, int c_) -> double& { return H[(size_t)r_ + (size_t)c_ * (m + 1)]; };
    auto rot = [](double& dx, double& dy, double c_, double s_) {
      const double t = c_ * dx + s_ * dy;
      dy = -s_ * dx + c_ * dy;
      dx = t;
    };
    auto update = [&](int k) {  // Update(x, k, H, s, v), gmres.h:53-69
      std::vector<double> y(s.begin(), s.begin() + k + 1);
      for (int i = k; i >= 0; i--) {
        y[i] /= Hm(i, i);
        for (int q = i - 1; q >= 0; q--) y[q] -= Hm(q, i) * y[i];
      }
      for (int q = 0; q <= k; q++)
        for (int e = 0; e < n; e++) x[e] += v[q][e] * y[q];
    };
    while (!conv && j <= max_iter) {
      v[0].assign(n, 0.0);
      for (int e = 0; e < n; e++) v[0][e] = r[e] * (1.0 / beta);
      std::fill(s.begin(), s.end(), 0.0);
      s[0] = beta;
      int i;
      for (i = 0; i < m && j <= max_iter; i++, j++) {
        std::vector<double> w = Msolve(Aapply(v[i]));
        for (int k = 0; k <= i; k++) {
          double d = 0;
          for (int e = 0; e < n; e++) d += w[e] * v[k][e];
          Hm(k, i) = d;
          for (int e = 0; e < n; e++) w[e] -= d * v[k][e];
        }
        Hm(i + 1, i) = norm(w);
        v[i + 1].assign(n, 0.0);
        for (int e = 0; e < n; e++) v[i + 1][e] = w[e] * (1.0 / Hm(i + 1, i));
        for (int k = 0; k < i; k++) rot(Hm(k, i), Hm(k + 1, i), cs[k], sn[k]);
        {  // GeneratePlaneRotation (gmres.h:28-43)
          const double dx = Hm(i, i), dy = Hm(i + 1, i);
          if (dy == 0.0) {
            cs[i] = 1.0;
            sn[i] = 0.0;
          } else if (std::fabs(dy) > std::fabs(dx)) {
            const double t = dx / dy;
            sn[i] = 1.0 / std::sqrt(1.0 + t * t);
            cs[i] = t * sn[i];
          } else {
            const double t = dy / dx;
            cs[i] = 1.0 / std::sqrt(1.0 + t * t);
            sn[i] = t * cs[i];
          }
        }
        rot(Hm(i, i), Hm(i + 1, i), cs[i], sn[i]);
        rot(s[i], s[i + 1], cs[i], sn[i]);
        if ((resid = std::fabs(s[i + 1]) / normb) < tol) {
          update(i);
          it = j;
          conv = true;
          break;
        }
      }
      if (conv) break;
      update(m - 1);
      r = presid();
      beta = norm(r);
      if ((resid = beta / normb) < tol) {
        it = j;
        conv = true;
      }
    }
    if (!conv) {
      status = -1;
      it = j - 1;
    }
    last_iter = it;
    worst = std::max(worst, resid);
    std::copy(x.begin(), x.end(), X + (size_t)c * n);
  }
  if (res_out) *res_out = worst;
  return status < 0 ? -1 : last_iter;
}

/* which: 0 compress, 1 recSM factor, 2 recSM solve, 3 recLU factor, 4 recLU solve */
double orc_time(void* p, int which) {
  Hodlr* h = (Hodlr*)p;
  switch (which) {
    case 0: return h->tCompress;
    case 1: return h->tSmFactor;
    case 2: return h->tSmSolve;
    case 3: return h->tLuFactor;
    default: return h->tLuSolve;
  }
}

/* Standalone partial-pivot ACA on a dense block (column-major, lda).
 * W: m x cap, V: n x cap.  Returns the rank, or -1 if cap is too small. */
int orc_aca(const double* A, int lda, int m, int n, double tol, int minRank, int maxRank,
            int minPivot, double* W, double* V, int cap, int* piv_rows, int* piv_cols,
            int* n_pivots) {
  struct DenseSrc {
    const double* A;
    int lda;
    double operator()(int i, int j) const { return A[(size_t)i + (size_t)j * lda]; }
  } s{A, lda};
  Mat Wm, Vm;
  int rank = 0;
  AcaTrace tr;
  aca_partial_pivot(s, 0, 0, m, n, tol, minRank, maxRank, minPivot, Wm, Vm, rank, &tr);
  if (n_pivots) *n_pivots = (int)tr.rows.size();
  if (piv_rows)
    for (size_t i = 0; i < tr.rows.size() && (int)i < cap; i++) {
      piv_rows[i] = tr.rows[i];
      piv_cols[i] = tr.cols[i];
    }
  if (Wm.c > cap) return -1;
  std::copy(Wm.a.begin(), Wm.a.end(), W);
  std::copy(Vm.a.begin(), Vm.a.end(), V);
  return rank;
}

/* Standalone SVD low-rank approximation; sigma receives min(m,n) values. */
int orc_svd_lowrank(const double* A, int lda, int m, int n, double tol, double* US, double* V,
                    double* sigma) {
  Mat B(m, n);
  for (int j = 0; j < n; j++)
    for (int i = 0; i < m; i++) B(i, j) = A[(size_t)i + (size_t)j * lda];
  Mat W, Vm;
  std::vector<double> sg;
  int r = svd_lowrank(B, tol, -1, W, Vm, &sg);
  std::copy(W.a.begin(), W.a.end(), US);
  std::copy(Vm.a.begin(), Vm.a.end(), V);
  std::copy(sg.begin(), sg.end(), sigma);
  return r;
}

/* Standalone partial-pivot LU (in place, column-major n x n); piv is 0-based
 * LAPACK-style interchanges. */
void orc_lu(double* A, int n, int* piv) {
  Mat M(n, n);
  std::copy(A, A + (size_t)n * n, M.a.begin());
  LU lu;
  lu.factor(M);
  std::copy(lu.lu.a.begin(), lu.lu.a.end(), A);
  for (int i = 0; i < n; i++) piv[i] = lu.perm[i];
}

/* Dense LU solve of A X = B (reference solution for small problems). */
double orc_dense_solve(const double* A, int n, const double* B, int nrhs, double* X) {
  Mat M(n, n);
  std::copy(A, A + (size_t)n * n, M.a.begin());
  LU lu;
  lu.factor(M);
  Mat R(n, nrhs);
  std::copy(B, B + (size_t)n * nrhs, R.a.begin());
  lu.solve_inplace(R);
  std::copy(R.a.begin(), R.a.end(), X);
  return lu.log_abs_det();
}

}  // extern "C"
