/*
 * iterative_Solve (src/HODLR_Matrix.cpp:1138-1213): restarted GMRES on the
 * exact operator, left-preconditioned by the HODLR direct solve, as the
 * reference runs IML++'s GMRES (IML/include/gmres.h:80-151) with the
 * matrix's recLU_Solve as the preconditioner (:1132-1136).
 *
 * The exact operator is the matrix the handle was built from: the dense
 * copy (a GEMM) or, for kernel instances, the kernel evaluated on the fly
 * (exact.hip); the reference iterates on its dense matrixData.  Per column
 * (and per instance of a batch), as in gmres.h: x0 = M^-1 b;
 * r = M^-1 (b - A x); the restart cycles build the Krylov basis of M^-1 A,
 * the Hessenberg matrix reduced by Givens rotations, and stop when
 * |s_(i+1)| / ||M^-1 b|| < tol (inside a cycle) or ||r|| / ||M^-1 b|| < tol
 * (after one).  Columns run in lockstep; a column that has met the
 * tolerance keeps its iterate.  The basis is orthogonalized by classical
 * Gram-Schmidt applied twice (IML uses modified Gram-Schmidt: the same
 * space, different rounding).
 */
#include "engine.h"
#include "exact.h"

#include <algorithm>
#include <cmath>
#include <cstdio>

namespace hodlr {

void Engine::exact_apply(const double* dX, int ldx, int nrhs, double* dY, int ldy) {
  if (nrhs <= 0) return;
  std::vector<KopDesc> kd;
  std::vector<GemmDesc> g;
  int tiles = 0;
  for (int q = 0; q < (int)specs_.size(); q++) {
    const int r0 = inst_row0_[q], n = specs_[q].n;
    if (specs_[q].kind == 1) {  // the stored dense matrix
      GemmDesc e{};
      e.M = n;
      e.N = nrhs;
      e.K = n;
      e.trans_a = 0;
      e.A = d_dense_[q].p;
      e.lda = n;
      e.B = dX + r0;
      e.ldb = ldx;
      e.C = dY + r0;
      e.ldc = ldy;
      e.alpha = 1.0;
      e.beta = 0.0;
      e.ksplit = 1;
      g.push_back(e);
    } else {
      KopDesc k{};
      k.X = dX + r0;
      k.Y = dY + r0;
      k.inst = q;
      k.ldx = ldx;
      k.ldy = ldy;
      k.tile0 = tiles;
      tiles += kop_tiles(n, nrhs);
      kd.push_back(k);
    }
  }
  run_gemm(g);
  if (!kd.empty())
    launch_kernel_operator(arena_.push(kd, stream_), (int)kd.size(), tiles, d_srcs_.p, nrhs,
                           stream_);
}

void Engine::iterative_solve(int which, const double* dB, int ldb, int nrhs, double* dX, int ldx,
                             int max_iters, double tol) {
  if (nrhs <= 0) return;
  if (which != 0 && which != 1) throw std::invalid_argument("unknown direct solver");
  if (ldb < ntot_ || ldx < ntot_) throw std::invalid_argument("leading dimension < n");
  if (max_iters < 0 || !(tol >= 0.0)) throw std::invalid_argument("bad iteration settings");
  factorize_fp64(which);
  HODLR_HIP_CHECK(hipEventRecord(ev0_, stream_));
  const int ld = ntot_;
  const size_t N = (size_t)ntot_;
  const int m = ITER_RESTART;
  const size_t vs = N * nrhs;
  grow(xw_, vs);
  grow(rbuf_, vs);
  grow(zbuf_, vs);
  grow(basis_, (size_t)(m + 1) * vs);
  grow(nbuf_, (size_t)(m + 2) * nrhs + nrhs);
  double* V = basis_.p;
  double* dcoef = nbuf_.p;
  double* dnorm = nbuf_.p + (size_t)(m + 1) * nrhs;
  // norms per instance and column (every instance meets the tolerance on its own)
  const int nseg = (int)specs_.size();
  if (seg_rows_.n < (size_t)nseg + 1) {
    std::vector<int> r0(inst_row0_);
    r0.push_back(ntot_);
    seg_rows_.alloc(r0.size());
    HODLR_HIP_CHECK(hipMemcpy(seg_rows_.p, r0.data(), r0.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  const int nsplit = std::max(1, std::min(16, 2048 / std::max(1, nseg * nrhs)));
  grow(segbuf_, (size_t)nseg * nrhs * nsplit);
  std::vector<double> parts;
  auto segnorms = [&](const double* X, std::vector<double>& out) {
    launch_segnorm2(X, ld, seg_rows_.p, nseg, nrhs, nsplit, segbuf_.p, stream_);
    parts.resize((size_t)nseg * nrhs * nsplit);
    HODLR_HIP_CHECK(hipMemcpyAsync(parts.data(), segbuf_.p, parts.size() * sizeof(double),
                                   hipMemcpyDeviceToHost, stream_));
    HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
    out.assign((size_t)nseg * nrhs, 0.0);
    for (size_t e = 0; e < out.size(); e++) {
      double v = 0.0;
      for (int p = 0; p < nsplit; p++) v += parts[e * nsplit + p];
      out[e] = std::sqrt(v);
    }
  };
  auto upload = [&](const std::vector<double>& h, double* d) {
    HODLR_HIP_CHECK(hipMemcpyAsync(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                                   stream_));
    HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
  };
  auto precond = [&](double* X) { solve_fp64_inplace(which, X, nrhs); };  // X <- M^-1 X
  // the preconditioned residual r = M^-1 (b - A x) into rbuf_
  auto presidual = [&]() {
    exact_apply(xw_.p, ld, nrhs, zbuf_.p, ld);
    launch_sub(dB, ldb, zbuf_.p, ld, rbuf_.p, ld, ntot_, nrhs, stream_);
    precond(rbuf_.p);
  };
  // A batch is one block-diagonal system: the segment (instance) norms of a
  // column combine into the column's norm; the tolerance holds per segment
  auto colnorm = [&](const std::vector<double>& sg, int c) {
    double s = 0.0;
    for (int q = 0; q < nseg; q++) s += sg[(size_t)q + (size_t)c * nseg] * sg[(size_t)q + (size_t)c * nseg];
    return std::sqrt(s);
  };
  // x0 = M^-1 b; normb = ||M^-1 b|| (gmres.h:89-90, HODLR_Matrix.cpp:1175-1186)
  HODLR_HIP_CHECK(hipMemcpy2DAsync(xw_.p, N * sizeof(double), dB, (size_t)ldb * sizeof(double),
                                   N * sizeof(double), nrhs, hipMemcpyDeviceToDevice, stream_));
  precond(xw_.p);
  std::vector<double> bn, rs;
  segnorms(xw_.p, bn);
  for (auto& v : bn)
    if (!(v > 0)) v = 1.0;  // (:94-95)
  std::vector<double> bmin(nrhs);
  for (int c = 0; c < nrhs; c++) {
    bmin[c] = bn[(size_t)c * nseg];
    for (int q = 1; q < nseg; q++) bmin[c] = std::min(bmin[c], bn[(size_t)q + (size_t)c * nseg]);
  }
  presidual();
  segnorms(rbuf_.p, rs);
  auto seg_resid = [&](int c) {  // the largest segment residual of column c
    double r = 0.0;
    for (int q = 0; q < nseg; q++) {
      const size_t e = (size_t)q + (size_t)c * nseg;
      r = std::max(r, rs[e] / bn[e]);
    }
    return r;
  };
  std::vector<char> done(nrhs, 0);
  double res = 0.0;
  for (int c = 0; c < nrhs; c++) {
    done[c] = seg_resid(c) <= tol;  // (:97-101)
    res = std::max(res, seg_resid(c));
  }
  std::vector<double> H((size_t)(m + 1) * m * nrhs), cs((size_t)m * nrhs), sn((size_t)m * nrhs),
      sv((size_t)(m + 1) * nrhs), beta(nrhs), coef((size_t)(m + 1) * nrhs), hn, y;
  auto Hc = [&](int c) { return H.data() + (size_t)c * (m + 1) * m; };
  int j = 1;  // iteration counter of gmres.h (j <= max_iter)
  std::vector<char> active(nrhs);
  std::vector<int> upd(nrhs, 0);  // per column: basis vectors of its update this cycle
  // x += V y per column (the back substitution of Update, gmres.h:53-69)
  auto update = [&](int len_max) {
    if (len_max <= 0) return;
    y.assign((size_t)len_max * nrhs, 0.0);
    for (int c = 0; c < nrhs; c++) {
      const int k = upd[c];
      const double* Hm = Hc(c);
      double* yc = y.data() + (size_t)c * len_max;
      for (int i = 0; i < k; i++) yc[i] = sv[(size_t)c * (m + 1) + i];
      for (int i = k - 1; i >= 0; i--) {
        yc[i] /= Hm[(size_t)i * (m + 1) + i];
        for (int r = i - 1; r >= 0; r--) yc[r] -= Hm[(size_t)i * (m + 1) + r] * yc[i];
      }
    }
    upload(y, dcoef);
    launch_multi_axpy(V, vs, len_max, dcoef, xw_.p, ntot_, ld, nrhs, 1.0, 0, stream_);
  };
  while (j <= max_iters) {
    bool any = false;
    for (int c = 0; c < nrhs; c++) {
      active[c] = !done[c];
      any = any || active[c];
      beta[c] = colnorm(rs, c);
      upd[c] = 0;
    }
    if (!any) break;
    // v_0 = r / beta (:104); converged columns get a zero basis
    std::vector<double> inv(nrhs);
    for (int c = 0; c < nrhs; c++) {
      inv[c] = active[c] && beta[c] > 0 ? 1.0 / beta[c] : 0.0;
      sv[(size_t)c * (m + 1)] = active[c] ? beta[c] : 0.0;
      for (int i = 1; i <= m; i++) sv[(size_t)c * (m + 1) + i] = 0.0;
    }
    HODLR_HIP_CHECK(hipMemcpyAsync(V, rbuf_.p, vs * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    upload(inv, dnorm);
    launch_col_scale(V, ld, ntot_, nrhs, dnorm, stream_);
    int ilen = 0;
    bool all_done = false;
    for (int i = 0; i < m && j <= max_iters; i++, j++) {
      double* Vi = V + (size_t)i * vs;
      double* Vn = V + (size_t)(i + 1) * vs;
      exact_apply(Vi, ld, nrhs, Vn, ld);  // w = M^-1 A v_i (:108)
      precond(Vn);
      std::vector<double> h((size_t)(i + 1) * nrhs, 0.0);
      for (int pass = 0; pass < 2; pass++) {
        launch_multi_dot(V, vs, i + 1, Vn, ntot_, ld, nrhs, dcoef, stream_);
        HODLR_HIP_CHECK(hipMemcpyAsync(coef.data(), dcoef, (size_t)(i + 1) * nrhs * sizeof(double),
                                       hipMemcpyDeviceToHost, stream_));
        HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
        launch_multi_axpy(V, vs, i + 1, dcoef, Vn, ntot_, ld, nrhs, -1.0, 0, stream_);
        for (size_t e = 0; e < h.size(); e++) h[e] += coef[e];
      }
      launch_colnorm2(Vn, ntot_, ld, nrhs, dnorm, stream_);
      hn.resize(nrhs);
      HODLR_HIP_CHECK(hipMemcpyAsync(hn.data(), dnorm, nrhs * sizeof(double), hipMemcpyDeviceToHost,
                                     stream_));
      HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
      for (auto& v : hn) v = std::sqrt(v);
      all_done = true;
      for (int c = 0; c < nrhs; c++) {
        double* Hi = Hc(c) + (size_t)i * (m + 1);
        if (!active[c]) {  // frozen: an identity column keeps the solve regular
          for (int k = 0; k <= i + 1; k++) Hi[k] = 0.0;
          Hi[i] = 1.0;
          cs[(size_t)c * m + i] = 1.0;
          sn[(size_t)c * m + i] = 0.0;
          inv[c] = 0.0;
          continue;
        }
        for (int k = 0; k <= i; k++) Hi[k] = h[k + (size_t)c * (i + 1)];
        Hi[i + 1] = hn[c];  // (:113)
        inv[c] = hn[c] > 0 ? 1.0 / hn[c] : 0.0;
        for (int k = 0; k < i; k++) {  // ApplyPlaneRotation (:116-117)
          const double a = Hi[k], b = Hi[k + 1];
          const double cc = cs[(size_t)c * m + k], ss = sn[(size_t)c * m + k];
          Hi[k] = cc * a + ss * b;
          Hi[k + 1] = -ss * a + cc * b;
        }
        // GeneratePlaneRotation (gmres.h:28-43), applied to H and s (:119-121)
        double dx = Hi[i], dy = Hi[i + 1], cc, ss;
        if (dy == 0.0) {
          cc = 1.0;
          ss = 0.0;
        } else if (std::fabs(dy) > std::fabs(dx)) {
          const double tq = dx / dy;
          ss = 1.0 / std::sqrt(1.0 + tq * tq);
          cc = tq * ss;
        } else {
          const double tq = dy / dx;
          cc = 1.0 / std::sqrt(1.0 + tq * tq);
          ss = tq * cc;
        }
        cs[(size_t)c * m + i] = cc;
        sn[(size_t)c * m + i] = ss;
        Hi[i] = cc * dx + ss * dy;
        Hi[i + 1] = -ss * dx + cc * dy;
        double* s = sv.data() + (size_t)c * (m + 1);
        const double s0 = s[i], s1 = s[i + 1];
        s[i] = cc * s0 + ss * s1;
        s[i + 1] = -ss * s0 + cc * s1;
        upd[c] = i + 1;
        // (:123-129) a column that meets the tolerance stops here
        if (std::fabs(s[i + 1]) / bmin[c] < tol || !(hn[c] > 0)) {
          active[c] = 0;
          done[c] = 1;
          inv[c] = 0.0;
        } else {
          all_done = false;
        }
      }
      upload(inv, dnorm);
      launch_col_scale(Vn, ld, ntot_, nrhs, dnorm, stream_);  // v_{i+1} (:114)
      ilen = i + 1;
      if (all_done) {
        j++;
        break;
      }
    }
    update(ilen);  // (:125, :132)
    presidual();   // r = M^-1 (b - A x) (:133-134)
    segnorms(rbuf_.p, rs);
    res = 0.0;
    for (int c = 0; c < nrhs; c++) {
      const double rc = seg_resid(c);
      res = std::max(res, rc);
      if (rc < tol) done[c] = 1;  // (:135-140)
    }
    if (all_done) break;
  }
  res = 0.0;
  for (int c = 0; c < nrhs; c++) res = std::max(res, seg_resid(c));
  stats_.iter_iterations = j - 1;
  stats_.iter_precond_residual = res;
  // the true relative residual ||b - A x|| / ||b|| of the result (FP64)
  exact_apply(xw_.p, ld, nrhs, zbuf_.p, ld);
  launch_sub(dB, ldb, zbuf_.p, ld, rbuf_.p, ld, ntot_, nrhs, stream_);
  std::vector<double> trn, bnn;
  segnorms(rbuf_.p, trn);
  HODLR_HIP_CHECK(hipMemcpy2DAsync(zbuf_.p, N * sizeof(double), dB, (size_t)ldb * sizeof(double),
                                   N * sizeof(double), nrhs, hipMemcpyDeviceToDevice, stream_));
  segnorms(zbuf_.p, bnn);
  double tr = 0.0;
  for (size_t e = 0; e < trn.size(); e++) tr = std::max(tr, bnn[e] > 0 ? trn[e] / bnn[e] : trn[e]);
  stats_.iter_residual = tr;
  HODLR_HIP_CHECK(hipMemcpy2DAsync(dX, (size_t)ldx * sizeof(double), xw_.p, N * sizeof(double),
                                   N * sizeof(double), nrhs, hipMemcpyDeviceToDevice, stream_));
  HODLR_HIP_CHECK(hipEventRecord(ev1_, stream_));
  HODLR_HIP_CHECK(hipEventSynchronize(ev1_));
  float ms = 0;
  HODLR_HIP_CHECK(hipEventElapsedTime(&ms, ev0_, ev1_));
  stats_.solve_time = ms * 1e-3;
  arena_.reset();
  // converged as gmres.h decides: the Arnoldi estimate |s_{i+1}| / ||M^-1 b||
  // below the tolerance (:123-129) or the residual at a restart (:135-140);
  // the residual evaluated at the end (stats) can sit above the tolerance at
  // the rounding floor of its FP64 evaluation
  bool all_conv = true;
  for (int c = 0; c < nrhs; c++) all_conv = all_conv && (done[c] || seg_resid(c) <= 0.0);
  if (!all_conv) {
    char msg[200];
    std::snprintf(msg, sizeof msg,
                  "GMRES stopped at preconditioned relative residual %.3e (tolerance %.3e) after %d "
                  "iterations",
                  res, tol, j - 1);
    throw NotConverged(msg);
  }
}

}  // namespace hodlr
