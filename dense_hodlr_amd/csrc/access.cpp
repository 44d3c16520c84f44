/*
 * Entry access of a HODLR matrix (block, row, col of the reference,
 * src/HODLR_Matrix.cpp:1396-1480), on the device.
 *
 *  source      the matrix the handle was built from (the dense copy or the
 *              kernel evaluated on the fly), which is what the reference's
 *              block() returns whenever it holds the data (:1404-1413);
 *  compressed  the HODLR operator A_H: leaf blocks plus U V^T of every
 *              off-diagonal block that meets the requested block, the
 *              reference's fill_Block (:1424-1480) for a matrix that only
 *              holds its tree.
 *
 * For a batch handle the matrix is the block-diagonal system of its
 * instances (entries between instances are zero).
 */
#include "engine.h"

#include <algorithm>
#include <climits>
#include <cstdint>
#include <map>

namespace hodlr {

namespace {
/* [a0, a0 + an) intersected with [b0, b0 + bn): start and length (<= 0: empty) */
inline void cut(int a0, int an, int b0, int bn, int& s, int& len) {
  s = std::max(a0, b0);
  len = std::min(a0 + an, b0 + bn) - s;
}
}  // namespace

void Engine::block_device(int which, int i0, int j0, int m, int n, double* dOut, int ldo) {
  if (which != 0 && which != 1) throw std::invalid_argument("unknown block kind");
  if (m < 0 || n < 0 || i0 < 0 || j0 < 0 || (long long)i0 + m > ntot_ ||
      (long long)j0 + n > ntot_ || ldo < std::max(1, m))
    throw std::invalid_argument("block outside the matrix");
  if (m == 0 || n == 0) return;
  HODLR_HIP_CHECK(hipMemset2DAsync(dOut, (size_t)ldo * sizeof(double), 0, (size_t)m * sizeof(double),
                                   n, stream_));
  if (which == 0) {
    std::vector<AssembleDesc> ad;
    size_t maxe = 0;
    for (int q = 0; q < (int)specs_.size(); q++) {
      const int r0 = inst_row0_[q], nq = specs_[q].n;
      int rs, rl, cs, cl;
      cut(i0, m, r0, nq, rs, rl);
      cut(j0, n, r0, nq, cs, cl);
      if (rl <= 0 || cl <= 0) continue;
      AssembleDesc a{};
      a.out = dOut + (rs - i0) + (size_t)(cs - j0) * ldo;
      a.inst = q;
      a.i0 = rs - r0;
      a.j0 = cs - r0;
      a.m = rl;
      a.n = cl;
      a.ldo = ldo;
      a.trans = 0;
      ad.push_back(a);
      maxe = std::max(maxe, (size_t)rl * cl);
    }
    for (size_t c0 = 0; c0 < ad.size(); c0 += 4096) {  // (max_elems is an int)
      std::vector<AssembleDesc> part(ad.begin() + c0, ad.begin() + std::min(ad.size(), c0 + 4096));
      size_t pe = 0;
      for (auto& a : part) pe = std::max(pe, (size_t)a.m * a.n);
      launch_assemble(arena_.push(part, stream_), (int)part.size(),
                      (int)std::min<size_t>(pe, 1u << 30), d_srcs_.p, stream_);
    }
  } else {
    compress();
    std::vector<CopyDesc> cp;
    for (auto& lf : leaves_) {
      int rs, rl, cs, cl;
      cut(i0, m, lf.row0, lf.s, rs, rl);
      cut(j0, n, lf.row0, lf.s, cs, cl);
      if (rl <= 0 || cl <= 0) continue;
      cp.push_back({Lorig_.p + lf.mat_off + (rs - lf.row0) + (size_t)(cs - lf.row0) * lf.s,
                    dOut + (rs - i0) + (size_t)(cs - j0) * ldo, rl, cl, lf.s, ldo});
    }
    launch_copy(arena_.push(cp, stream_), (int)cp.size(), stream_);
    std::vector<LrFillDesc> lr;
    size_t maxe = 0;
    for (auto& nd : nodes_) {
      const int st = nd.split - nd.min_i + 1, sb = nd.max_i - nd.split;
      const int top = nd.row0, bot = nd.row0 + st;
      const int d = nd.level;
      for (int h = 0; h < 2; h++) {  // top block: rows top.., cols bot..; bottom the reverse
        const int ur = h == 0 ? top : bot, ul = h == 0 ? st : sb;
        const int vr = h == 0 ? bot : top, vl = h == 0 ? sb : st;
        const int k = h == 0 ? nd.kt : nd.kb;
        int rs, rl, cs, cl;
        cut(i0, m, ur, ul, rs, rl);
        cut(j0, n, vr, vl, cs, cl);
        if (rl <= 0 || cl <= 0 || k <= 0) continue;
        LrFillDesc f{};
        f.U = U_[d].p + rs;
        f.V = V_[d].p + cs;
        f.out = dOut + (rs - i0) + (size_t)(cs - j0) * ldo;
        f.m = rl;
        f.n = cl;
        f.k = k;
        f.ldu = ntot_;
        f.ldv = ntot_;
        f.ldo = ldo;
        lr.push_back(f);
        maxe = std::max(maxe, (size_t)rl * cl);
      }
    }
    launch_lr_fill(arena_.push(lr, stream_), (int)lr.size(), maxe, stream_);
  }
  HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
  arena_.reset();
}

void Engine::block_host(int which, int i0, int j0, int m, int n, double* out, int ldo) {
  if (m < 0 || n < 0 || ldo < std::max(1, m)) throw std::invalid_argument("bad block size");
  if (m == 0 || n == 0) return;
  // columns in chunks of at most 2^26 entries of device staging
  const int cchunk = (int)std::max<long long>(1, std::min<long long>(n, (1LL << 26) / m));
  DBuf<double> buf;
  buf.alloc((size_t)m * cchunk);
  for (int c0 = 0; c0 < n; c0 += cchunk) {
    const int cn = std::min(cchunk, n - c0);
    block_device(which, i0, j0 + c0, m, cn, buf.p, m);
    HODLR_HIP_CHECK(hipMemcpy2D(out + (size_t)c0 * ldo, (size_t)ldo * sizeof(double), buf.p,
                                (size_t)m * sizeof(double), (size_t)m * sizeof(double), cn,
                                hipMemcpyDeviceToHost));
  }
}

}  // namespace hodlr

namespace hodlr {

/* The root's left (side 0) or right (side 1) diagonal block as a matrix of
 * its own, with its part of the tree and of the compressed factors (no
 * recompression): topDiag / bottDiag and the two halves of splitAtTop
 * (HODLR_Matrix.cpp:1507-1603).  The new matrix keeps the source of its
 * rows (the dense block, or the kernel on its points), its tree as an
 * explicit partition (levels one up, rank bounds and methods kept) and the
 * handle's settings. */
std::unique_ptr<Engine> Engine::diagonal_block(int side) {
  if (side != 0 && side != 1) throw std::invalid_argument("side must be 0 (top) or 1 (bottom)");
  if (specs_.size() != 1) throw std::invalid_argument("a batch of instances cannot be split");
  compress();  // (the reference's splitAtTop and topDiag store the factors first)
  if (nodes_.empty()) throw std::logic_error("the root is a leaf: the matrix cannot be split");
  const InstanceSpec& ps = specs_[0];
  const TreeNodeH& root = nodes_[level_nodes_[0][0]];
  const int lo = side == 0 ? 0 : root.split + 1, hi = side == 0 ? root.split : ps.n - 1;
  const int s = hi - lo + 1;
  InstanceSpec cs;
  cs.n = s;
  cs.leaf = ps.leaf;
  cs.kind = ps.kind;
  cs.kp = ps.kp;
  if (ps.kind == 0) {
    cs.xs.assign(ps.xs.begin() + lo, ps.xs.begin() + hi + 1);
    if (!ps.ys.empty()) cs.ys.assign(ps.ys.begin() + lo, ps.ys.begin() + hi + 1);
    if (!ps.zs.empty()) cs.zs.assign(ps.zs.begin() + lo, ps.zs.begin() + hi + 1);
  } else {
    cs.dense.resize((size_t)s * s);
    HODLR_HIP_CHECK(hipMemcpy2D(cs.dense.data(), (size_t)s * sizeof(double),
                                d_dense_[0].p + lo + (size_t)lo * ps.n, (size_t)ps.n * sizeof(double),
                                (size_t)s * sizeof(double), s, hipMemcpyDeviceToHost));
  }
  // the subtree as an explicit partition, root first
  std::map<std::pair<int, int>, int> at;  // (level, min_i) -> node
  for (int i = 0; i < (int)nodes_.size(); i++) at[{nodes_[i].level, nodes_[i].min_i}] = i;
  struct Rec {
    int level, lo, hi, parent, which;
  };
  std::vector<Rec> st{{1, lo, hi, -1, 0}};
  while (!st.empty()) {
    const Rec r = st.back();
    st.pop_back();
    const int me = (int)cs.tree.size();
    cs.tree.emplace_back();
    if (r.parent >= 0) (r.which == 0 ? cs.tree[r.parent].left : cs.tree[r.parent].right) = me;
    auto it = at.find({r.level, r.lo});
    if (it == at.end() || nodes_[it->second].max_i != r.hi) continue;  // a leaf
    const TreeNodeH& nd = nodes_[it->second];
    UserNode& u = cs.tree[me];
    u.split = nd.split - lo;
    u.lr = nd.lr;
    for (int h = 0; h < 2; h++) {
      u.min_rank[h] = nd.min_rank[h];
      u.max_rank[h] = nd.max_rank[h];
    }
    st.push_back({r.level + 1, nd.split + 1, r.hi, me, 1});
    st.push_back({r.level + 1, r.lo, nd.split, me, 0});
  }
  std::vector<InstanceSpec> v;
  v.push_back(std::move(cs));
  std::unique_ptr<Engine> c(new Engine(std::move(v)));
  c->lr_method_ = lr_method_;
  c->tol_ = tol_;
  c->min_pivot_ = min_pivot_;
  c->precision_ = precision_;
  c->refine_target_ = refine_target_;
  c->refine_max_ = refine_max_;
  c->tune_ = tune_;
  c->pred_params_ = pred_params_;
  c->pred_margin_ = pred_margin_;
  // ranks and factors: child level l is level l + 1 here
  for (auto& cn : c->nodes_) {
    auto it = at.find({cn.level + 1, cn.min_i + lo});
    if (it == at.end()) throw std::logic_error("split: the subtree does not match");
    cn.kt = nodes_[it->second].kt;
    cn.kb = nodes_[it->second].kb;
  }
  c->U_.clear();
  c->V_.clear();
  c->U_.resize(c->nlev_);
  c->V_.resize(c->nlev_);
  c->K_.assign(c->nlev_, 0);
  c->off_.assign(c->nlev_ + 1, 0);
  for (auto& cn : c->nodes_) c->K_[cn.level] = std::max(c->K_[cn.level], std::max(cn.kt, cn.kb));
  for (int l = 0; l < c->nlev_; l++) {
    const int k = c->K_[l];
    if (k == 0) continue;
    for (int w = 0; w < 2; w++) {
      DBuf<double>& dst = w == 0 ? c->U_[l] : c->V_[l];
      const DBuf<double>& src = w == 0 ? U_[l + 1] : V_[l + 1];
      dst.alloc((size_t)s * k);
      HODLR_HIP_CHECK(hipMemcpy2DAsync(dst.p, (size_t)s * sizeof(double), src.p + lo,
                                       (size_t)ntot_ * sizeof(double), (size_t)s * sizeof(double), k,
                                       hipMemcpyDeviceToDevice, stream_));
    }
  }
  for (int l = 0; l < c->nlev_; l++) c->off_[l + 1] = c->off_[l] + c->K_[l];
  c->ktot_ = c->off_[c->nlev_];
  // leaves: a contiguous range of this matrix's leaf buffer, in the same order
  size_t first = SIZE_MAX, end = 0;
  for (auto& lf : leaves_)
    if (lf.min_i >= lo && lf.max_i <= hi) {
      first = std::min(first, lf.mat_off);
      end = std::max(end, lf.mat_off + (size_t)lf.s * lf.s);
    }
  if (first == SIZE_MAX || end - first != c->leaf_elems_)
    throw std::logic_error("split: the leaves do not match");
  c->Lorig_.alloc(c->leaf_elems_);
  HODLR_HIP_CHECK(hipMemcpyAsync(c->Lorig_.p, Lorig_.p + first, c->leaf_elems_ * sizeof(double),
                                 hipMemcpyDeviceToDevice, stream_));
  HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
  c->compressed_ = true;
  return c;
}

}  // namespace hodlr
