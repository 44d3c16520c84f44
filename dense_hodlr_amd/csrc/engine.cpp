/*
 * Host orchestration of the HODLR engine (see engine.h for the layout).
 */
#include "engine.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace hodlr {

/* ------------------------------------------------------------------ */
/* DescArena                                                           */
/* ------------------------------------------------------------------ */
template <class T>
T* DescArena::push(const std::vector<T>& v, hipStream_t s) {
  const size_t bytes = v.size() * sizeof(T);
  const size_t need = (bytes + 255) & ~size_t(255);
  if (need == 0) return nullptr;
  if (chunks_.empty() || chunks_.back().used + need > chunks_.back().buf.n) {
    Chunk c;
    c.buf.alloc(std::max<size_t>(need, 8u << 20));
    chunks_.push_back(std::move(c));
  }
  Chunk& c = chunks_.back();
  keep_.emplace_back(reinterpret_cast<const char*>(v.data()),
                     reinterpret_cast<const char*>(v.data()) + bytes);
  char* dst = c.buf.p + c.used;
  HODLR_HIP_CHECK(hipMemcpyAsync(dst, keep_.back().data(), bytes, hipMemcpyHostToDevice, s));
  c.used += need;
  return reinterpret_cast<T*>(dst);
}

void DescArena::reset() {
  for (auto& c : chunks_) c.used = 0;
  while (chunks_.size() > 1) chunks_.pop_front();
  keep_.clear();
}

size_t DescArena::bytes() const {
  size_t b = 0;
  for (auto& c : chunks_) b += c.buf.bytes();
  return b;
}

/* ------------------------------------------------------------------ */
/* construction                                                        */
/* ------------------------------------------------------------------ */
Engine::Engine(std::vector<InstanceSpec> specs) : specs_(std::move(specs)) {
  HODLR_HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  HODLR_HIP_CHECK(hipEventCreate(&ev0_));
  HODLR_HIP_CHECK(hipEventCreate(&ev1_));
  build_trees();
  upload_sources();
}

Engine::~Engine() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

/* Default tree of each instance (src/HODLR_Tree.cpp:33-132): the root splits
 * at n/2-1; a left child of size s splits at min + s/2 - 1, a right child at
 * max - s/2; the root is a leaf if n < leaf, a child if s <= leaf. */
void Engine::build_trees() {
  struct Rec {
    int level, lo, sp, hi;
  };
  std::vector<TreeNodeH> all;
  std::vector<LeafH> lv;
  ntot_ = 0;
  nlev_ = 0;
  inst_row0_.clear();
  for (int q = 0; q < (int)specs_.size(); q++) {
    const InstanceSpec& sp = specs_[q];
    const int n = sp.n, leaf = sp.leaf;
    const int row0 = ntot_;
    inst_row0_.push_back(row0);
    ntot_ += n;
    if (n < leaf) {
      lv.push_back({q, 0, 0, n - 1, row0, n, 0, 0});
      nlev_ = std::max(nlev_, 1);
      continue;
    }
    std::vector<Rec> stack{{0, 0, n / 2 - 1, n - 1}};
    while (!stack.empty()) {
      Rec r = stack.back();
      stack.pop_back();
      TreeNodeH nd;
      nd.inst = q;
      nd.level = r.level;
      nd.min_i = r.lo;
      nd.split = r.sp;
      nd.max_i = r.hi;
      nd.row0 = row0 + r.lo;
      all.push_back(nd);
      const int ts = r.sp - r.lo + 1, bs = r.hi - r.sp;
      nlev_ = std::max(nlev_, r.level + 2);
      if (ts > leaf)
        stack.push_back({r.level + 1, r.lo, r.lo + ts / 2 - 1, r.sp});
      else
        lv.push_back({q, r.level + 1, r.lo, r.sp, row0 + r.lo, ts, 0, 0});
      if (bs > leaf)
        stack.push_back({r.level + 1, r.sp + 1, r.hi - bs / 2, r.hi});
      else
        lv.push_back({q, r.level + 1, r.sp + 1, r.hi, row0 + r.sp + 1, bs, 0, 0});
    }
  }
  std::sort(all.begin(), all.end(), [](const TreeNodeH& a, const TreeNodeH& b) {
    if (a.level != b.level) return a.level < b.level;
    if (a.inst != b.inst) return a.inst < b.inst;
    return a.min_i < b.min_i;
  });
  std::sort(lv.begin(), lv.end(), [](const LeafH& a, const LeafH& b) {
    if (a.inst != b.inst) return a.inst < b.inst;
    return a.min_i < b.min_i;
  });
  nodes_ = all;
  leaves_ = lv;
  level_nodes_.assign(nlev_, {});
  level_leaves_.assign(nlev_, {});
  for (int i = 0; i < (int)nodes_.size(); i++) level_nodes_[nodes_[i].level].push_back(i);
  size_t mo = 0, po = 0;
  for (int i = 0; i < (int)leaves_.size(); i++) {
    leaves_[i].mat_off = mo;
    leaves_[i].piv_off = po;
    mo += (size_t)leaves_[i].s * leaves_[i].s;
    po += leaves_[i].s;
    level_leaves_[leaves_[i].level].push_back(i);
  }
  leaf_elems_ = mo;
  leaf_piv_elems_ = po;
}

void Engine::upload_sources() {
  h_srcs_.resize(specs_.size());
  d_coords_.clear();
  d_dense_.clear();
  d_coords_.resize(specs_.size());
  d_dense_.resize(specs_.size());
  for (size_t q = 0; q < specs_.size(); q++) {
    InstanceSpec& sp = specs_[q];
    EntrySource es{};
    es.n = sp.n;
    if (sp.kind == 0) {
      const int n = sp.n;
      d_coords_[q].alloc((size_t)3 * n);
      std::vector<double> c((size_t)3 * n, 0.0);
      std::copy(sp.xs.begin(), sp.xs.end(), c.begin());
      if (!sp.ys.empty()) std::copy(sp.ys.begin(), sp.ys.end(), c.begin() + n);
      if (!sp.zs.empty()) std::copy(sp.zs.begin(), sp.zs.end(), c.begin() + 2 * n);
      HODLR_HIP_CHECK(hipMemcpy(d_coords_[q].p, c.data(), c.size() * sizeof(double),
                                hipMemcpyHostToDevice));
      es.kind = 0;
      es.kp = sp.kp;
      es.kp.x = d_coords_[q].p;
      es.kp.y = d_coords_[q].p + n;
      es.kp.z = d_coords_[q].p + 2 * n;
    } else {
      d_dense_[q].alloc((size_t)sp.n * sp.n);
      HODLR_HIP_CHECK(hipMemcpy(d_dense_[q].p, sp.dense.data(),
                                (size_t)sp.n * sp.n * sizeof(double), hipMemcpyHostToDevice));
      es.kind = 1;
      es.ld = sp.n;
      es.dense = d_dense_[q].p;
      std::vector<double>().swap(sp.dense);
    }
    h_srcs_[q] = es;
  }
  d_srcs_.alloc(h_srcs_.size());
  HODLR_HIP_CHECK(hipMemcpy(d_srcs_.p, h_srcs_.data(), h_srcs_.size() * sizeof(EntrySource),
                            hipMemcpyHostToDevice));
}

/* ------------------------------------------------------------------ */
/* settings (reset_attributes semantics, HODLR_Matrix.cpp:1232-1286)    */
/* ------------------------------------------------------------------ */
void Engine::reset_attributes() {
  compressed_ = false;
  leaves_factored_ = false;
  for (auto& f : fac_) {
    f.done = false;
    f.W.release();
    f.C.release();
    f.piv.release();
  }
  U_.clear();
  V_.clear();
  stats_ = Stats();
}

void Engine::set_lr_method(int m) {
  if (m != 0 && m != 1) throw std::invalid_argument("unknown low-rank method");
  if (m != lr_method_) {
    lr_method_ = m;
    reset_attributes();
  }
}

void Engine::set_tolerance(double t) {
  if (!(t >= 0.0)) throw std::invalid_argument("tolerance must be >= 0");
  if (t != tol_) {
    tol_ = t;
    reset_attributes();
  }
}

void Engine::set_min_pivot(double p) {
  const int v = (int)p;  // int parameter of the ACA (include/lowRank.hpp:25)
  if (v != min_pivot_) {
    min_pivot_ = v;
    reset_attributes();
  }
}

void Engine::set_precision(int p) {
  if (p != 0 && p != 1) throw std::invalid_argument("unknown precision");
  if (p != precision_) {
    precision_ = p;
    for (auto& f : fac_) {
      f.done = false;
      f.W.release();
      f.C.release();
      f.piv.release();
    }
  }
}

void Engine::set_refinement(double target, int max_iters) {
  if (!(target > 0.0) || max_iters < 0) throw std::invalid_argument("bad refinement settings");
  refine_target_ = target;
  refine_max_ = max_iters;
}

long long Engine::device_bytes() const {
  long long b = 0;
  for (auto& x : U_) b += x.bytes();
  for (auto& x : V_) b += x.bytes();
  b += Lorig_.bytes() + Lfac_.bytes() + Linv_.bytes() + Lpiv_.bytes();
  for (auto& f : fac_) b += f.W.bytes() + f.C.bytes() + f.piv.bytes();
  b += ws_.bytes() + tmp_.bytes() + tmp2_.bytes() + xw_.bytes();
  for (auto& x : d_coords_) b += x.bytes();
  for (auto& x : d_dense_) b += x.bytes();
  b += arena_.bytes();
  return b;
}

/* ------------------------------------------------------------------ */
/* helpers                                                             */
/* ------------------------------------------------------------------ */
void Engine::ensure_ws(size_t doubles) {
  if (ws_.n < doubles) {
    HODLR_HIP_CHECK(hipStreamSynchronize(stream_));  // queued kernels may use the old buffer
    ws_.alloc(doubles + doubles / 4);
  }
}

void Engine::grow(DBuf<double>& b, size_t doubles) {
  if (b.n < doubles) {
    HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
    b.alloc(doubles);
  }
}

void Engine::run_gemm(std::vector<GemmDesc>& descs) {
  if (descs.empty()) return;
  std::vector<int> split_ids;
  int ntiles = 0;
  const int bm = gemm_tile_m(descs);
  const size_t ws = gemm_plan(descs, split_ids, ntiles, 2048, bm);
  if (ws) ensure_ws(ws);
  const GemmDesc* d = arena_.push(descs, stream_);
  const int* ids = split_ids.empty() ? nullptr : arena_.push(split_ids, stream_);
  launch_gemm(d, (int)descs.size(), ntiles, ws_.p, ids, (int)split_ids.size(), bm, stream_);
}

static GemmDesc gemm_desc(int M, int N, int K, int ta, const double* A, int lda, const double* B,
                          int ldb, double* C, int ldc, double alpha, double beta) {
  GemmDesc g{};
  g.M = M;
  g.N = N;
  g.K = K;
  g.trans_a = ta;
  g.A = A;
  g.lda = lda;
  g.B = B;
  g.ldb = ldb;
  g.C = C;
  g.ldc = ldc;
  g.alpha = alpha;
  g.beta = beta;
  g.ksplit = 1;
  return g;
}

/* Blocked right-looking LU with partial pivoting of a batch of matrices
 * (getrf semantics): per panel of LU_NB columns, the panel factorization with
 * whole-row swaps, the triangular solve for U12 and the trailing update
 * A22 -= L21 U12 as one grouped GEMM over the batch. */
void Engine::lu_batch(std::vector<LuDesc>& descs) {
  if (descs.empty()) return;
  int maxn = 0;
  for (auto& d : descs) maxn = std::max(maxn, d.n);
  const LuDesc* dd = arena_.push(descs, stream_);
  for (int p0 = 0; p0 < maxn; p0 += LU_NB) {
    launch_lu_panel(dd, (int)descs.size(), p0, LU_NB, stream_);
    std::vector<TrsmDesc> tr;
    std::vector<GemmDesc> g;
    int nwg = 0;
    for (auto& d : descs) {
      const int pw = std::min(LU_NB, d.n - p0);
      const int rest = d.n - p0 - pw;
      if (pw <= 0 || rest <= 0) continue;
      TrsmDesc t{};
      t.T = d.A + p0 + (size_t)p0 * d.ld;
      t.B = d.A + p0 + (size_t)(p0 + pw) * d.ld;
      t.ldt = d.ld;
      t.ldb = d.ld;
      t.pw = pw;
      t.upper = 0;
      t.ncols = rest;
      t.chunk0 = nwg;
      nwg += (rest + 63) / 64;
      tr.push_back(t);
      g.push_back(gemm_desc(rest, rest, pw, 0, d.A + (p0 + pw) + (size_t)p0 * d.ld, d.ld,
                            d.A + p0 + (size_t)(p0 + pw) * d.ld, d.ld,
                            d.A + (p0 + pw) + (size_t)(p0 + pw) * d.ld, d.ld, -1.0, 1.0));
    }
    if (!tr.empty()) {
      launch_trsm(arena_.push(tr, stream_), (int)tr.size(), nwg, stream_);
      run_gemm(g);
    }
  }
}

/* B <- A^-1 B for a batch of LU-factored matrices: row interchanges, then
 * blocked forward (unit L) and backward (U) substitution, each block a small
 * triangular solve followed by a grouped GEMM update of the remaining rows. */
void Engine::lu_solve_batch(std::vector<SolveDesc>& descs) {
  if (descs.empty()) return;
  int maxn = 0;
  {
    std::vector<PermDesc> pd;
    int nwg = 0;
    for (auto& d : descs) {
      maxn = std::max(maxn, d.n);
      PermDesc p{};
      p.B = d.B;
      p.piv = d.piv;
      p.ldb = d.ldb;
      p.n = d.n;
      p.ncols = d.ncols;
      p.chunk0 = nwg;
      nwg += (d.ncols + 255) / 256;
      pd.push_back(p);
    }
    launch_permute(arena_.push(pd, stream_), (int)pd.size(), nwg, stream_);
  }
  for (int dir = 0; dir < 2; dir++) {
    const int nblk = (maxn + LU_NB - 1) / LU_NB;
    for (int bi = 0; bi < nblk; bi++) {
      std::vector<TrsmDesc> tr;
      std::vector<GemmDesc> g;
      int nwg = 0;
      for (auto& d : descs) {
        const int nb_d = (d.n + LU_NB - 1) / LU_NB;
        const int blk = dir == 0 ? bi : nb_d - 1 - bi;  // backward: from the last block
        if (blk < 0 || blk >= nb_d) continue;
        const int p0 = blk * LU_NB;
        const int pw = std::min(LU_NB, d.n - p0);
        TrsmDesc t{};
        t.T = d.LU + p0 + (size_t)p0 * d.ld_lu;
        t.B = d.B + p0;
        t.ldt = d.ld_lu;
        t.ldb = d.ldb;
        t.pw = pw;
        t.upper = dir;
        t.ncols = d.ncols;
        t.chunk0 = nwg;
        nwg += (d.ncols + 63) / 64;
        tr.push_back(t);
        if (dir == 0 && p0 + pw < d.n)  // rows below: B2 -= L21 B1
          g.push_back(gemm_desc(d.n - p0 - pw, d.ncols, pw, 0, d.LU + (p0 + pw) + (size_t)p0 * d.ld_lu,
                                d.ld_lu, d.B + p0, d.ldb, d.B + p0 + pw, d.ldb, -1.0, 1.0));
        if (dir == 1 && p0 > 0)  // rows above: B0 -= U01 B1
          g.push_back(gemm_desc(p0, d.ncols, pw, 0, d.LU + (size_t)p0 * d.ld_lu, d.ld_lu, d.B + p0,
                                d.ldb, d.B, d.ldb, -1.0, 1.0));
      }
      launch_trsm(arena_.push(tr, stream_), (int)tr.size(), nwg, stream_);
      run_gemm(g);
    }
  }
}


/* ------------------------------------------------------------------ */
/* compression: storeLRinTree (HODLR_Matrix.cpp:257-430)               */
/* ------------------------------------------------------------------ */
void Engine::compress() {
  if (compressed_) return;
  if (lr_method_ != 0)
    throw std::runtime_error("SVD low-rank method is not available in this build");
  HODLR_HIP_CHECK(hipEventRecord(ev0_, stream_));
  aca_compress();
  assemble_leaves();
  HODLR_HIP_CHECK(hipEventRecord(ev1_, stream_));
  HODLR_HIP_CHECK(hipEventSynchronize(ev1_));
  float ms = 0;
  HODLR_HIP_CHECK(hipEventElapsedTime(&ms, ev0_, ev1_));
  stats_.compress_time = ms * 1e-3;
  arena_.reset();
  compressed_ = true;
}

void Engine::aca_compress() {
  const int nn = (int)nodes_.size();
  const int nl = std::max(1, nlev_ - 1);
  U_.clear();
  V_.clear();
  U_.resize(nlev_);
  V_.resize(nlev_);
  K_.assign(nlev_, 0);
  off_.assign(nlev_ + 1, 0);
  if (nn == 0) {
    ktot_ = 0;
    return;
  }
  const int jpt = 2, ipt = 2;
  const int cpw = 256 * jpt, rpw = 256 * ipt;
  const int nb = 2 * nn;
  std::vector<AcaBlock> blk(nb);
  std::vector<int> lvl_count(nl, 0), lvl_slots(nl, 0), upper_max(nl, 1);
  size_t words = 0;
  for (int id = 0; id < nn; id++) {
    const TreeNodeH& nd = nodes_[id];
    const int st = nd.split - nd.min_i + 1, sb = nd.max_i - nd.split;
    const int g0 = inst_row0_[nd.inst];
    for (int h = 0; h < 2; h++) {
      AcaBlock& b = blk[2 * id + h];
      b.inst = nd.inst;
      b.level = nd.level;
      if (h == 0) {  // top block: rows [min, split], cols [split+1, max]
        b.i0 = nd.min_i;
        b.j0 = nd.split + 1;
        b.m = st;
        b.n = sb;
        b.w_row0 = g0 + nd.min_i;
        b.v_row0 = g0 + nd.split + 1;
      } else {  // bottom block: rows [split+1, max], cols [min, split]
        b.i0 = nd.split + 1;
        b.j0 = nd.min_i;
        b.m = sb;
        b.n = st;
        b.w_row0 = g0 + nd.split + 1;
        b.v_row0 = g0 + nd.min_i;
      }
      b.row_nwg = (b.n + cpw - 1) / cpw;
      b.col_nwg = (b.m + rpw - 1) / rpw;
      b.bits_row0 = (int)words;
      words += (b.m + 31) / 32;
      b.bits_col0 = (int)words;
      words += (b.n + 31) / 32;
      b.lidx = lvl_count[b.level]++;
      b.part_off = lvl_slots[b.level];
      lvl_slots[b.level] += b.row_nwg + b.col_nwg;
      upper_max[b.level] = std::max(upper_max[b.level], std::min(b.m, b.n));
    }
  }
  std::vector<int> cap(nl, 0);
  std::vector<DBuf<double>> P(nl);
  AcaLevels L{};
  L.ld = ntot_;
  auto alloc_level = [&](int l, int c) {
    DBuf<double> u, v;
    u.alloc((size_t)ntot_ * c);
    v.alloc((size_t)ntot_ * c);
    HODLR_HIP_CHECK(hipMemsetAsync(u.p, 0, u.bytes(), stream_));
    HODLR_HIP_CHECK(hipMemsetAsync(v.p, 0, v.bytes(), stream_));
    if (cap[l] > 0) {
      const size_t old = (size_t)ntot_ * cap[l];
      HODLR_HIP_CHECK(hipMemcpyAsync(u.p, U_[l].p, old * sizeof(double), hipMemcpyDeviceToDevice, stream_));
      HODLR_HIP_CHECK(hipMemcpyAsync(v.p, V_[l].p, old * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    }
    DBuf<double> p;
    p.alloc((size_t)lvl_slots[l] * (c + 4));
    HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
    U_[l] = std::move(u);
    V_[l] = std::move(v);
    P[l] = std::move(p);
    cap[l] = c;
    L.W[l] = U_[l].p;
    L.V[l] = V_[l].p;
    L.P[l] = P[l].p;
    L.cap[l] = c;
  };
  for (int l = 0; l < nl; l++) {
    if (lvl_count[l] == 0) continue;
    int c = init_cap_ > 0 ? init_cap_ : 32;
    if (const char* e = std::getenv("HODLR_ACA_INIT_CAP")) c = std::max(1, std::atoi(e));
    c = std::max(1, std::min(c, upper_max[l]));
    alloc_level(l, c);
  }

  DBuf<AcaBlock> d_blk;
  DBuf<AcaState> d_state;
  DBuf<uint32_t> d_bits;
  DBuf<double> d_fin;
  DBuf<int> d_cnt, d_ranks, d_nchunk;
  d_blk.alloc(nb);
  d_state.alloc(nb);
  d_bits.alloc(words);
  d_fin.alloc((size_t)nb * MAX_FIN_CHUNKS);
  d_cnt.alloc(nb);
  d_ranks.alloc(nb);
  d_nchunk.alloc(nb);
  HODLR_HIP_CHECK(hipMemsetAsync(d_bits.p, 0, d_bits.bytes(), stream_));
  HODLR_HIP_CHECK(hipMemsetAsync(d_cnt.p, 0, d_cnt.bytes(), stream_));
  HODLR_HIP_CHECK(hipMemcpyAsync(d_blk.p, blk.data(), nb * sizeof(AcaBlock), hipMemcpyHostToDevice, stream_));
  aca_launch_init(d_blk.p, d_state.p, d_bits.p, nb, stream_);

  std::vector<int> active(nb);
  for (int b = 0; b < nb; b++) active[b] = b;
  std::vector<int> m_row, m_col, m_fin, m_chunk, nchunk(nb, 1);
  AcaLaunch a{};
  auto build_maps = [&]() {
    m_row.clear();
    m_col.clear();
    m_fin.clear();
    m_chunk.clear();
    for (int b = 0; b < nb; b++) {
      nchunk[b] = std::max(1, (cap[blk[b].level] + 63) / 64);
      if (nchunk[b] > MAX_FIN_CHUNKS) throw std::runtime_error("ACA rank capacity exceeds 16384");
    }
    for (int b : active) {
      blk[b].row_wg0 = (int)m_row.size();
      for (int w = 0; w < blk[b].row_nwg; w++) m_row.push_back(b);
      blk[b].col_wg0 = (int)m_col.size();
      for (int w = 0; w < blk[b].col_nwg; w++) m_col.push_back(b);
      for (int c = 0; c < nchunk[b]; c++) {
        m_fin.push_back(b);
        m_chunk.push_back(c);
      }
    }
    arena_.reset();
    HODLR_HIP_CHECK(hipMemcpyAsync(d_blk.p, blk.data(), nb * sizeof(AcaBlock), hipMemcpyHostToDevice, stream_));
    HODLR_HIP_CHECK(hipMemcpyAsync(d_nchunk.p, nchunk.data(), nb * sizeof(int), hipMemcpyHostToDevice, stream_));
    a.blocks = d_blk.p;
    a.state = d_state.p;
    a.bits = d_bits.p;
    a.wg2blk_row = arena_.push(m_row, stream_);
    a.wg2blk_col = arena_.push(m_col, stream_);
    a.wg2blk_fin = arena_.push(m_fin, stream_);
    a.wg2chunk_fin = arena_.push(m_chunk, stream_);
    a.nchunk = d_nchunk.p;
    a.fin_part = d_fin.p;
    a.counters = d_cnt.p;
    a.n_row_wg = (int)m_row.size();
    a.n_col_wg = (int)m_col.size();
    a.n_fin_wg = (int)m_fin.size();
    a.jpt = jpt;
    a.ipt = ipt;
    a.tol = tol_;
    a.min_rank = -1;  // default tree: topOffDiag_minRank = -1 (HODLR_Tree.cpp:44-47)
    a.max_rank = -1;
    a.min_pivot = min_pivot_;
  };
  build_maps();

  std::vector<AcaState> hs(nb);
  int check = 4;
  // every iteration either accepts a term or consumes an unchosen row, so a
  // block ends within m + n iterations
  long long max_dim = 0;
  for (auto& b : blk) max_dim = std::max<long long>(max_dim, (long long)b.m + b.n);
  long long launched = 0;
  while (true) {
    if (launched > max_dim + 64)
      throw std::runtime_error("ACA did not terminate within m + n iterations");
    for (int c = 0; c < check; c++) aca_launch_iteration(L, a, d_srcs_.p, stream_);
    launched += check;
    HODLR_HIP_CHECK(hipMemcpyAsync(hs.data(), d_state.p, nb * sizeof(AcaState), hipMemcpyDeviceToHost, stream_));
    HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
    std::vector<int> na;
    std::vector<char> grow(nl, 0);
    bool paused = false;
    for (int b : active) {
      if (hs[b].status == 0) na.push_back(b);
      if (hs[b].status == 2) {
        grow[blk[b].level] = 1;
        paused = true;
      }
    }
    if (paused) {
      for (int l = 0; l < nl; l++)
        if (grow[l]) alloc_level(l, std::min(upper_max[l], std::max(cap[l] + 1, 2 * cap[l])));
      for (int b : active)
        if (hs[b].status == 2) {
          hs[b].status = 0;
          na.push_back(b);
        }
      std::sort(na.begin(), na.end());
      HODLR_HIP_CHECK(hipMemcpyAsync(d_state.p, hs.data(), nb * sizeof(AcaState), hipMemcpyHostToDevice, stream_));
    }
    if (na.empty()) break;
    const bool changed = paused || na.size() != active.size();
    active.swap(na);
    if (changed) build_maps();
    check = std::min(16, check * 2);
  }
  int maxit = 0;
  for (int b = 0; b < nb; b++) maxit = std::max(maxit, hs[b].iters);
  stats_.aca_iterations = maxit;
  // dense fallback needs min(m, n) columns (lowRank.cpp:260-274)
  std::vector<int> need(nl, 0);
  for (int b = 0; b < nb; b++) {
    const int up = std::min(blk[b].m, blk[b].n);
    if (hs[b].k >= up - 1) need[blk[b].level] = std::max(need[blk[b].level], up);
  }
  for (int l = 0; l < nl; l++)
    if (need[l] > cap[l]) alloc_level(l, need[l]);
  for (int l = 0; l < nl; l++) P[l].release();
  aca_launch_post(L, d_blk.p, d_state.p, d_srcs_.p, d_ranks.p, nb, stream_);
  std::vector<int> ranks(nb);
  HODLR_HIP_CHECK(hipMemcpyAsync(ranks.data(), d_ranks.p, nb * sizeof(int), hipMemcpyDeviceToHost, stream_));
  HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
  for (int id = 0; id < nn; id++) {
    nodes_[id].kt = ranks[2 * id];
    nodes_[id].kb = ranks[2 * id + 1];
    const int l = nodes_[id].level;
    K_[l] = std::max(K_[l], std::max(nodes_[id].kt, nodes_[id].kb));
  }
  off_[0] = 0;
  for (int l = 0; l < nlev_; l++) off_[l + 1] = off_[l] + K_[l];
  ktot_ = off_[nlev_];
}

void Engine::assemble_leaves() {
  Lorig_.alloc(leaf_elems_);
  std::vector<AssembleDesc> ad;
  int maxe = 0;
  for (auto& lf : leaves_) {
    ad.push_back({Lorig_.p + lf.mat_off, lf.inst, lf.min_i, lf.min_i, lf.s, lf.s, lf.s});
    maxe = std::max(maxe, lf.s * lf.s);
  }
  launch_assemble(arena_.push(ad, stream_), (int)ad.size(), maxe, d_srcs_.p, stream_);
}

/* ------------------------------------------------------------------ */
/* factorization                                                       */
/* ------------------------------------------------------------------ */
void Engine::factor_leaves() {
  if (leaves_factored_) return;
  Lfac_.alloc(leaf_elems_);
  Linv_.alloc(leaf_elems_);
  Lpiv_.alloc(leaf_piv_elems_);
  HODLR_HIP_CHECK(hipMemcpyAsync(Lfac_.p, Lorig_.p, leaf_elems_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
  std::vector<LuDesc> ld;
  std::vector<CopyDesc> id;
  std::vector<SolveDesc> sd;
  int maxs = 1;
  for (auto& lf : leaves_) {
    ld.push_back({Lfac_.p + lf.mat_off, Lpiv_.p + lf.piv_off, lf.s, lf.s});
    id.push_back({nullptr, Linv_.p + lf.mat_off, lf.s, lf.s, 0, lf.s});
    SolveDesc s{};
    s.LU = Lfac_.p + lf.mat_off;
    s.piv = Lpiv_.p + lf.piv_off;
    s.B = Linv_.p + lf.mat_off;
    s.n = lf.s;
    s.ld_lu = lf.s;
    s.ldb = lf.s;
    s.ncols = lf.s;
    sd.push_back(s);
    maxs = std::max(maxs, lf.s);
  }
  lu_batch(ld);
  launch_set_identity(arena_.push(id, stream_), (int)id.size(), stream_);
  lu_solve_batch(sd);
  leaves_factored_ = true;
}

/* recSM_Factorize (HODLR_Matrix.cpp:520-592) and recLU_Factorize (:612-692),
 * level by level from the deepest: leaves of level d and the Woodbury
 * (or Schur) factors of the nodes of level d are applied to the rows that
 * the ancestors' U^SM columns hold. */
void Engine::factorize(int which) {
  if (which != 0 && which != 1) throw std::invalid_argument("unknown factorization");
  compress();
  Factorization& F = fac_[which];
  if (F.done) return;
  HODLR_HIP_CHECK(hipEventRecord(ev0_, stream_));
  factor_leaves();
  const size_t N = (size_t)ntot_;
  F.W.alloc(std::max<size_t>(1, N * ktot_));
  for (int l = 0; l < nlev_; l++)
    if (K_[l] > 0)
      HODLR_HIP_CHECK(hipMemcpyAsync(F.W.p + N * off_[l], U_[l].p, N * K_[l] * sizeof(double),
                                     hipMemcpyDeviceToDevice, stream_));
  const int nn = (int)nodes_.size();
  F.c_off.assign(nn, 0);
  F.p_off.assign(nn, 0);
  size_t co = 0, po = 0;
  for (int i = 0; i < nn; i++) {
    const int K = nodes_[i].kt + nodes_[i].kb;
    F.c_off[i] = co;
    F.p_off[i] = po;
    co += (size_t)K * K;
    po += K;
  }
  F.C.alloc(std::max<size_t>(co, 1));
  F.piv.alloc(std::max<size_t>(po, 1));
  double* W = F.W.p;
  const int ld = ntot_;
  for (int d = nlev_ - 1; d >= 0; d--) {
    const int offd = off_[d];
    // leaves of level d: W[rows, 0:off_d] <- A_leaf^-1 W[rows, 0:off_d]
    if (offd > 0 && !level_leaves_[d].empty()) {
      size_t need = 0;
      for (int li : level_leaves_[d]) need += (size_t)leaves_[li].s * offd;
      grow(tmp_, need);
      std::vector<GemmDesc> g;
      std::vector<CopyDesc> cp;
      size_t to = 0;
      for (int li : level_leaves_[d]) {
        const LeafH& lf = leaves_[li];
        g.push_back(gemm_desc(lf.s, offd, lf.s, 0, Linv_.p + lf.mat_off, lf.s, W + lf.row0, ld,
                              tmp_.p + to, lf.s, 1.0, 0.0));
        cp.push_back({tmp_.p + to, W + lf.row0, lf.s, offd, lf.s, ld});
        to += (size_t)lf.s * offd;
      }
      run_gemm(g);
      launch_copy(arena_.push(cp, stream_), (int)cp.size(), stream_);
    }
    if (level_nodes_[d].empty()) continue;
    // nodes of level d
    const int ncols = offd + K_[d];
    size_t zneed = 0, yneed = 0;
    for (int id : level_nodes_[d]) {
      const int K = nodes_[id].kt + nodes_[id].kb;
      zneed += (size_t)K * ncols;
      yneed += (size_t)K * std::max(offd, 1);
    }
    grow(tmp2_, zneed);
    grow(tmp_, yneed);
    std::vector<GemmDesc> gz, gu;
    std::vector<CapDesc> cd;
    std::vector<LuDesc> lu;
    std::vector<SolveDesc> sd;
    size_t zo = 0, yo = 0;
    int maxK = 1;
    for (int id : level_nodes_[d]) {
      const TreeNodeH& nd = nodes_[id];
      const int kt = nd.kt, kb = nd.kb, K = kt + kb;
      const int st = nd.split - nd.min_i + 1, sb = nd.max_i - nd.split;
      const int top = nd.row0, bot = nd.row0 + st;
      double* Zt = tmp2_.p + zo;
      double* Zb = Zt + (size_t)kb * ncols;
      zo += (size_t)K * ncols;
      // Zt = V_b^T [W_top], Zb = V_t^T [W_bot]
      gz.push_back(gemm_desc(kb, ncols, st, 1, V_[d].p + top, ld, W + top, ld, Zt, kb, 1.0, 0.0));
      gz.push_back(gemm_desc(kt, ncols, sb, 1, V_[d].p + bot, ld, W + bot, ld, Zb, kt, 1.0, 0.0));
      double* Cn = F.C.p + F.c_off[id];
      double* Y = tmp_.p + yo;
      yo += (size_t)K * std::max(offd, 1);
      CapDesc c{};
      c.Zt = Zt;
      c.Zb = Zb;
      c.C = Cn;
      c.Y = Y;
      c.kt = kt;
      c.kb = kb;
      c.nanc = offd;
      c.own0 = offd;
      c.lu_order = which;
      cd.push_back(c);
      lu.push_back({Cn, F.piv.p + F.p_off[id], K, K});
      maxK = std::max(maxK, K);
      if (offd > 0) {
        SolveDesc s{};
        s.LU = Cn;
        s.piv = F.piv.p + F.p_off[id];
        s.B = Y;
        s.n = K;
        s.ld_lu = K;
        s.ldb = K;
        s.ncols = offd;
        sd.push_back(s);
        const int qt = which == 0 ? 0 : kb;  // rows of Q multiplying U_t^SM
        const int qb = which == 0 ? kt : 0;  // rows of Q multiplying U_b^SM
        gu.push_back(gemm_desc(st, offd, kt, 0, W + top + (size_t)offd * ld, ld, Y + qt, K, W + top,
                               ld, -1.0, 1.0));
        gu.push_back(gemm_desc(sb, offd, kb, 0, W + bot + (size_t)offd * ld, ld, Y + qb, K, W + bot,
                               ld, -1.0, 1.0));
      }
    }
    run_gemm(gz);
    launch_capacitance(arena_.push(cd, stream_), (int)cd.size(), stream_);
    lu_batch(lu);
    if (offd > 0) {
      lu_solve_batch(sd);
      run_gemm(gu);
    }
  }
  HODLR_HIP_CHECK(hipEventRecord(ev1_, stream_));
  HODLR_HIP_CHECK(hipEventSynchronize(ev1_));
  float ms = 0;
  HODLR_HIP_CHECK(hipEventElapsedTime(&ms, ev0_, ev1_));
  stats_.factor_time = ms * 1e-3;
  arena_.reset();
  F.done = true;
}

/* ------------------------------------------------------------------ */
/* solve: recSM_Solve (HODLR_Matrix.cpp:594-610) / recLU_Solve (:694-755) */
/* ------------------------------------------------------------------ */
void Engine::solve_device(int which, const double* dB, int ldb, int nrhs, double* dX, int ldx) {
  if (nrhs <= 0) return;
  if (ldb < ntot_ || ldx < ntot_) throw std::invalid_argument("leading dimension < n");
  factorize(which);
  Factorization& F = fac_[which];
  HODLR_HIP_CHECK(hipEventRecord(ev0_, stream_));
  const int ld = ntot_;
  const size_t N = (size_t)ntot_;
  grow(xw_, N * nrhs);
  double* X = xw_.p;
  HODLR_HIP_CHECK(hipMemcpy2DAsync(X, N * sizeof(double), dB, (size_t)ldb * sizeof(double),
                                   N * sizeof(double), nrhs, hipMemcpyDeviceToDevice, stream_));
  // leaves: X[rows] <- A_leaf^-1 X[rows]
  {
    grow(tmp_, N * nrhs);
    std::vector<GemmDesc> g;
    std::vector<CopyDesc> cp;
    size_t to = 0;
    for (auto& lf : leaves_) {
      g.push_back(gemm_desc(lf.s, nrhs, lf.s, 0, Linv_.p + lf.mat_off, lf.s, X + lf.row0, ld,
                            tmp_.p + to, lf.s, 1.0, 0.0));
      cp.push_back({tmp_.p + to, X + lf.row0, lf.s, nrhs, lf.s, ld});
      to += (size_t)lf.s * nrhs;
    }
    run_gemm(g);
    launch_copy(arena_.push(cp, stream_), (int)cp.size(), stream_);
  }
  for (int d = nlev_ - 1; d >= 0; d--) {
    if (level_nodes_[d].empty()) continue;
    const int offd = off_[d];
    size_t zneed = 0;
    for (int id : level_nodes_[d]) zneed += (size_t)(nodes_[id].kt + nodes_[id].kb) * nrhs;
    grow(tmp2_, zneed);
    grow(tmp_, zneed);
    std::vector<GemmDesc> gz, gu;
    std::vector<CapDesc> cd;
    std::vector<SolveDesc> sd;
    size_t zo = 0;
    int maxK = 1;
    for (int id : level_nodes_[d]) {
      const TreeNodeH& nd = nodes_[id];
      const int kt = nd.kt, kb = nd.kb, K = kt + kb;
      const int st = nd.split - nd.min_i + 1, sb = nd.max_i - nd.split;
      const int top = nd.row0, bot = nd.row0 + st;
      double* Zt = tmp2_.p + zo;
      double* Zb = Zt + (size_t)kb * nrhs;
      double* Y = tmp_.p + zo;
      zo += (size_t)K * nrhs;
      gz.push_back(gemm_desc(kb, nrhs, st, 1, V_[d].p + top, ld, X + top, ld, Zt, kb, 1.0, 0.0));
      gz.push_back(gemm_desc(kt, nrhs, sb, 1, V_[d].p + bot, ld, X + bot, ld, Zb, kt, 1.0, 0.0));
      CapDesc c{};
      c.Zt = Zt;
      c.Zb = Zb;
      c.C = nullptr;
      c.Y = Y;
      c.kt = kt;
      c.kb = kb;
      c.nanc = nrhs;
      c.own0 = 0;
      c.lu_order = which;
      cd.push_back(c);
      SolveDesc s{};
      s.LU = F.C.p + F.c_off[id];
      s.piv = F.piv.p + F.p_off[id];
      s.B = Y;
      s.n = K;
      s.ld_lu = K;
      s.ldb = K;
      s.ncols = nrhs;
      sd.push_back(s);
      maxK = std::max(maxK, K);
      const int qt = which == 0 ? 0 : kb;
      const int qb = which == 0 ? kt : 0;
      gu.push_back(gemm_desc(st, nrhs, kt, 0, F.W.p + top + (size_t)offd * ld, ld, Y + qt, K,
                             X + top, ld, -1.0, 1.0));
      gu.push_back(gemm_desc(sb, nrhs, kb, 0, F.W.p + bot + (size_t)offd * ld, ld, Y + qb, K,
                             X + bot, ld, -1.0, 1.0));
    }
    run_gemm(gz);
    launch_capacitance(arena_.push(cd, stream_), (int)cd.size(), stream_);
    lu_solve_batch(sd);
    run_gemm(gu);
  }
  HODLR_HIP_CHECK(hipMemcpy2DAsync(dX, (size_t)ldx * sizeof(double), X, N * sizeof(double),
                                   N * sizeof(double), nrhs, hipMemcpyDeviceToDevice, stream_));
  HODLR_HIP_CHECK(hipEventRecord(ev1_, stream_));
  HODLR_HIP_CHECK(hipEventSynchronize(ev1_));
  float ms = 0;
  HODLR_HIP_CHECK(hipEventElapsedTime(&ms, ev0_, ev1_));
  stats_.solve_time = ms * 1e-3;
  arena_.reset();
}

void Engine::solve_host(int which, const double* B, int ldb, int nrhs, double* X, int ldx) {
  if (nrhs <= 0) return;
  if (ldb < ntot_ || ldx < ntot_) throw std::invalid_argument("leading dimension < n");
  const size_t N = (size_t)ntot_;
  DBuf<double> dB, dX;
  dB.alloc(N * nrhs);
  dX.alloc(N * nrhs);
  HODLR_HIP_CHECK(hipMemcpy2D(dB.p, N * sizeof(double), B, (size_t)ldb * sizeof(double),
                              N * sizeof(double), nrhs, hipMemcpyHostToDevice));
  solve_device(which, dB.p, (int)N, nrhs, dX.p, (int)N);
  HODLR_HIP_CHECK(hipMemcpy2D(X, (size_t)ldx * sizeof(double), dX.p, N * sizeof(double),
                              N * sizeof(double), nrhs, hipMemcpyDeviceToHost));
}

/* calcDeterminant (HODLR_Matrix.cpp:1671-1697): sum of log|diag LU| over the
 * leaves and the recSM capacitance matrices (perturbI.cpp:61-73). */
double Engine::logabsdet(double* det_prod) {
  factorize(0);
  Factorization& F = fac_[0];
  std::vector<LuDesc> ld;
  for (auto& lf : leaves_) ld.push_back({Lfac_.p + lf.mat_off, nullptr, lf.s, lf.s});
  for (int i = 0; i < (int)nodes_.size(); i++) {
    const int K = nodes_[i].kt + nodes_[i].kb;
    ld.push_back({F.C.p + F.c_off[i], nullptr, K, K});
  }
  DBuf<double> lg, pr;
  lg.alloc(ld.size());
  pr.alloc(ld.size());
  launch_logdet(arena_.push(ld, stream_), (int)ld.size(), lg.p, pr.p, stream_);
  std::vector<double> hl(ld.size()), hp(ld.size());
  HODLR_HIP_CHECK(hipMemcpyAsync(hl.data(), lg.p, hl.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
  HODLR_HIP_CHECK(hipMemcpyAsync(hp.data(), pr.p, hp.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
  HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
  arena_.reset();
  double s = 0, p = 1;
  for (size_t i = 0; i < hl.size(); i++) {
    s += hl[i];
    p *= hp[i];
  }
  if (det_prod) *det_prod = p;
  return s;
}

/* Y = A_H X: leaves plus, per node, U_t (V_t^T X_bot) and U_b (V_b^T X_top). */
void Engine::matvec_device(const double* dX, int ldx, int nrhs, double* dY, int ldy) {
  if (nrhs <= 0) return;
  compress();
  const int ld = ntot_;
  std::vector<GemmDesc> g;
  for (auto& lf : leaves_)
    g.push_back(gemm_desc(lf.s, nrhs, lf.s, 0, Lorig_.p + lf.mat_off, lf.s, dX + lf.row0, ldx,
                          dY + lf.row0, ldy, 1.0, 0.0));
  run_gemm(g);
  for (int d = 0; d < nlev_; d++) {
    if (level_nodes_[d].empty()) continue;
    size_t need = 0;
    for (int id : level_nodes_[d]) need += (size_t)(nodes_[id].kt + nodes_[id].kb) * nrhs;
    grow(tmp2_, need);
    std::vector<GemmDesc> g1, g2;
    size_t o = 0;
    for (int id : level_nodes_[d]) {
      const TreeNodeH& nd = nodes_[id];
      const int kt = nd.kt, kb = nd.kb;
      const int st = nd.split - nd.min_i + 1, sb = nd.max_i - nd.split;
      const int top = nd.row0, bot = nd.row0 + st;
      double* tt = tmp2_.p + o;
      double* tb = tt + (size_t)kt * nrhs;
      o += (size_t)(kt + kb) * nrhs;
      g1.push_back(gemm_desc(kt, nrhs, sb, 1, V_[d].p + bot, ld, dX + bot, ldx, tt, kt, 1.0, 0.0));
      g1.push_back(gemm_desc(kb, nrhs, st, 1, V_[d].p + top, ld, dX + top, ldx, tb, kb, 1.0, 0.0));
      g2.push_back(gemm_desc(st, nrhs, kt, 0, U_[d].p + top, ld, tt, kt, dY + top, ldy, 1.0, 1.0));
      g2.push_back(gemm_desc(sb, nrhs, kb, 0, U_[d].p + bot, ld, tb, kb, dY + bot, ldy, 1.0, 1.0));
    }
    run_gemm(g1);
    run_gemm(g2);
  }
  HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
  arena_.reset();
}

void Engine::get_factor(int idx, int which, double* host_out) {
  compress();
  if (idx < 0 || idx >= (int)nodes_.size() || which < 0 || which > 3)
    throw std::invalid_argument("bad factor index");
  const TreeNodeH& nd = nodes_[idx];
  const int st = nd.split - nd.min_i + 1, sb = nd.max_i - nd.split;
  const int top = nd.row0, bot = nd.row0 + st;
  const int d = nd.level;
  const double* src;
  int rows, rank;
  switch (which) {
    case 0: src = U_[d].p + top; rows = st; rank = nd.kt; break;
    case 1: src = V_[d].p + bot; rows = sb; rank = nd.kt; break;
    case 2: src = U_[d].p + bot; rows = sb; rank = nd.kb; break;
    default: src = V_[d].p + top; rows = st; rank = nd.kb; break;
  }
  HODLR_HIP_CHECK(hipMemcpy2D(host_out, (size_t)rows * sizeof(double), src,
                              (size_t)ntot_ * sizeof(double), (size_t)rows * sizeof(double), rank,
                              hipMemcpyDeviceToHost));
}

template double* DescArena::push<double>(const std::vector<double>&, hipStream_t);

}  // namespace hodlr
