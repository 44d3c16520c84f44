/*
 * Host orchestration of the HODLR engine (see engine.h for the layout).
 */
#include "engine.h"
#include "mlp.h"
#include "svd.h"

#include <algorithm>
#include <map>
#include <mutex>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <cstring>

namespace hodlr {

/* ------------------------------------------------------------------ */
/* DescArena                                                           */
/* ------------------------------------------------------------------ */
void DescArena::reset() {
  for (auto& c : chunks_) c.used = 0;
  while (chunks_.size() > 1) chunks_.pop_front();
  keep_.clear();
  dev_.clear();
  replay_ = false;
}

size_t DescArena::bytes() const {
  size_t b = 0;
  for (auto& c : chunks_) b += c.buf.bytes();
  return b;
}

/* ------------------------------------------------------------------ */
/* construction                                                        */
/* ------------------------------------------------------------------ */
namespace {
std::mutex g_tune_mu;
std::map<std::string, Tuning> g_tune;  // autotuned shapes by problem shape
}  // namespace

Engine::Engine(std::vector<InstanceSpec> specs) : specs_(std::move(specs)) {
  trace_ = std::getenv("HODLR_TRACE") != nullptr;
  // largest system the one-launch LU solve takes (lowered to test the
  // blocked path on small problems)
  if (const char* e = std::getenv("HODLR_LU_SOLVE_SMALL_MAX"))
    solve_small_max_ = std::max(0, std::min(LU_SOLVE_SMALL_MAX, std::atoi(e)));
  // and the one-wave-per-system kernel (0: the workgroup kernel for all)
  if (const char* e = std::getenv("HODLR_LU_SOLVE_WAVE_MAX"))
    solve_wave_max_ = std::max(0, std::min(LU_SOLVE_WAVE_MAX, std::atoi(e)));
  // and the blocked one-launch solve above 512 rows (off: step by step)
  solve_blk_ = std::getenv("HODLR_NO_LU_SOLVE_BLOCKS") == nullptr;
  HODLR_HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  HODLR_HIP_CHECK(hipEventCreate(&ev0_));
  HODLR_HIP_CHECK(hipEventCreate(&ev1_));
  HODLR_HIP_CHECK(hipStreamCreateWithFlags(&aux_, hipStreamNonBlocking));
  HODLR_HIP_CHECK(hipEventCreateWithFlags(&ev_cap_, hipEventDisableTiming));
  HODLR_HIP_CHECK(hipEventCreateWithFlags(&ev_aca_, hipEventDisableTiming));
  HODLR_HIP_CHECK(hipEventCreateWithFlags(&ev_w0_, hipEventDisableTiming));
  HODLR_HIP_CHECK(hipEventCreateWithFlags(&ev_w1_, hipEventDisableTiming));
  HODLR_HIP_CHECK(hipEventCreateWithFlags(&ev_root_, hipEventDisableTiming));
  HODLR_HIP_CHECK(hipEventCreate(&ev_f0_));
  HODLR_HIP_CHECK(hipEventCreate(&ev_f1_));
  HODLR_HIP_CHECK(hipEventCreate(&ev_froot_));
  // (off: the roots' LU on the engine's stream, as every other level's)
  overlap_root_ = std::getenv("HODLR_NO_ROOT_OVERLAP") == nullptr;
  adaptive_aca_pt_ = std::getenv("HODLR_FIXED_ACA_PT") == nullptr;
  // (off: the factorization's small-rank updates through the tiled GEMM)
  stream_update_ = std::getenv("HODLR_NO_STREAM_UPDATE") == nullptr;
  build_trees();
  upload_sources();
  noalloc_setup();  // the allocation-free product's scratch address on this device
  // problem shape: instance sizes, leaf sizes and kinds
  shape_key_ = std::to_string(specs_.size());
  for (auto& sp : specs_)
    shape_key_ += ":" + std::to_string(sp.n) + "/" + std::to_string(sp.leaf) + "/" +
                  std::to_string(sp.kind);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(shape_key_);
  if (it != g_tune.end()) tune_ = it->second;
}

void Engine::set_tuning(const Tuning& t) {
  if (t.aca_pt != 1 && t.aca_pt != 2 && t.aca_pt != 4)
    throw std::invalid_argument("ACA rows per thread must be 1, 2 or 4");
  if (t.gemm_bm != 0 && t.gemm_bm != 64 && t.gemm_bm != 128)
    throw std::invalid_argument("GEMM tile rows must be 0, 64 or 128");
  if (t.gemm_target < 1) throw std::invalid_argument("split-K target must be positive");
  if (t.lu_nb < 1 || t.lu_nb > LU_NB) throw std::invalid_argument("LU panel width must be 1..64");
  if (t.gemm_mfma != 0 && t.gemm_mfma != 1) throw std::invalid_argument("mfma must be 0 or 1");
  if (t.leaf_solve < 0 || t.leaf_solve > 4)
    throw std::invalid_argument("the leaf solve variant must be 0 .. 4");
  tune_ = t;
  tune_.tuned = false;
  reset_attributes();
}

void Engine::autotune() {
  Tuning best = tune_;
  auto time_compress = [&](const Tuning& t) {
    tune_ = t;
    reset_attributes();
    compress();
    return stats_.compress_time;
  };
  auto time_factor = [&](const Tuning& t) {
    tune_ = t;
    finish_root();
    for (auto& f : fac_) f.done = false;
    leaves_factored_ = false;
    factorize_fp64(0);
    sync_stats();  // the full factorization, the roots' LU on the side stream included
    return stats_.factor_time;
  };
  double bt = 1e300;
  for (int pt : {1, 2, 4}) {
    Tuning t = best;
    t.aca_pt = pt;
    const double v = std::min(time_compress(t), time_compress(t));
    if (v < bt) {
      bt = v;
      best = t;
    }
  }
  time_compress(best);
  bt = 1e300;
  Tuning fb = best;
  // {tile rows (0: per launch), split-K target, matrix cores}
  const int shapes[][3] = {{0, 2048, 0}, {0, 1024, 0}, {0, 4096, 0}, {64, 2048, 0},
                           {128, 2048, 0}, {128, 2048, 1}, {128, 4096, 1}, {0, 4096, 1}};
  for (int nb : {64, 32})
    for (auto& sh : shapes) {
      Tuning t = best;
      t.lu_nb = nb;
      t.gemm_bm = sh[0];
      t.gemm_target = sh[1];
      t.gemm_mfma = sh[2];
      const double v = std::min(time_factor(t), time_factor(t));
      if (v < bt) {
        bt = v;
        fb = t;
      }
    }
  {  // the leaf solves of the factorization: one wave, blocked, shared factor, matrix cores
    const Tuning base = fb;
    for (int ls : {0, 1, 2, 3, 4}) {
      if (ls == base.leaf_solve) continue;
      Tuning t = base;
      t.leaf_solve = ls;
      const double v = std::min(time_factor(t), time_factor(t));
      if (v < bt) {
        bt = v;
        fb = t;
      }
    }
  }
  fb.tuned = true;
  tune_ = fb;
  reset_attributes();
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[shape_key_] = fb;
}

Engine::~Engine() {
  if (aux_) (void)hipStreamSynchronize(aux_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (auto e : prof_ev_) (void)hipEventDestroy(e);
  for (auto e : span_events_) (void)hipEventDestroy(e);
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
  if (ev_cap_) (void)hipEventDestroy(ev_cap_);
  if (ev_aca_) (void)hipEventDestroy(ev_aca_);
  if (ev_w0_) (void)hipEventDestroy(ev_w0_);
  if (ev_w1_) (void)hipEventDestroy(ev_w1_);
  if (ev_root_) (void)hipEventDestroy(ev_root_);
  if (ev_f0_) (void)hipEventDestroy(ev_f0_);
  if (ev_f1_) (void)hipEventDestroy(ev_f1_);
  if (ev_froot_) (void)hipEventDestroy(ev_froot_);
  if (aux_) (void)hipStreamDestroy(aux_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void Engine::wait_root(Factorization& F) {
  if (F.root_pending) HODLR_HIP_CHECK(hipStreamWaitEvent(stream_, ev_root_, 0));
}

void Engine::sync_stats() {
  if (!factor_time_pending_) return;
  HODLR_HIP_CHECK(hipEventSynchronize(ev_froot_));
  float ms = 0;
  HODLR_HIP_CHECK(hipEventElapsedTime(&ms, ev_f0_, ev_froot_));
  stats_.factor_time = std::max(stats_.factor_stream_time, ms * 1e-3);
  factor_time_pending_ = false;
}

void Engine::finish_root() {
  bool any = false;
  for (auto& f : fac_) any = any || f.root_pending;
  if (!any) return;
  HODLR_HIP_CHECK(hipEventSynchronize(ev_root_));
  for (auto& f : fac_) f.root_pending = false;
  aux_arena_.reset();
}

/* Tree of each instance.  Default (src/HODLR_Tree.cpp:33-132): the root
 * splits at n/2-1; a left child of size s splits at min + s/2 - 1, a right
 * child at max - s/2; the root is a leaf if n < leaf, a child if s <= leaf.
 * User partition (userTree_To_HODLRTree, HODLR_Tree.cpp:207-321): a user node
 * with split -1 and at most leaf rows is a leaf, any other takes its split,
 * rank bounds and method; a user node without children gets default nodes
 * below it, the left one split at min + s/2 - 1 and the right one at
 * split + s/2 (not max - s/2: the reference's rule there), each with a
 * default subtree. */
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
    auto add_leaf = [&](int level, int lo, int hi) {
      lv.push_back({q, level, lo, hi, row0 + lo, hi - lo + 1, 0, 0});
      nlev_ = std::max(nlev_, level + 1);
    };
    auto add_node = [&](int level, int lo, int split, int hi) -> TreeNodeH& {
      TreeNodeH nd;
      nd.inst = q;
      nd.level = level;
      nd.min_i = lo;
      nd.split = split;
      nd.max_i = hi;
      nd.row0 = row0 + lo;
      all.push_back(nd);
      nlev_ = std::max(nlev_, level + 2);
      return all.back();
    };
    std::vector<Rec> stack;  // non-leaf nodes whose children follow the default rules
    auto default_below = [&]() {
      while (!stack.empty()) {
        Rec r = stack.back();
        stack.pop_back();
        add_node(r.level, r.lo, r.sp, r.hi);
        const int ts = r.sp - r.lo + 1, bs = r.hi - r.sp;
        if (ts > leaf)
          stack.push_back({r.level + 1, r.lo, r.lo + ts / 2 - 1, r.sp});
        else
          add_leaf(r.level + 1, r.lo, r.sp);
        if (bs > leaf)
          stack.push_back({r.level + 1, r.sp + 1, r.hi - bs / 2, r.hi});
        else
          add_leaf(r.level + 1, r.sp + 1, r.hi);
      }
    };
    if (sp.tree.empty()) {
      if (n < leaf) {
        add_leaf(0, 0, n - 1);
        continue;
      }
      stack.push_back({0, 0, n / 2 - 1, n - 1});
      default_below();
      continue;
    }
    // user partition: depth-first over the user nodes
    struct URec {
      int u, level, lo, hi;
    };
    std::vector<URec> us{{0, 0, 0, n - 1}};
    size_t visited = 0;
    while (!us.empty()) {
      const URec r = us.back();
      us.pop_back();
      if (r.u < 0 || r.u >= (int)sp.tree.size() || ++visited > sp.tree.size())
        throw std::invalid_argument("user tree: bad child index or a cycle");
      const UserNode& u = sp.tree[r.u];
      const int size = r.hi - r.lo + 1;
      if (size <= leaf && u.split == -1) {
        add_leaf(r.level, r.lo, r.hi);
        continue;
      }
      if (u.split == -1)
        throw std::invalid_argument("user tree: a node of " + std::to_string(size) +
                                    " rows (more than the leaf size) has no split");
      if (u.split < r.lo || u.split >= r.hi)
        throw std::invalid_argument("user tree: split " + std::to_string(u.split) +
                                    " outside its node [" + std::to_string(r.lo) + ", " +
                                    std::to_string(r.hi) + "]");
      if (u.lr < -1 || u.lr > 1) throw std::invalid_argument("user tree: unknown low-rank method");
      TreeNodeH& nd = add_node(r.level, r.lo, u.split, r.hi);
      nd.lr = u.lr;
      for (int h = 0; h < 2; h++) {
        nd.min_rank[h] = u.min_rank[h];
        nd.max_rank[h] = u.max_rank[h];
      }
      if (u.left < 0 && u.right < 0) {
        const int ls = u.split - r.lo + 1, rs = r.hi - u.split;
        if (ls <= leaf)
          add_leaf(r.level + 1, r.lo, u.split);
        else
          stack.push_back({r.level + 1, r.lo, r.lo + ls / 2 - 1, u.split});
        if (rs <= leaf)
          add_leaf(r.level + 1, u.split + 1, r.hi);
        else
          stack.push_back({r.level + 1, u.split + 1, u.split + rs / 2, r.hi});
        default_below();
      } else if (u.left >= 0 && u.right >= 0) {
        us.push_back({u.right, r.level + 1, u.split + 1, r.hi});
        us.push_back({u.left, r.level + 1, r.lo, u.split});
      } else {
        throw std::invalid_argument("user tree: a node has one child");
      }
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
  finish_root();
  buf_gen_++;
  graph_.release();
  compressed_ = false;
  leaves_factored_ = false;
  for (auto& f : fac_) {
    f.done = false;
    f.W.release();
    f.C.release();
    f.piv.release();
    f.perm.release();
  }
  for (auto& f : mfac_) f = MixedFactorization();
  U_.clear();
  V_.clear();
  stats_ = Stats();
  factor_time_pending_ = false;
}

/* set_LRMethod (HODLR_Matrix.cpp:1277-1286, HODLR_Tree.cpp:377-390): a
 * method other than the handle's default becomes every node's method (the
 * per-node choices of a user partition included); the default itself is a
 * no-op, as in the reference. */
void Engine::set_lr_method(int m) {
  if (m != 0 && m != 1) throw std::invalid_argument("unknown low-rank method");
  if (m != lr_method_) {
    lr_method_ = m;
    for (auto& nd : nodes_) nd.lr = -1;
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
  if (p < 0 || p > 2) throw std::invalid_argument("unknown precision");
  if (p != precision_) {
    finish_root();  // the roots' LU may still write F.C and F.piv on the side stream
    precision_ = p;
    for (auto& f : fac_) {
      f.done = false;
      f.W.release();
      f.C.release();
      f.piv.release();
      f.perm.release();
    }
    for (auto& f : mfac_) f = MixedFactorization();
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
  b += Lorig_.bytes() + Lfac_.bytes() + Lpiv_.bytes() + Lperm_.bytes();
  for (auto& f : fac_) b += f.W.bytes() + f.C.bytes() + f.piv.bytes() + f.perm.bytes();
  b += ws_.bytes() + tmp_.bytes() + tmp2_.bytes() + xw_.bytes();
  for (auto& f : mfac_) {
    b += f.W.bytes() + f.C.bytes() + f.piv.bytes();
    for (auto& x : f.V) b += x.bytes();
  }
  b += t32a_.bytes() + t32b_.bytes() + x32_.bytes();
  b += rbuf_.bytes() + ybuf_.bytes() + zbuf_.bytes() + nbuf_.bytes() + basis_.bytes() +
       zbasis_.bytes() + ddy_.bytes() + ddpart_.bytes() + ddt_.bytes();
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
    buf_gen_++;
  }
}

void Engine::grow(DBuf<double>& b, size_t doubles) {
  if (b.n < doubles) {
    HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
    b.alloc(doubles);
    buf_gen_++;
  }
}

/* ------------------------------------------------------------------ */
/* profiling of the factorization's kernel groups                      */
/* ------------------------------------------------------------------ */
namespace {
constexpr double FP64_PEAK = 78.6e12;  // MI355X dense FP64, vector and matrix cores
constexpr double HBM_PEAK = 8.0e12;

template <class T>
void gemm_work(const std::vector<GemmDescT<T>>& d, double& fl, double& by) {
  fl = by = 0;
  for (auto& g : d) {
    fl += 2.0 * g.M * (double)g.N * g.K;
    by += sizeof(T) * ((double)g.M * g.K + (double)g.K * g.N +
                       (double)g.M * g.N * (g.beta != T(0) ? 2 : 1));
  }
}

void lu_work(const std::vector<LuDesc>& d, double& fl, double& by) {
  fl = by = 0;
  for (auto& l : d) {
    fl += 2.0 / 3.0 * (double)l.n * l.n * l.n;
    by += 16.0 * (double)l.n * l.n;
  }
}
}  // namespace

int Engine::prof_begin(hipStream_t s) {
  if (!prof_fac_) return -1;
  if (prof_depth_++ > 0) return -1;  // outermost group only
  while (span_next_ + 2 > span_events_.size()) {
    hipEvent_t e;
    HODLR_HIP_CHECK(hipEventCreate(&e));
    span_events_.push_back(e);
  }
  const int tok = (int)span_next_;
  span_next_ += 2;
  HODLR_HIP_CHECK(hipEventRecord(span_events_[tok], s));
  return tok;
}

void Engine::prof_end(int tok, hipStream_t s, int cat, double flops, double bytes) {
  if (!prof_fac_) return;
  prof_depth_--;
  if (tok < 0) return;
  HODLR_HIP_CHECK(hipEventRecord(span_events_[tok + 1], s));
  spans_.push_back({cat, flops, bytes, span_events_[tok], span_events_[tok + 1]});
}

/* after both streams are idle */
void Engine::prof_collect() {
  for (int c = 0; c < 4; c++) {
    stats_.fac_time[c] = stats_.fac_flops[c] = stats_.fac_bytes[c] = stats_.fac_bound[c] = 0;
    stats_.fac_launches[c] = 0;
  }
  for (auto& sp : spans_) {
    float ms = 0;
    HODLR_HIP_CHECK(hipEventElapsedTime(&ms, sp.a, sp.b));
    stats_.fac_time[sp.cat] += ms * 1e-3;
    stats_.fac_flops[sp.cat] += sp.flops;
    stats_.fac_bytes[sp.cat] += sp.bytes;
    stats_.fac_bound[sp.cat] += std::max(sp.flops / FP64_PEAK, sp.bytes / HBM_PEAK);
    stats_.fac_launches[sp.cat]++;
  }
  spans_.clear();
  span_next_ = 0;
}

template <class T>
void Engine::run_gemm_t(std::vector<GemmDescT<T>>& descs) {
  if (descs.empty()) return;
  if constexpr (std::is_same<T, double>::value) {
    // the low-rank updates of small rank: streamed on the matrix cores
    // (bit-identical to the tiled kernel), the rest tiled
    if (tune_.gemm_mfma && stream_update_) {
      std::vector<GemmDesc> up, rest;
      int maxk = 0;
      for (auto& g : descs) {
        if (gemm_update_eligible(g)) {
          up.push_back(g);
          maxk = std::max(maxk, g.K);
        } else {
          rest.push_back(g);
        }
      }
      const int nwg = up.empty() ? 0 : gemm_update_plan(up, gemm_update_rows(maxk));
      if (nwg >= 1024) {  // (fewer: the tiled kernel's shorter tiles fill the GPU better)
        launch_gemm_update(arena_.push(up, stream_), (int)up.size(), nwg, maxk, stream_);
        if (!rest.empty()) run_gemm_tiled(rest);
        return;
      }
    }
  }
  run_gemm_tiled(descs);
}

template <class T>
void Engine::run_gemm_tiled(std::vector<GemmDescT<T>>& descs) {
  std::vector<int> split_ids;
  int ntiles = 0;
  const bool mfma = tune_.gemm_mfma != 0;
  const int bm = tune_.gemm_bm ? tune_.gemm_bm : gemm_tile_m(descs, mfma);
  const size_t ws = gemm_plan(descs, split_ids, ntiles, tune_.gemm_target, bm, mfma);
  if (ws) ensure_ws((ws * sizeof(T) + 7) / 8);
  const GemmDescT<T>* d = arena_.push(descs, stream_);
  const int* ids = split_ids.empty() ? nullptr : arena_.push(split_ids, stream_);
  launch_gemm<T>(d, (int)descs.size(), ntiles, reinterpret_cast<T*>(ws_.p), ids,
                 (int)split_ids.size(), bm, stream_, mfma);
}

void Engine::run_gemm(std::vector<GemmDesc>& descs) { run_gemm_t<double>(descs); }

template <class T>
static GemmDescT<T> gemm_desc(int M, int N, int K, int ta, const T* A, int lda, const T* B, int ldb,
                              T* C, int ldc, T alpha, T beta) {
  GemmDescT<T> g{};
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
 * (getrf semantics): per panel of LU_NB columns, the panel factorization,
 * the step kernel (the panel's row interchanges in the columns outside it,
 * the triangular solve for U12) and the trailing update A22 -= L21 U12 as
 * one grouped GEMM over the batch. */
void Engine::lu_batch(std::vector<LuDesc>& descs) {
  finish_root();  // (the side-stream LU shares the interchange scratch)
  lu_batch(descs, stream_, arena_);
}

void Engine::lu_batch(std::vector<LuDesc>& descs, hipStream_t s, DescArena& arena) {
  if (descs.empty()) return;
  int maxn = 0;
  for (auto& d : descs) maxn = std::max(maxn, d.n);
  const int nb = lu_panel_width(maxn, tune_.lu_nb);
  // the register panels record their interchanges for the step kernel
  const bool defer = lu_panel_defers_swaps(maxn, nb);
  grow_t(lu_swp_, descs.size() * LU_SWP_INTS);
  for (size_t i = 0; i < descs.size(); i++) descs[i].swp = lu_swp_.p + i * LU_SWP_INTS;
  const LuDesc* dd = arena.push(descs, s);
  // Every panel step's TRSM and trailing-update descriptors are planned on
  // the host first and uploaded together: one copy per batch instead of two
  // or three per panel (a 1090-row matrix has 137 panels of 8).
  struct Step {
    int tr0, ntr, nwg;            // step descriptors [tr0, tr0 + ntr)
    int g0, ng, ntiles, bm;       // GEMM descriptors [g0, g0 + ng)
    int id0, nid;                 // split-K problem ids [id0, id0 + nid)
  };
  std::vector<Step> steps;
  std::vector<LuStepDesc> tr_all;
  std::vector<GemmDesc> g_all;
  std::vector<int> id_all;
  size_t ws = 0;
  for (int p0 = 0; p0 < maxn; p0 += nb) {
    Step st{(int)tr_all.size(), 0, 0, (int)g_all.size(), 0, 0, 0, (int)id_all.size(), 0};
    std::vector<GemmDesc> g;
    for (auto& d : descs) {
      const int pw = std::min(nb, d.n - p0);
      const int rest = d.n - p0 - pw;
      if (pw <= 0) continue;
      const int nleft = defer ? (p0 + 63) / 64 : 0, nright = (rest + 63) / 64;
      if (nleft + nright > 0) {
        LuStepDesc t{};
        t.A = d.A;
        t.swp = d.swp;
        t.n = d.n;
        t.ld = d.ld;
        t.p0 = p0;
        t.pw = pw;
        t.chunk0 = st.nwg;
        t.nleft = nleft;
        st.nwg += nleft + nright;
        tr_all.push_back(t);
      }
      if (rest <= 0) continue;
      g.push_back(gemm_desc(rest, rest, pw, 0, d.A + (p0 + pw) + (size_t)p0 * d.ld, d.ld,
                            d.A + p0 + (size_t)(p0 + pw) * d.ld, d.ld,
                            d.A + (p0 + pw) + (size_t)(p0 + pw) * d.ld, d.ld, -1.0, 1.0));
    }
    st.ntr = (int)tr_all.size() - st.tr0;
    if (!g.empty()) {
      std::vector<int> ids;
      const bool mfma = tune_.gemm_mfma != 0;
      st.bm = tune_.gemm_bm ? tune_.gemm_bm : gemm_tile_m(g, mfma);
      ws = std::max(ws, gemm_plan(g, ids, st.ntiles, tune_.gemm_target, st.bm, mfma));
      st.ng = (int)g.size();
      st.nid = (int)ids.size();
      g_all.insert(g_all.end(), g.begin(), g.end());
      id_all.insert(id_all.end(), ids.begin(), ids.end());
    }
    steps.push_back(st);
  }
  // (the panel updates have K = panel width <= 64: never split, no workspace;
  // ensure_ws would reallocate a buffer another stream may be using)
  if (ws && s != stream_) throw std::logic_error("split-K update in a side-stream LU");
  if (ws) ensure_ws(ws);
  const LuStepDesc* dtr = tr_all.empty() ? nullptr : arena.push(tr_all, s);
  const GemmDesc* dg = g_all.empty() ? nullptr : arena.push(g_all, s);
  const int* did = id_all.empty() ? nullptr : arena.push(id_all, s);
  int p = 0;
  for (int p0 = 0; p0 < maxn; p0 += nb, p++) {
    launch_lu_panel(dd, (int)descs.size(), p0, nb, maxn, s);
    const Step& st = steps[p];
    if (st.ntr > 0) launch_lu_step(dtr + st.tr0, st.ntr, st.nwg, defer, nb, s);
    if (st.ng > 0)
      launch_gemm<double>(dg + st.g0, st.ng, st.ntiles, ws_.p, did ? did + st.id0 : nullptr,
                          st.nid, st.bm, s, tune_.gemm_mfma != 0);
  }
}

/* B <- A^-1 B for a batch of LU-factored matrices: row interchanges, then
 * blocked forward (unit L) and backward (U) substitution, each block a small
 * triangular solve followed by a grouped GEMM update of the remaining rows. */
void Engine::lu_solve_batch(std::vector<SolveDesc>& descs, int mode) {
  if (descs.empty()) return;
  int maxn = 0;
  bool perms = true;
  for (auto& d : descs) {
    maxn = std::max(maxn, d.n);
    perms = perms && d.perm != nullptr;
  }
  if (mode == 4 && maxn <= LU_SOLVE_WIDE_MAX && maxn <= solve_wave_max_) {
    int nu = 0;
    for (auto& d : descs) {
      d.chunk0 = nu;
      nu += lu_solve_wide_units(d.ncols);
    }
    launch_lu_solve_wide(arena_.push(descs, stream_), (int)descs.size(), nu, maxn, stream_);
    return;
  }
  if (mode == 3 && maxn <= LU_SOLVE_MFMA_MAX && perms) {
    int nwg = 0;
    for (auto& d : descs) {
      d.chunk0 = nwg;
      nwg += lu_solve_mfma_wgs(d.ncols);
    }
    launch_lu_solve_mfma(arena_.push(descs, stream_), (int)descs.size(), nwg, maxn, stream_);
    return;
  }
  if (mode == 2 && maxn <= LU_SOLVE_WAVE_MAX && maxn <= solve_wave_max_) {
    int nwg = 0;
    for (auto& d : descs) {
      d.chunk0 = nwg;
      nwg += lu_solve_shared_wgs(d.ncols, maxn);
    }
    launch_lu_solve_shared(arena_.push(descs, stream_), (int)descs.size(), nwg, maxn, stream_);
    return;
  }
  if (maxn <= solve_small_max_ && mode != 1) {  // every system in one launch
    const bool wave = maxn <= solve_wave_max_;
    int nwg = 0;
    for (auto& d : descs) {
      d.chunk0 = nwg;
      nwg += lu_solve_small_wgs(d.ncols, maxn, wave);
    }
    launch_lu_solve_small(arena_.push(descs, stream_), (int)descs.size(), nwg, maxn, wave,
                          solve_blk_, stream_);
    return;
  }
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
    const int nblk = (maxn + tune_.lu_nb - 1) / tune_.lu_nb;
    for (int bi = 0; bi < nblk; bi++) {
      std::vector<TrsmDesc> tr;
      std::vector<GemmDesc> g;
      int nwg = 0;
      for (auto& d : descs) {
        const int nb_d = (d.n + tune_.lu_nb - 1) / tune_.lu_nb;
        const int blk = dir == 0 ? bi : nb_d - 1 - bi;  // backward: from the last block
        if (blk < 0 || blk >= nb_d) continue;
        const int p0 = blk * tune_.lu_nb;
        const int pw = std::min(tune_.lu_nb, d.n - p0);
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
  HODLR_HIP_CHECK(hipEventRecord(ev0_, stream_));
  const int nn = (int)nodes_.size();
  U_.clear();
  V_.clear();
  U_.resize(nlev_);
  V_.resize(nlev_);
  K_.assign(nlev_, 0);
  off_.assign(nlev_ + 1, 0);
  bool any_aca = false, any_svd = false;
  for (auto& nd : nodes_) {
    if (node_method(nd) == 1)
      any_svd = true;
    else
      any_aca = true;
  }
  std::vector<int> blk_rank(2 * nn, -1);
  if (any_aca) aca_compress(blk_rank);
  if (any_svd) svd_compress(blk_rank);
  for (int id = 0; id < nn; id++) {
    nodes_[id].kt = blk_rank[2 * id];
    nodes_[id].kb = blk_rank[2 * id + 1];
    if (nodes_[id].kt < 1 || nodes_[id].kb < 1) throw std::logic_error("a block was not compressed");
    const int l = nodes_[id].level;
    K_[l] = std::max(K_[l], std::max(nodes_[id].kt, nodes_[id].kb));
  }
  off_[0] = 0;
  for (int l = 0; l < nlev_; l++) off_[l + 1] = off_[l] + K_[l];
  ktot_ = off_[nlev_];
  assemble_leaves();
  HODLR_HIP_CHECK(hipEventRecord(ev1_, stream_));
  HODLR_HIP_CHECK(hipEventSynchronize(ev1_));
  float ms = 0;
  HODLR_HIP_CHECK(hipEventElapsedTime(&ms, ev0_, ev1_));
  stats_.compress_time = ms * 1e-3;
  arena_.reset();
  compressed_ = true;
}

void Engine::aca_compress(std::vector<int>& blk_rank) {
  stats_.aca_panel_grows = 0;
  const int nn = (int)nodes_.size();
  const int nl = std::max(1, nlev_ - 1);
  if (nn == 0) return;
  const int base_pt = tune_.aca_pt;
  const int cpw = 256 * base_pt, rpw = 256 * base_pt;
  int cur_pt = base_pt;  // panel rows per thread of the current launches
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
      b.min_rank = nd.min_rank[h];  // -1 in default trees (HODLR_Tree.cpp:44-47)
      b.max_rank = nd.max_rank[h];
      b.skip = node_method(nd) != 0 ? 2 : (nd.max_rank[h] == 0 ? 1 : 0);
      b.pad_ = 0;
      b.lidx = lvl_count[b.level]++;
      b.part_off = lvl_slots[b.level];
      lvl_slots[b.level] += b.row_nwg + b.col_nwg;
      upper_max[b.level] = std::max(upper_max[b.level], std::min(b.m, b.n));
    }
  }
  std::vector<int> cap(nl, 0);
  std::vector<DBuf<double>> P(nl);
  AcaLevels L{};
  // (the panel passes address 16 columns of a panel by 32-bit byte offsets)
  if (ntot_ >= (1 << 24)) throw std::invalid_argument("ACA: more than 2^24 - 1 rows in one handle");
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
  // initial panel capacity per level: the rank predictor's capacities when
  // one is attached (all blocks of kernel instances), else a fixed start
  std::vector<int> pcap;
  if (!pred_params_.empty()) pcap = predicted_capacities();
  std::vector<int> lvl_pcap(nl, 0);
  for (int b = 0; b < (int)pcap.size(); b++)
    lvl_pcap[blk[b].level] = std::max(lvl_pcap[blk[b].level], pcap[b]);
  for (int l = 0; l < nl; l++) {
    if (lvl_count[l] == 0) continue;
    int c = init_cap_ > 0 ? init_cap_ : 32;
    if (!pcap.empty()) c = lvl_pcap[l] + 1;  // + the speculative column
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

  std::vector<int> active;
  for (int b = 0; b < nb; b++)
    if (blk[b].skip == 0 && blk[b].m > 0 && blk[b].n > 0) active.push_back(b);
  std::vector<int> m_row, m_col, m_fin, m_chunk, nchunk(nb, 1);
  AcaLaunch a{};
  // When the active blocks would give fewer row-pass workgroups than the
  // device has CUs (the last one or two blocks of the top level), their
  // panel rows are spread over more workgroups (fewer rows per thread):
  // the partial slots of the active blocks are laid out again, compactly per
  // level (they only carry values within an iteration).
  auto relayout = [&](int pt) {
    std::vector<int> used(nl, 0);
    for (int b : active) {
      AcaBlock& x = blk[b];
      x.row_nwg = (x.n + 256 * pt - 1) / (256 * pt);
      x.col_nwg = (x.m + 256 * pt - 1) / (256 * pt);
      x.part_off = used[x.level];
      used[x.level] += x.row_nwg + x.col_nwg;
    }
    for (int l = 0; l < nl; l++)
      if (used[l] > lvl_slots[l]) {
        HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
        lvl_slots[l] = used[l];
        P[l].alloc((size_t)lvl_slots[l] * (cap[l] + 4));
        L.P[l] = P[l].p;
      }
    cur_pt = pt;
  };
  auto build_maps = [&]() {
    if (adaptive_aca_pt_) {
      long long rw = 0;
      for (int b : active) rw += (blk[b].n + 256 * base_pt - 1) / (256 * base_pt);
      int pt = base_pt;
      while (pt > 1 && rw > 0 && rw * (base_pt / pt) < aca_min_wg_) pt /= 2;
      if (pt != cur_pt) relayout(pt);
    }
    m_row.clear();
    m_col.clear();
    m_fin.clear();
    m_chunk.clear();
    for (int b = 0; b < nb; b++) {
      nchunk[b] = std::max(1, (cap[blk[b].level] + 63) / 64);
      if (nchunk[b] > MAX_FIN_CHUNKS) throw std::runtime_error("ACA rank capacity exceeds 16384");
    }
    for (int b : active) {
      // the finalize addresses a block's partial slots with 32-bit byte
      // offsets from its first slot
      if ((size_t)(blk[b].row_nwg + blk[b].col_nwg) * (cap[blk[b].level] + 4) * sizeof(double) >
          (size_t)INT32_MAX)
        throw std::runtime_error("ACA: a block's partial slots exceed 2 GiB (rank capacity too large)");
      blk[b].row_wg0 = (int)m_row.size();
      for (int w = 0; w < blk[b].row_nwg; w++) m_row.push_back(b);
      blk[b].col_wg0 = (int)m_col.size();
      for (int w = 0; w < blk[b].col_nwg; w++) m_col.push_back(b);
      const int nfin = aca_fin_packed(blk[b]) ? (nchunk[b] + 3) / 4 : nchunk[b];
      for (int c = 0; c < nfin; c++) {
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
    a.jpt = cur_pt;
    a.ipt = cur_pt;
    a.tol = tol_;
    a.min_pivot = min_pivot_;
  };
  build_maps();

  std::vector<AcaState> hs(nb);
  // the state snapshot in pinned memory, so that its copy is asynchronous
  struct PinnedStates {
    AcaState* p = nullptr;
    explicit PinnedStates(int n) {
      HODLR_HIP_CHECK(hipHostMalloc((void**)&p, (size_t)std::max(n, 1) * sizeof(AcaState)));
    }
    ~PinnedStates() { (void)hipHostFree(p); }
  } pin(nb);
  int check = 4;
  // every iteration either accepts a term or consumes an unchosen row, so a
  // block ends within m + n iterations
  long long max_dim = 0;
  for (auto& b : blk) max_dim = std::max<long long>(max_dim, (long long)b.m + b.n);
  long long launched = 0;
  auto launch_batch = [&](int count) {
    if (profiling_) {
      while ((int)prof_ev_.size() < 3 * count) {
        hipEvent_t e;
        HODLR_HIP_CHECK(hipEventCreate(&e));
        prof_ev_.push_back(e);
      }
      for (int c = 0; c < count; c++)
        aca_launch_iteration(L, a, d_srcs_.p, stream_, prof_ev_.data() + 3 * c);
    } else {
      for (int c = 0; c < count; c++) aca_launch_iteration(L, a, d_srcs_.p, stream_);
    }
    launched += count;
  };
  // Without profiling the next batch is launched before the state of the
  // previous one is read (the GPU never waits for the host): blocks that
  // finished or paused meanwhile skip it (their state says so), and the
  // maps rebuilt from the state take effect with the batch after it.
  const bool ahead = !profiling_;
  launch_batch(check);
  while (!active.empty()) {
    if (launched > max_dim + 64)
      throw std::runtime_error("ACA did not terminate within m + n iterations");
    HODLR_HIP_CHECK(hipMemcpyAsync(pin.p, d_state.p, nb * sizeof(AcaState), hipMemcpyDeviceToHost, stream_));
    HODLR_HIP_CHECK(hipEventRecord(ev_aca_, stream_));
    const int next = std::min(16, check * 2);
    if (ahead) launch_batch(next);
    HODLR_HIP_CHECK(hipEventSynchronize(ev_aca_));
    std::memcpy(hs.data(), pin.p, nb * sizeof(AcaState));
    if (profiling_) {
      for (int c = 0; c < check; c++) {
        float r = 0, q = 0;
        HODLR_HIP_CHECK(hipEventElapsedTime(&r, prof_ev_[3 * c], prof_ev_[3 * c + 1]));
        HODLR_HIP_CHECK(hipEventElapsedTime(&q, prof_ev_[3 * c + 1], prof_ev_[3 * c + 2]));
        stats_.aca_row_time += r * 1e-3;
        stats_.aca_col_time += q * 1e-3;
        stats_.aca_launches += 1;
      }
    }
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
        if (grow[l]) {
          alloc_level(l, std::min(upper_max[l], std::max(cap[l] + 1, 2 * cap[l])));
          stats_.aca_panel_grows++;
        }
      // (a paused block made no progress since the snapshot: only its own
      // state is written back)
      for (int b : active)
        if (hs[b].status == 2) {
          hs[b].status = 0;
          pin.p[b].status = 0;
          na.push_back(b);
          HODLR_HIP_CHECK(hipMemcpyAsync(d_state.p + b, pin.p + b, sizeof(AcaState),
                                         hipMemcpyHostToDevice, stream_));
        }
      std::sort(na.begin(), na.end());
    }
    if (na.empty()) break;
    const bool changed = paused || na.size() != active.size();
    active.swap(na);
    if (changed) build_maps();
    check = next;
    if (!ahead) launch_batch(check);
  }
  HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
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
    if (need[l] > cap[l]) {
      alloc_level(l, need[l]);
      stats_.aca_panel_grows++;
    }
  stats_.aca_panel_bytes = 0;
  for (int l = 0; l < nl; l++) stats_.aca_panel_bytes += 2.0 * U_[l].bytes();
  for (int l = 0; l < nl; l++) P[l].release();
  aca_launch_post(L, d_blk.p, d_state.p, d_srcs_.p, d_ranks.p, nb, stream_);
  std::vector<int> ranks(nb);
  HODLR_HIP_CHECK(hipMemcpyAsync(ranks.data(), d_ranks.p, nb * sizeof(int), hipMemcpyDeviceToHost, stream_));
  HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
  for (int b = 0; b < nb; b++)
    if (blk[b].skip != 2) blk_rank[b] = ranks[b];
  if (!pcap.empty()) {  // how the predictor did: misses need the adaptive growth
    stats_.pred_blocks = 0;
    stats_.pred_misses = 0;
    double err = 0.0;
    for (int b = 0; b < nb; b++) {
      if (blk[b].skip != 0) continue;
      stats_.pred_blocks++;
      if (ranks[b] > pcap[b]) stats_.pred_misses++;
      err += std::fabs(std::log(std::max(pred_rank_[b], 1.0)) - std::log((double)std::max(ranks[b], 1)));
    }
    stats_.pred_log_err = stats_.pred_blocks ? err / stats_.pred_blocks : 0.0;
  }
  // algorithmic bytes of the residual passes: 8 (m+n) k(k+1)/2 per block
  double bytes = 0;
  for (int b = 0; b < nb; b++) {
    const double k = hs[b].k;
    bytes += 8.0 * (double)(blk[b].m + blk[b].n) * k * (k + 1) / 2;
  }
  stats_.aca_bytes = bytes;
}

/* SVD_LowRankApprox (lowRank.cpp:520-569) for every off-diagonal block: the
 * block is assembled (transposed when it has fewer rows than columns, so the
 * working matrix has m >= n), orthogonalized by batched one-sided Jacobi
 * sweeps, and its singular values sigma_j = |a_j| sorted in decreasing order
 * (stable).  The rank is the number of sigma_j >= tol (absolute, :529-533);
 * rank 0 stores one zero column (:549-561).  Stored: U Sigma and V, as the
 * reference does. */
void Engine::svd_compress(std::vector<int>& blk_rank) {
  const int nn_all = (int)nodes_.size();
  std::vector<int> bid;  // global block index (2 * node + half) of each SVD block
  for (int id = 0; id < nn_all; id++)
    if (node_method(nodes_[id]) == 1) {
      bid.push_back(2 * id);
      bid.push_back(2 * id + 1);
    }
  if (bid.empty()) return;
  const int nb = (int)bid.size();
  struct Blk {
    int m, n, mm, nn_, trans, row_u, row_v, level, min_rank;
    size_t a_off, v_off;
  };
  std::vector<Blk> bl(nb);
  size_t a_tot = 0, v_tot = 0;
  std::vector<AssembleDesc> ad;
  std::vector<CopyDesc> idd;
  int max_e = 0;
  for (int bi = 0; bi < nb; bi++) {
    const int id = bid[bi] / 2, h = bid[bi] % 2;
    const TreeNodeH& nd = nodes_[id];
    const int st = nd.split - nd.min_i + 1, sb = nd.max_i - nd.split;
    {
      Blk& b = bl[bi];
      b.min_rank = nd.min_rank[h];
      b.m = h == 0 ? st : sb;
      b.n = h == 0 ? sb : st;
      b.trans = b.m < b.n;
      b.mm = std::max(b.m, b.n);
      b.nn_ = std::min(b.m, b.n);
      b.row_u = h == 0 ? nd.row0 : nd.row0 + st;  // rows of U (block rows)
      b.row_v = h == 0 ? nd.row0 + st : nd.row0;  // rows of V (block columns)
      b.level = nd.level;
      b.a_off = a_tot;
      b.v_off = v_tot;
      a_tot += (size_t)b.mm * b.nn_;
      v_tot += (size_t)b.nn_ * b.nn_;
      AssembleDesc a{};
      a.inst = nd.inst;
      a.i0 = h == 0 ? nd.min_i : nd.split + 1;
      a.j0 = h == 0 ? nd.split + 1 : nd.min_i;
      a.m = b.m;
      a.n = b.n;
      a.ldo = b.mm;
      a.trans = b.trans;
      ad.push_back(a);
      max_e = std::max(max_e, b.m * b.n);
    }
  }
  DBuf<double> A, Vw;
  A.alloc(a_tot);
  Vw.alloc(v_tot);
  std::vector<SvdDesc> sd(nb);
  int wg = 0, max_steps = 0, maxn = 0;
  for (int b = 0; b < nb; b++) {
    ad[b].out = A.p + bl[b].a_off;
    idd.push_back({nullptr, Vw.p + bl[b].v_off, bl[b].nn_, bl[b].nn_, 0, bl[b].nn_});
    SvdDesc& d = sd[b];
    d.A = A.p + bl[b].a_off;
    d.V = Vw.p + bl[b].v_off;
    d.m = bl[b].mm;
    d.n = bl[b].nn_;
    d.wg0 = wg;
    wg += svd_step_wgs(d);
    const int ne = d.n + (d.n & 1);
    max_steps = std::max(max_steps, ne - 1);
    maxn = std::max(maxn, d.n);
  }
  launch_assemble(arena_.push(ad, stream_), nb, max_e, d_srcs_.p, stream_);
  launch_set_identity(arena_.push(idd, stream_), nb, stream_);
  std::vector<int> soff(nb + 1, 0);
  for (int b = 0; b < nb; b++) soff[b + 1] = soff[b] + bl[b].nn_;
  DBuf<int> doff;
  doff.alloc(nb + 1);
  HODLR_HIP_CHECK(hipMemcpyAsync(doff.p, soff.data(), soff.size() * sizeof(int), hipMemcpyHostToDevice, stream_));
  DBuf<double> dsig;
  dsig.alloc(std::max(soff[nb], 1));
  std::vector<double> sig(soff[nb]);
  {  // noise floor of the rotations from the initial column norms
    const SvdDesc* d0 = arena_.push(sd, stream_);
    launch_svd_norms(d0, nb, doff.p, maxn, dsig.p, stream_);
    HODLR_HIP_CHECK(hipMemcpyAsync(sig.data(), dsig.p, sig.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
    const double eps = 2.220446049250313e-16;
    for (int b = 0; b < nb; b++) {
      double c2 = 0.0;
      for (int j = soff[b]; j < soff[b + 1]; j++) c2 = std::max(c2, sig[j] * sig[j]);
      sd[b].floor2 = eps * eps * c2;
    }
  }
  const SvdDesc* dsd = arena_.push(sd, stream_);
  DBuf<unsigned> cnt;
  cnt.alloc(1);
  int sweeps = 0;
  for (; sweeps < 60; sweeps++) {  // at most 60 sweeps, as the oracle's Jacobi
    HODLR_HIP_CHECK(hipMemsetAsync(cnt.p, 0, sizeof(unsigned), stream_));
    for (int step = 0; step < max_steps; step++)
      launch_svd_step(dsd, nb, wg, step, cnt.p, stream_);
    unsigned h = 0;
    HODLR_HIP_CHECK(hipMemcpyAsync(&h, cnt.p, sizeof(unsigned), hipMemcpyDeviceToHost, stream_));
    HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
    if (h == 0) break;
  }
  stats_.aca_iterations = sweeps + 1;
  // singular values, sorted; ranks
  launch_svd_norms(dsd, nb, doff.p, maxn, dsig.p, stream_);
  HODLR_HIP_CHECK(hipMemcpyAsync(sig.data(), dsig.p, sig.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
  HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
  std::vector<std::vector<int>> ord(nb);
  std::vector<int> rank(nb);
  for (int b = 0; b < nb; b++) {
    const double* sgm = sig.data() + soff[b];
    std::vector<int>& o = ord[b];
    o.resize(bl[b].nn_);
    for (int j = 0; j < bl[b].nn_; j++) o[j] = j;
    std::stable_sort(o.begin(), o.end(), [&](int x, int y) { return sgm[x] > sgm[y]; });
    int r = bl[b].nn_;
    for (int j = 0; j < bl[b].nn_; j++)
      if (sgm[o[j]] < tol_) {
        r = j;
        break;
      }
    // at least the node's minimum rank (lowRank.cpp:534-535), at most min(m, n)
    r = std::min(bl[b].nn_, std::max(r, bl[b].min_rank));
    rank[b] = r;
  }
  // panel columns each level needs (rank 0 keeps one zero column); panels
  // that ACA already holds for other blocks of the level grow if needed
  std::vector<int> need(nlev_, 0);
  for (int b = 0; b < nb; b++) need[bl[b].level] = std::max(need[bl[b].level], std::max(rank[b], 1));
  for (int l = 0; l < nlev_; l++) {
    if (need[l] == 0) continue;
    const size_t cur = U_[l].n / std::max<size_t>(1, (size_t)ntot_);
    if ((size_t)need[l] <= cur) continue;
    DBuf<double> u, v;
    u.alloc((size_t)ntot_ * need[l]);
    v.alloc((size_t)ntot_ * need[l]);
    HODLR_HIP_CHECK(hipMemsetAsync(u.p, 0, u.bytes(), stream_));
    HODLR_HIP_CHECK(hipMemsetAsync(v.p, 0, v.bytes(), stream_));
    if (cur > 0) {
      HODLR_HIP_CHECK(hipMemcpyAsync(u.p, U_[l].p, U_[l].bytes(), hipMemcpyDeviceToDevice, stream_));
      HODLR_HIP_CHECK(hipMemcpyAsync(v.p, V_[l].p, V_[l].bytes(), hipMemcpyDeviceToDevice, stream_));
    }
    HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
    U_[l] = std::move(u);
    V_[l] = std::move(v);
  }
  // U Sigma and V into the level panels (rank 0 keeps its zero column)
  std::vector<ColCopyDesc> cc;
  for (int b = 0; b < nb; b++) {
    const Blk& B = bl[b];
    const double* sgm = sig.data() + soff[b];
    for (int c = 0; c < rank[b]; c++) {
      const int j = ord[b][c];
      const double* acol = A.p + B.a_off + (size_t)j * B.mm;
      const double* vcol = Vw.p + B.v_off + (size_t)j * B.nn_;
      double* ucol = U_[B.level].p + B.row_u + (size_t)c * ntot_;
      double* vout = V_[B.level].p + B.row_v + (size_t)c * ntot_;
      if (!B.trans) {  // A = (A V) V^T: U Sigma = A V, V = V
        cc.push_back({acol, ucol, 1.0, B.m, 0});
        cc.push_back({vcol, vout, 1.0, B.n, 0});
      } else {  // A^T = (A^T W) W^T: U Sigma = W Sigma, V = (A^T W) / Sigma
        cc.push_back({vcol, ucol, sgm[j], B.m, 0});
        cc.push_back({acol, vout, sgm[j], B.n, 1});
      }
    }
  }
  for (size_t c0 = 0; c0 < cc.size(); c0 += 65535) {
    const int cnt_c = (int)std::min<size_t>(65535, cc.size() - c0);
    std::vector<ColCopyDesc> part(cc.begin() + c0, cc.begin() + c0 + cnt_c);
    launch_col_copy(arena_.push(part, stream_), cnt_c, stream_);
  }
  HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
  for (int b = 0; b < nb; b++) blk_rank[bid[b]] = std::max(rank[b], 1);
}

/* Features of every off-diagonal block (2 per node, node order: top, then
 * bottom) for the rank predictor; empty when an instance has no points. */
void Engine::block_features_device(double tol, DBuf<float>& out) {
  const int nn = (int)nodes_.size();
  std::vector<FeatDesc> fd;
  for (int id = 0; id < nn; id++) {
    const TreeNodeH& nd = nodes_[id];
    const EntrySource& es = h_srcs_[nd.inst];
    if (es.kind != 0) throw std::invalid_argument("block features need a kernel (point) instance");
    for (int h = 0; h < 2; h++) {
      FeatDesc f{};
      f.x = es.kp.x;
      f.y = es.kp.y;
      f.z = es.kp.z;
      const int top0 = nd.min_i, bot0 = nd.split + 1;
      const int st = nd.split - nd.min_i + 1, sb = nd.max_i - nd.split;
      f.r0 = h == 0 ? top0 : bot0;
      f.nr = h == 0 ? st : sb;
      f.c0 = h == 0 ? bot0 : top0;
      f.nc = h == 0 ? sb : st;
      f.dim = es.kp.dim;
      f.level = nd.level;
      f.length_scale = (float)(1.0 / es.kp.inv_l);
      f.log10_tol = (float)std::log10(tol);
      f.family = (float)es.kp.family;
      fd.push_back(f);
    }
  }
  out.alloc(std::max<size_t>(1, fd.size() * MLP_IN));
  launch_block_features(arena_.push(fd, stream_), (int)fd.size(), out.p, stream_);
}

void Engine::block_features(double tol, float* host_out) {
  DBuf<float> f;
  block_features_device(tol, f);
  HODLR_HIP_CHECK(hipMemcpyAsync(host_out, f.p, nodes_.size() * 2 * MLP_IN * sizeof(float),
                                 hipMemcpyDeviceToHost, stream_));
  HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
  arena_.reset();
}

void Engine::set_rank_predictor(const float* params, double margin) {
  if (params) {
    if (!(margin >= 0.0)) throw std::invalid_argument("margin must be >= 0");
    pred_params_.assign(params, params + MLP_NPARAM);
  } else {
    pred_params_.clear();
  }
  pred_margin_ = margin;
  reset_attributes();
}

/* Rank capacity of every block from the predictor (BF16 inference):
 * ceil(exp(y) (1 + margin)), y the predicted log-rank. */
std::vector<int> Engine::predicted_capacities() {
  const int nb = 2 * (int)nodes_.size();
  DBuf<float> feat, P, y;
  block_features_device(tol_, feat);
  P.alloc(MLP_NPARAM);
  y.alloc(std::max(nb, 1));
  HODLR_HIP_CHECK(hipMemcpyAsync(P.p, pred_params_.data(), MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice, stream_));
  launch_mlp_infer_bf16(P.p, feat.p, nb, y.p, stream_);
  std::vector<float> hy(nb);
  HODLR_HIP_CHECK(hipMemcpyAsync(hy.data(), y.p, nb * sizeof(float), hipMemcpyDeviceToHost, stream_));
  HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
  std::vector<int> cap(nb);
  pred_rank_.assign(nb, 0.0);
  for (int b = 0; b < nb; b++) {
    const double r = std::exp(std::min(20.0, (double)hy[b]));
    pred_rank_[b] = r;
    cap[b] = std::max(1, (int)std::ceil(r * (1.0 + pred_margin_)));
  }
  return cap;
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
/* The leaves' LU factors (PartialPivLU, HODLR_Matrix.cpp:541, :552). */
void Engine::factor_leaves() {
  if (leaves_factored_) return;
  Lfac_.alloc(leaf_elems_);
  Lpiv_.alloc(leaf_piv_elems_);
  HODLR_HIP_CHECK(hipMemcpyAsync(Lfac_.p, Lorig_.p, leaf_elems_ * sizeof(double), hipMemcpyDeviceToDevice, stream_));
  std::vector<LuDesc> ld;
  for (auto& lf : leaves_) ld.push_back({Lfac_.p + lf.mat_off, Lpiv_.p + lf.piv_off, lf.s, lf.s});
  double fl = 0, by = 0;
  lu_work(ld, fl, by);
  const int t = prof_begin(stream_);
  lu_batch(ld);
  prof_end(t, stream_, FAC_LU, fl, by);
  // the interchanges as permutations, for the matrix-core leaf solves
  Lperm_.alloc(std::max<size_t>(leaf_piv_elems_, 1));
  std::vector<PermBuild> pb;
  for (auto& lf : leaves_)
    if (lf.s <= LU_PERM_MAX) pb.push_back({Lpiv_.p + lf.piv_off, Lperm_.p + lf.piv_off, lf.s, 0});
  launch_lu_perm(arena_.push(pb, stream_), (int)pb.size(), stream_);
  leaves_factored_ = true;
}

/* B[rows of leaf] <- A_leaf^-1 B[rows of leaf] for the given leaves, by
 * their LU factors (the reference's leafLU.solve, HODLR_Matrix.cpp:598-601,
 * :569-572): B points at global row 0 (ld ldb) unless compact, where the
 * leaves' blocks follow each other (s x ncols each, ld s). */
void Engine::leaf_solves(const std::vector<int>& leaf_ids, double* B, int ldb, int ncols,
                         bool compact) {
  if (leaf_ids.empty() || ncols <= 0) return;
  std::vector<SolveDesc> sd;
  size_t o = 0;
  for (int li : leaf_ids) {
    const LeafH& lf = leaves_[li];
    SolveDesc s{};
    s.LU = Lfac_.p + lf.mat_off;
    s.piv = Lpiv_.p + lf.piv_off;
    s.perm = lf.s <= LU_PERM_MAX ? Lperm_.p + lf.piv_off : nullptr;
    s.B = compact ? B + o : B + lf.row0;
    s.n = lf.s;
    s.ld_lu = lf.s;
    s.ldb = compact ? lf.s : ldb;
    s.ncols = ncols;
    sd.push_back(s);
    o += (size_t)lf.s * ncols;
  }
  // many columns (the factorization's leaf step, solves of many right-hand
  // sides): the tuned variant
  const int mode = ncols < 64 ? 0 : tune_.leaf_solve;
  lu_solve_batch(sd, mode);
}

/* recSM_Factorize (HODLR_Matrix.cpp:520-592) and recLU_Factorize (:612-692),
 * level by level from the deepest: leaves of level d and the Woodbury
 * (or Schur) factors of the nodes of level d are applied to the rows that
 * the ancestors' U^SM columns hold. */
void Engine::factorize(int which) {
  if (which != 0 && which != 1) throw std::invalid_argument("unknown factorization");
  if (precision_ == 1) {
    factorize_mixed(which);
    return;
  }
  factorize_fp64(which);
}

void Engine::factorize_fp64(int which) {
  compress();
  Factorization& F = fac_[which];
  if (F.done) return;
  finish_root();  // a previous factorization's root LU may still use F's buffers
  buf_gen_++;  // captured solves refer to the factorization's buffers
  factor_time_pending_ = false;
  prof_fac_ = profiling_;
  prof_depth_ = 0;
  spans_.clear();
  span_next_ = 0;
  HODLR_HIP_CHECK(hipEventRecord(ev_f0_, stream_));
  const size_t N = (size_t)ntot_;
  F.W.alloc(std::max<size_t>(1, N * ktot_));
  // W starts as a copy of the panels (the reference's U_SM = U, :561-565):
  // the copies (HBM-bound) run on the side stream while the leaves' LU
  // (latency-bound) runs here; the leaf solves wait for them
  HODLR_HIP_CHECK(hipEventRecord(ev_w0_, stream_));
  HODLR_HIP_CHECK(hipStreamWaitEvent(aux_, ev_w0_, 0));
  for (int l = 0; l < nlev_; l++)
    if (K_[l] > 0)
      HODLR_HIP_CHECK(hipMemcpyAsync(F.W.p + N * off_[l], U_[l].p, N * K_[l] * sizeof(double),
                                     hipMemcpyDeviceToDevice, aux_));
  HODLR_HIP_CHECK(hipEventRecord(ev_w1_, aux_));
  factor_leaves();
  HODLR_HIP_CHECK(hipStreamWaitEvent(stream_, ev_w1_, 0));
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
  F.perm.alloc(std::max<size_t>(po, 1));
  F.perm_level.assign(nlev_, 0);
  double* W = F.W.p;
  const int ld = ntot_;
  for (int d = nlev_ - 1; d >= 0; d--) {
    const int offd = off_[d];
    // leaves of level d: W[rows, 0:off_d] <- A_leaf^-1 W[rows, 0:off_d] (in
    // place by the LU factors; the columns of level d and deeper are zero
    // padding in a leaf's rows)
    if (offd > 0 && !level_leaves_[d].empty()) {
      double fl = 0, by = 0;
      for (int li : level_leaves_[d]) {
        const double sz = leaves_[li].s;
        fl += 2.0 * sz * sz * offd;
        by += 8.0 * (sz * sz + 2.0 * sz * offd);
      }
      const int t = prof_begin(stream_);
      leaf_solves(level_leaves_[d], W, ld, offd, false);
      prof_end(t, stream_, FAC_LEAF_SOLVE, fl, by);
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
    double fl = 0, by = 0;
    gemm_work(gz, fl, by);
    int t = prof_begin(stream_);
    run_gemm(gz);
    prof_end(t, stream_, FAC_GEMM, fl, by);
    launch_capacitance(arena_.push(cd, stream_), (int)cd.size(), stream_);
    lu_work(lu, fl, by);
    if (offd == 0 && overlap_root_) {
      // the roots: nothing later in the factorization reads their LU, so it
      // runs on aux_ and the solve that follows waits for it at its last level
      HODLR_HIP_CHECK(hipEventRecord(ev_cap_, stream_));
      HODLR_HIP_CHECK(hipStreamWaitEvent(aux_, ev_cap_, 0));
      t = prof_begin(aux_);
      lu_batch(lu, aux_, aux_arena_);
      prof_end(t, aux_, FAC_LU, fl, by);
      HODLR_HIP_CHECK(hipEventRecord(ev_root_, aux_));
      HODLR_HIP_CHECK(hipEventRecord(ev_froot_, aux_));
      F.root_pending = true;
      factor_time_pending_ = true;
    } else {
      t = prof_begin(stream_);
      lu_batch(lu);
      prof_end(t, stream_, FAC_LU, fl, by);
    }
    if (offd > 0) {
      fl = by = 0;
      for (auto& v : sd) {
        fl += 2.0 * v.n * (double)v.n * v.ncols;
        by += 8.0 * ((double)v.n * v.n + 2.0 * v.n * (double)v.ncols);
      }
      t = prof_begin(stream_);
      // node matrices of at most 256 rows: on the matrix cores with the leaf
      // solves, their interchanges composed into permutations first
      const bool mc = tune_.leaf_solve == 3 && maxK <= LU_SOLVE_MFMA_MAX;
      if (mc) {
        std::vector<PermBuild> pb;
        for (size_t i = 0; i < sd.size(); i++) {
          const int id = level_nodes_[d][i];
          int* pm = F.perm.p + F.p_off[id];
          pb.push_back({sd[i].piv, pm, sd[i].n, 0});
          sd[i].perm = pm;
        }
        launch_lu_perm(arena_.push(pb, stream_), (int)pb.size(), stream_);
        F.perm_level[d] = 1;
      }
      lu_solve_batch(sd, mc ? 3 : 0);
      prof_end(t, stream_, FAC_NODE_SOLVE, fl, by);
      gemm_work(gu, fl, by);
      t = prof_begin(stream_);
      run_gemm(gu);
      prof_end(t, stream_, FAC_GEMM, fl, by);
    }
  }
  HODLR_HIP_CHECK(hipEventRecord(ev_f1_, stream_));
  HODLR_HIP_CHECK(hipEventSynchronize(ev_f1_));
  float ms = 0;
  HODLR_HIP_CHECK(hipEventElapsedTime(&ms, ev_f0_, ev_f1_));
  stats_.factor_stream_time = ms * 1e-3;
  stats_.factor_time = ms * 1e-3;  // completed by sync_stats while the roots' LU is pending
  if (prof_fac_) {
    HODLR_HIP_CHECK(hipStreamSynchronize(aux_));
    prof_collect();
    prof_fac_ = false;
  }
  arena_.reset();
  F.done = true;
}

/* ------------------------------------------------------------------ */
/* solve: recSM_Solve (HODLR_Matrix.cpp:594-610) / recLU_Solve (:694-755) */
/* ------------------------------------------------------------------ */
void Engine::solve_device(int which, const double* dB, int ldb, int nrhs, double* dX, int ldx) {
  if (nrhs <= 0) return;
  if (ldb < ntot_ || ldx < ntot_) throw std::invalid_argument("leading dimension < n");
  if (precision_ != 0) {
    solve_gmres(which, precision_ == 1, dB, ldb, nrhs, dX, ldx);
    return;
  }
  factorize_fp64(which);
  HODLR_HIP_CHECK(hipEventRecord(ev0_, stream_));
  const size_t N = (size_t)ntot_;
  grow(xw_, N * nrhs);
  HODLR_HIP_CHECK(hipMemcpy2DAsync(xw_.p, N * sizeof(double), dB, (size_t)ldb * sizeof(double),
                                   N * sizeof(double), nrhs, hipMemcpyDeviceToDevice, stream_));
  if (use_graphs_)
    solve_fp64_graph(which, xw_.p, nrhs);
  else
    solve_fp64_inplace(which, xw_.p, nrhs);
  HODLR_HIP_CHECK(hipMemcpy2DAsync(dX, (size_t)ldx * sizeof(double), xw_.p, N * sizeof(double),
                                   N * sizeof(double), nrhs, hipMemcpyDeviceToDevice, stream_));
  HODLR_HIP_CHECK(hipEventRecord(ev1_, stream_));
  HODLR_HIP_CHECK(hipEventSynchronize(ev1_));
  float ms = 0;
  HODLR_HIP_CHECK(hipEventElapsedTime(&ms, ev0_, ev1_));
  stats_.solve_time = ms * 1e-3;
  arena_.reset();
}

void Engine::set_graphs(bool on) {
  use_graphs_ = on;
  if (!on) graph_.release();
}

/* The FP64 solve through a captured hipGraph.  On first use for a
 * (factorization, nrhs) and whenever a work buffer or the factorization has
 * been reallocated since, the solve runs once eagerly with its descriptor
 * arrays uploaded into the graph's own arena, and is then captured with
 * stream capture while that arena replays the same device pointers (no copy
 * is captured, only kernel launches).  Later calls copy X into the graph's
 * buffer, launch the graph and copy the result back. */
void Engine::solve_fp64_graph(int which, double* X, int nrhs) {
  finish_root();  // (captured and replayed work stays on stream_)
  const size_t N = (size_t)ntot_;
  grow(gx_, N * nrhs);
  const size_t bytes = N * nrhs * sizeof(double);
  HODLR_HIP_CHECK(hipMemcpyAsync(gx_.p, X, bytes, hipMemcpyDeviceToDevice, stream_));
  const bool stale = !graph_.valid || graph_.which != which || graph_.nrhs != nrhs ||
                     graph_.gen != buf_gen_;
  if (!stale) {
    HODLR_HIP_CHECK(hipGraphLaunch(graph_.exec, stream_));
    stats_.graph_launches++;
  } else {
    graph_.release();
    graph_.arena.reset();
    std::swap(arena_, graph_.arena);
    try {
      solve_fp64_inplace(which, gx_.p, nrhs);  // eager run: the result of this call
      HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
      const unsigned long long gen = buf_gen_;
      arena_.begin_replay();
      HODLR_HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeGlobal));
      hipGraph_t g = nullptr;
      try {
        solve_fp64_inplace(which, gx_.p, nrhs);
      } catch (...) {
        (void)hipStreamEndCapture(stream_, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      HODLR_HIP_CHECK(hipStreamEndCapture(stream_, &g));
      graph_.graph = g;
      HODLR_HIP_CHECK(hipGraphInstantiate(&graph_.exec, g, nullptr, nullptr, 0));
      if (buf_gen_ != gen) throw std::runtime_error("work buffers changed during capture");
      graph_.valid = true;
      graph_.which = which;
      graph_.nrhs = nrhs;
      graph_.gen = gen;
      stats_.graph_captures++;
    } catch (...) {
      arena_.end_replay();
      std::swap(arena_, graph_.arena);
      graph_.release();
      throw;
    }
    arena_.end_replay();
    std::swap(arena_, graph_.arena);
  }
  HODLR_HIP_CHECK(hipMemcpyAsync(X, gx_.p, bytes, hipMemcpyDeviceToDevice, stream_));
}

/* X <- A_H^-1 X with the FP64 recSM (recLU) factorization, X is ntot x nrhs
 * with leading dimension ntot: all leaves, then the nodes level by level from
 * the deepest (the post-order of HODLR_Matrix.cpp:594-610 regrouped by level;
 * nodes of one level own disjoint rows). */
void Engine::solve_fp64_inplace(int which, double* X, int nrhs) {
  Factorization& F = fac_[which];
  const int ld = ntot_;
  const size_t N = (size_t)ntot_;
  // leaves: X[rows] <- A_leaf^-1 X[rows] (in place by the LU factors)
  {
    std::vector<int> all(leaves_.size());
    for (size_t i = 0; i < all.size(); i++) all[i] = (int)i;
    leaf_solves(all, X, ld, nrhs, false);
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
    if (offd == 0) wait_root(F);
    // many right-hand sides: the matrix cores where the factorization
    // composed the level's permutations
    const bool mc = nrhs >= 64 && (size_t)d < F.perm_level.size() && F.perm_level[d];
    if (mc) {
      const std::vector<int>& ids = level_nodes_[d];
      for (size_t i = 0; i < sd.size(); i++) sd[i].perm = F.perm.p + F.p_off[ids[i]];
    }
    lu_solve_batch(sd, mc ? 3 : 0);
    run_gemm(gu);
  }
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
  auto copy_back = [&] {
    HODLR_HIP_CHECK(hipMemcpy2D(X, (size_t)ldx * sizeof(double), dX.p, N * sizeof(double),
                                N * sizeof(double), nrhs, hipMemcpyDeviceToHost));
  };
  try {
    solve_device(which, dB.p, (int)N, nrhs, dX.p, (int)N);
  } catch (const NotConverged&) {
    copy_back();  // the best iterate
    throw;
  }
  copy_back();
}

/* ------------------------------------------------------------------ */
/* mixed precision: FP32 bulk, FP64 leaf LU / node matrices / residuals */
/* ------------------------------------------------------------------ */
template <class T>
void Engine::grow_t(DBuf<T>& b, size_t n) {
  if (b.n < n) {
    HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
    b.alloc(n);
  }
}

template <class T>
long long Engine::count_nonfinite(const T* x, size_t n) {
  DBuf<unsigned long long> c;
  c.alloc(1);
  HODLR_HIP_CHECK(hipMemsetAsync(c.p, 0, sizeof(unsigned long long), stream_));
  launch_nonfinite<T>(x, n, c.p, stream_);
  unsigned long long h = 0;
  HODLR_HIP_CHECK(hipMemcpyAsync(&h, c.p, sizeof(h), hipMemcpyDeviceToHost, stream_));
  HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
  return (long long)h;
}

/* The recSM (or recLU) factorization of factorize(), with the U^SM panel and
 * the V panels in FP32 and every bulk product an FP32 GEMM; the leaves' LU and
 * inverses and the node matrices (built, factored and solved) stay FP64. */
void Engine::factorize_mixed(int which) {
  compress();
  MixedFactorization& M = mfac_[which];
  if (M.done) return;
  HODLR_HIP_CHECK(hipEventRecord(ev0_, stream_));
  factor_leaves();
  const size_t N = (size_t)ntot_;
  const int ld = ntot_;
  M.V.clear();
  M.V.resize(nlev_);
  M.W.alloc(std::max<size_t>(1, N * ktot_));
  HODLR_HIP_CHECK(hipMemsetAsync(M.W.p, 0, M.W.bytes(), stream_));
  {
    // rebalanced FP32 copies of every node's two blocks; rows and terms
    // outside them stay zero
    std::vector<BalanceDesc> bd;
    int kmax = 0;
    for (int d = 0; d < nlev_; d++) {
      if (K_[d] == 0) continue;
      M.V[d].alloc(N * K_[d]);
      HODLR_HIP_CHECK(hipMemsetAsync(M.V[d].p, 0, M.V[d].bytes(), stream_));
      kmax = std::max(kmax, K_[d]);
      float* W32 = M.W.p + N * off_[d];
      for (int id : level_nodes_[d]) {
        const TreeNodeH& nd = nodes_[id];
        const int st = nd.split - nd.min_i + 1, sb = nd.max_i - nd.split;
        const int top = nd.row0, bot = nd.row0 + st;
        // top block: U rows [top, top+st), V rows [bot, bot+sb); bottom the reverse
        bd.push_back({U_[d].p + top, V_[d].p + bot, W32 + top, M.V[d].p + bot, st, sb, nd.kt, ld,
                      ld, ld, ld, 0});
        bd.push_back({U_[d].p + bot, V_[d].p + top, W32 + bot, M.V[d].p + top, sb, st, nd.kb, ld,
                      ld, ld, ld, 0});
      }
    }
    launch_balance(arena_.push(bd, stream_), (int)bd.size(), kmax, stream_);
  }
  if (trace_) {
    std::fprintf(stderr, "[mixed] leaves: non-finite LU %lld, W %lld\n",
                 count_nonfinite(Lfac_.p, leaf_elems_), count_nonfinite(M.W.p, N * ktot_));
    for (int d = 0; d < nlev_; d++)
      if (K_[d] > 0)
        std::fprintf(stderr, "[mixed] level %d panels: non-finite U %lld, V %lld, V32 %lld\n", d,
                     count_nonfinite(U_[d].p, N * K_[d]), count_nonfinite(V_[d].p, N * K_[d]),
                     count_nonfinite(M.V[d].p, N * K_[d]));
  }
  const int nn = (int)nodes_.size();
  M.c_off.assign(nn, 0);
  M.p_off.assign(nn, 0);
  size_t co = 0, po = 0;
  for (int i = 0; i < nn; i++) {
    const int K = nodes_[i].kt + nodes_[i].kb;
    M.c_off[i] = co;
    M.p_off[i] = po;
    co += (size_t)K * K;
    po += K;
  }
  M.C.alloc(std::max<size_t>(co, 1));
  M.piv.alloc(std::max<size_t>(po, 1));
  float* W = M.W.p;
  for (int d = nlev_ - 1; d >= 0; d--) {
    const int offd = off_[d];
    if (offd > 0 && !level_leaves_[d].empty()) {
      // leaf solves in FP64: the leaves' condition numbers reach 1e6 on
      // smooth kernels with small shifts, and FP32 solves lose every digit
      // of the (much smaller) result
      size_t need = 0;
      for (int li : level_leaves_[d]) need += (size_t)leaves_[li].s * offd;
      grow(tmp2_, need);
      std::vector<ConvDesc<float, double>> up;
      std::vector<ConvDesc<double, float>> down;
      size_t to = 0;
      for (int li : level_leaves_[d]) {
        const LeafH& lf = leaves_[li];
        double* x64 = tmp2_.p + to;
        up.push_back({W + lf.row0, x64, lf.s, offd, ld, lf.s});
        down.push_back({x64, W + lf.row0, lf.s, offd, lf.s, ld});
        to += (size_t)lf.s * offd;
      }
      launch_convert(arena_.push(up, stream_), (int)up.size(), stream_);
      leaf_solves(level_leaves_[d], tmp2_.p, 0, offd, true);
      launch_convert(arena_.push(down, stream_), (int)down.size(), stream_);
      if (trace_)
        std::fprintf(stderr, "[mixed] level %d leaf apply: non-finite W %lld\n", d,
                     count_nonfinite(M.W.p, N * ktot_));
    }
    if (level_nodes_[d].empty()) continue;
    const int ncols = offd + K_[d];
    size_t zneed = 0, yneed = 0;
    for (int id : level_nodes_[d]) {
      const int K = nodes_[id].kt + nodes_[id].kb;
      zneed += (size_t)K * ncols;
      yneed += (size_t)K * std::max(offd, 1);
    }
    grow_t(t32a_, zneed);
    grow_t(t32b_, yneed);
    grow(tmp_, yneed);
    std::vector<GemmDescF> gz, gu;
    std::vector<CapDescF> cd;
    std::vector<LuDesc> lu;
    std::vector<SolveDesc> sd;
    std::vector<ConvDesc<double, float>> cq;
    size_t zo = 0, yo = 0;
    for (int id : level_nodes_[d]) {
      const TreeNodeH& nd = nodes_[id];
      const int kt = nd.kt, kb = nd.kb, K = kt + kb;
      const int st = nd.split - nd.min_i + 1, sb = nd.max_i - nd.split;
      const int top = nd.row0, bot = nd.row0 + st;
      float* Zt = t32a_.p + zo;
      float* Zb = Zt + (size_t)kb * ncols;
      zo += (size_t)K * ncols;
      gz.push_back(gemm_desc<float>(kb, ncols, st, 1, M.V[d].p + top, ld, W + top, ld, Zt, kb, 1.0f,
                                    0.0f));
      gz.push_back(gemm_desc<float>(kt, ncols, sb, 1, M.V[d].p + bot, ld, W + bot, ld, Zb, kt, 1.0f,
                                    0.0f));
      double* Cn = M.C.p + M.c_off[id];
      double* Y = tmp_.p + yo;
      float* Q = t32b_.p + yo;
      yo += (size_t)K * std::max(offd, 1);
      CapDescF c{};
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
      lu.push_back({Cn, M.piv.p + M.p_off[id], K, K});
      if (offd > 0) {
        SolveDesc sv{};
        sv.LU = Cn;
        sv.piv = M.piv.p + M.p_off[id];
        sv.B = Y;
        sv.n = K;
        sv.ld_lu = K;
        sv.ldb = K;
        sv.ncols = offd;
        sd.push_back(sv);
        cq.push_back({Y, Q, K, offd, K, K});
        const int qt = which == 0 ? 0 : kb;
        const int qb = which == 0 ? kt : 0;
        gu.push_back(gemm_desc<float>(st, offd, kt, 0, W + top + (size_t)offd * ld, ld, Q + qt, K,
                                      W + top, ld, -1.0f, 1.0f));
        gu.push_back(gemm_desc<float>(sb, offd, kb, 0, W + bot + (size_t)offd * ld, ld, Q + qb, K,
                                      W + bot, ld, -1.0f, 1.0f));
      }
    }
    run_gemm_t<float>(gz);
    launch_capacitance(arena_.push(cd, stream_), (int)cd.size(), stream_);
    lu_batch(lu);
    if (offd > 0) {
      lu_solve_batch(sd);
      launch_convert(arena_.push(cq, stream_), (int)cq.size(), stream_);
      run_gemm_t<float>(gu);
    }
    if (trace_)
      std::fprintf(stderr, "[mixed] level %d: non-finite W %lld, Z %lld, C %lld\n", d,
                   count_nonfinite(M.W.p, N * ktot_), count_nonfinite(t32a_.p, zneed),
                   count_nonfinite(M.C.p, co));
  }
  HODLR_HIP_CHECK(hipEventRecord(ev1_, stream_));
  HODLR_HIP_CHECK(hipEventSynchronize(ev1_));
  float ms = 0;
  HODLR_HIP_CHECK(hipEventElapsedTime(&ms, ev0_, ev1_));
  stats_.factor_time = ms * 1e-3;
  stats_.factor_stream_time = ms * 1e-3;
  factor_time_pending_ = false;
  arena_.reset();
  M.done = true;
}

/* One application of the mixed factorization: X <- M^-1 X (FP32, in place). */
void Engine::solve_mixed_once(int which, float* X, int nrhs) {
  MixedFactorization& M = mfac_[which];
  const int ld = ntot_;
  const size_t N = (size_t)ntot_;
  {
    // leaf solves in FP64 (see factorize_mixed)
    const size_t need = N * nrhs;
    grow(tmp2_, need);
    std::vector<ConvDesc<float, double>> up;
    std::vector<ConvDesc<double, float>> down;
    std::vector<int> all(leaves_.size());
    size_t to = 0;
    for (size_t i = 0; i < leaves_.size(); i++) {
      const LeafH& lf = leaves_[i];
      all[i] = (int)i;
      double* x64 = tmp2_.p + to;
      up.push_back({X + lf.row0, x64, lf.s, nrhs, ld, lf.s});
      down.push_back({x64, X + lf.row0, lf.s, nrhs, lf.s, ld});
      to += (size_t)lf.s * nrhs;
    }
    launch_convert(arena_.push(up, stream_), (int)up.size(), stream_);
    leaf_solves(all, tmp2_.p, 0, nrhs, true);
    launch_convert(arena_.push(down, stream_), (int)down.size(), stream_);
  }
  for (int d = nlev_ - 1; d >= 0; d--) {
    if (level_nodes_[d].empty()) continue;
    const int offd = off_[d];
    size_t zneed = 0;
    for (int id : level_nodes_[d]) zneed += (size_t)(nodes_[id].kt + nodes_[id].kb) * nrhs;
    grow_t(t32a_, zneed);
    grow_t(t32b_, zneed);
    grow(tmp_, zneed);
    std::vector<GemmDescF> gz, gu;
    std::vector<CapDescF> cd;
    std::vector<SolveDesc> sd;
    std::vector<ConvDesc<double, float>> cq;
    size_t zo = 0;
    for (int id : level_nodes_[d]) {
      const TreeNodeH& nd = nodes_[id];
      const int kt = nd.kt, kb = nd.kb, K = kt + kb;
      const int st = nd.split - nd.min_i + 1, sb = nd.max_i - nd.split;
      const int top = nd.row0, bot = nd.row0 + st;
      float* Zt = t32a_.p + zo;
      float* Zb = Zt + (size_t)kb * nrhs;
      double* Y = tmp_.p + zo;
      float* Q = t32b_.p + zo;
      zo += (size_t)K * nrhs;
      gz.push_back(gemm_desc<float>(kb, nrhs, st, 1, M.V[d].p + top, ld, X + top, ld, Zt, kb, 1.0f, 0.0f));
      gz.push_back(gemm_desc<float>(kt, nrhs, sb, 1, M.V[d].p + bot, ld, X + bot, ld, Zb, kt, 1.0f, 0.0f));
      CapDescF c{};
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
      SolveDesc sv{};
      sv.LU = M.C.p + M.c_off[id];
      sv.piv = M.piv.p + M.p_off[id];
      sv.B = Y;
      sv.n = K;
      sv.ld_lu = K;
      sv.ldb = K;
      sv.ncols = nrhs;
      sd.push_back(sv);
      cq.push_back({Y, Q, K, nrhs, K, K});
      const int qt = which == 0 ? 0 : kb;
      const int qb = which == 0 ? kt : 0;
      gu.push_back(gemm_desc<float>(st, nrhs, kt, 0, M.W.p + top + (size_t)offd * ld, ld, Q + qt, K,
                                    X + top, ld, -1.0f, 1.0f));
      gu.push_back(gemm_desc<float>(sb, nrhs, kb, 0, M.W.p + bot + (size_t)offd * ld, ld, Q + qb, K,
                                    X + bot, ld, -1.0f, 1.0f));
    }
    run_gemm_t<float>(gz);
    launch_capacitance(arena_.push(cd, stream_), (int)cd.size(), stream_);
    lu_solve_batch(sd);
    launch_convert(arena_.push(cq, stream_), (int)cq.size(), stream_);
    run_gemm_t<float>(gu);
  }
}

/* Refinement by restarted flexible GMRES (FGMRES-IR): solves A_H x = b with
 * the FP64 compressed operator A_H (leaves plus U V^T) and the preconditioner
 * M^-1 = the mixed-precision factorization (mixed) or the FP64 one, until
 * every column satisfies ||b - A_H x|| <= target ||b|| (true residual in
 * double-double arithmetic, checked at each restart) or the step budget is
 * spent.  The preconditioned vectors
 * z_j = M^-1 v_j are kept and the update is x += Z y: an FP32 preconditioner
 * is not exactly linear, and applying it again to V y would break the
 * agreement between the Arnoldi residual estimate and the true residual.
 * Columns run in lockstep: one preconditioner application and one operator
 * product per step serve all right-hand sides; orthogonalization is classical
 * Gram-Schmidt applied twice, per column. */
void Engine::solve_gmres(int which, bool mixed, const double* dB, int ldb, int nrhs, double* dX,
                         int ldx) {
  if (mixed)
    factorize_mixed(which);
  else
    factorize_fp64(which);
  HODLR_HIP_CHECK(hipEventRecord(ev0_, stream_));
  const int ld = ntot_;
  const size_t N = (size_t)ntot_;
  const int m = std::max(1, std::min(gmres_restart_, std::max(1, refine_max_)));
  grow(xw_, N * nrhs);
  grow(rbuf_, N * nrhs);
  grow(basis_, (size_t)(m + 1) * N * nrhs);
  grow(zbasis_, (size_t)m * N * nrhs);
  grow(nbuf_, (size_t)(m + 2) * nrhs + nrhs);
  if (mixed) grow_t(x32_, N * nrhs);
  const size_t vstride = N * nrhs;
  double* V = basis_.p;
  double* dcoef = nbuf_.p;            // (m + 1) x nrhs coefficients
  double* dnorm = nbuf_.p + (size_t)(m + 1) * nrhs;  // nrhs values
  // Norms per instance segment and column: every instance of a batch must
  // reach the target on its own (rows of instance q: inst_row0_[q] ..).
  const int nseg = (int)specs_.size();
  if (seg_rows_.n < (size_t)nseg + 1) {
    std::vector<int> r0(inst_row0_);
    r0.push_back(ntot_);
    seg_rows_.alloc(r0.size());
    HODLR_HIP_CHECK(hipMemcpy(seg_rows_.p, r0.data(), r0.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  // norms: enough workgroups to stream the vectors at bandwidth (parts
  // summed on the host in order)
  const int nsplit = std::max(1, std::min(16, 2048 / std::max(1, nseg * nrhs)));
  grow(segbuf_, (size_t)nseg * nrhs * nsplit);
  auto colnorms = [&](const double* X, std::vector<double>& out) {
    launch_colnorm2(X, ntot_, ld, nrhs, dnorm, stream_);
    out.resize(nrhs);
    HODLR_HIP_CHECK(hipMemcpyAsync(out.data(), dnorm, nrhs * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
    for (auto& v : out) v = std::sqrt(v);
  };
  std::vector<double> parts;
  auto segnorms = [&](const double* X, std::vector<double>& out) {
    launch_segnorm2(X, ld, seg_rows_.p, nseg, nrhs, nsplit, segbuf_.p, stream_);
    parts.resize((size_t)nseg * nrhs * nsplit);
    HODLR_HIP_CHECK(hipMemcpyAsync(parts.data(), segbuf_.p, parts.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
    out.assign((size_t)nseg * nrhs, 0.0);
    for (size_t e = 0; e < out.size(); e++) {
      double v = 0.0;
      for (int p = 0; p < nsplit; p++) v += parts[e * nsplit + p];
      out[e] = std::sqrt(v);
    }
  };
  auto upload = [&](const std::vector<double>& h, double* d) {
    HODLR_HIP_CHECK(hipMemcpyAsync(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, stream_));
    HODLR_HIP_CHECK(hipStreamSynchronize(stream_));  // h is reused right away
  };
  auto precond = [&](const double* in, double* out) {  // out = M^-1 in
    if (mixed) {
      std::vector<ConvDesc<double, float>> cv{{in, x32_.p, ntot_, nrhs, ld, ld}};
      launch_convert(arena_.push(cv, stream_), 1, stream_);
      solve_mixed_once(which, x32_.p, nrhs);
      std::vector<ConvDesc<float, double>> cb{{x32_.p, out, ntot_, nrhs, ld, ld}};
      launch_convert(arena_.push(cb, stream_), 1, stream_);
    } else {
      if (in != out)
        HODLR_HIP_CHECK(hipMemcpyAsync(out, in, N * nrhs * sizeof(double), hipMemcpyDeviceToDevice, stream_));
      if (use_graphs_)
        solve_fp64_graph(which, out, nrhs);
      else
        solve_fp64_inplace(which, out, nrhs);
    }
  };
  std::vector<double> bn, rs, rn(nrhs), hn, coef((size_t)(m + 1) * nrhs);
  HODLR_HIP_CHECK(hipMemsetAsync(xw_.p, 0, N * nrhs * sizeof(double), stream_));
  HODLR_HIP_CHECK(hipMemcpy2DAsync(rbuf_.p, N * sizeof(double), dB, (size_t)ldb * sizeof(double),
                                   N * sizeof(double), nrhs, hipMemcpyDeviceToDevice, stream_));
  segnorms(rbuf_.p, bn);
  for (auto& v : bn)
    if (!(v > 0)) v = 1.0;
  // A batch is solved with every instance's right-hand side scaled to unit
  // norm (the Krylov minimization is over all instances at once) and the
  // solution scaled back at the end.
  const double* Beff = dB;
  int ldbeff = ldb;
  const std::vector<double> bscale = bn;
  if (nseg > 1) {
    grow(bscaled_, N * nrhs);
    HODLR_HIP_CHECK(hipMemcpyAsync(bscaled_.p, rbuf_.p, N * nrhs * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    std::vector<double> inv(bn.size());
    for (size_t e = 0; e < bn.size(); e++) inv[e] = 1.0 / bn[e];
    upload(inv, segbuf_.p);
    launch_segscale(bscaled_.p, ld, seg_rows_.p, nseg, nrhs, segbuf_.p, stream_);
    HODLR_HIP_CHECK(hipMemcpyAsync(rbuf_.p, bscaled_.p, N * nrhs * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    Beff = bscaled_.p;
    ldbeff = ld;
    segnorms(rbuf_.p, bn);
    for (auto& v : bn)
      if (!(v > 0)) v = 1.0;
  }
  // The best iterate of every instance column (lowest true relative
  // residual seen), returned where the refinement ends above its target
  // with a worse one.
  grow(xbest_, N * nrhs);
  grow_t(bestmask_, (size_t)nseg * nrhs);
  std::vector<double> best((size_t)nseg * nrhs, HUGE_VAL);
  std::vector<int> hmask((size_t)nseg * nrhs, 0);
  auto masked_copy = [&](const double* from, double* to) {
    HODLR_HIP_CHECK(hipMemcpyAsync(bestmask_.p, hmask.data(), hmask.size() * sizeof(int),
                                   hipMemcpyHostToDevice, stream_));
    launch_segcopy(from, to, ld, seg_rows_.p, nseg, nrhs, nsplit, bestmask_.p, stream_);
    HODLR_HIP_CHECK(hipStreamSynchronize(stream_));  // hmask is reused
  };
  auto note_best = [&](const std::vector<double>& r) {  // r: segnorms of the residual of xw_
    bool any = false;
    for (size_t e = 0; e < r.size(); e++) {
      const double rc = r[e] / bn[e];
      hmask[e] = rc < best[e];
      if (hmask[e]) {
        best[e] = rc;
        any = true;
      }
    }
    if (any) masked_copy(xw_.p, xbest_.p);
  };
  // Hessenberg, Givens rotations and right-hand side per column (host)
  std::vector<double> H((size_t)(m + 1) * m * nrhs), cs((size_t)m * nrhs), sn((size_t)m * nrhs),
      g((size_t)(m + 1) * nrhs);
  auto Hc = [&](int c) { return H.data() + (size_t)c * (m + 1) * m; };
  int total = 0;
  double res = 0.0;
  // With the FP64 factorization as preconditioner, plain refinement
  // x += M^-1 r, one double-double residual per step, contracts by orders of
  // magnitude per step (M^-1 solves A_H to about 1e-6 relative on the hard
  // systems) and needs one preconditioner application and one residual per
  // step; a Krylov step also spends an FP64 product and orthogonalization,
  // and a cycle's FP64 products cannot see below their own evaluation floor,
  // so it needs a second cycle after the first residual.  Refinement runs
  // while every step gains at least a factor 8; the flexible GMRES cycles
  // below continue from its iterate otherwise (and always serve the FP32
  // factorization, which contracts too little).
  if (!mixed && refine_max_ > 0) {
    std::vector<double> ones(nrhs, 1.0);
    upload(ones, dcoef);
    double prev = 0.0;
    segnorms(rbuf_.p, rs);
    note_best(rs);
    for (size_t e = 0; e < rs.size(); e++) prev = std::max(prev, rs[e] / bn[e]);
    std::vector<double> dn, xn;
    while (prev > refine_target_ && total < refine_max_) {
      precond(rbuf_.p, zbasis_.p);  // the step d = M^-1 r
      // A step small against the iterate (every instance and column) updates
      // the residual exactly enough in FP64: r_new = r - A_H (x_new - x_old),
      // with x_new - x_old = d - e from the TwoSum x_old + d = x_new + e.  The
      // FP64 product of the step errs by its size relative to x times the
      // FP64 evaluation floor of A_H x (~1e-10 at N = 262144), i.e. by
      // 1e-3 of that floor at most; r itself came exact from double-double
      // arithmetic (the first step, from x = 0, always takes that path).
      segnorms(zbasis_.p, dn);
      segnorms(xw_.p, xn);
      double ratio = 0.0;
      for (size_t e = 0; e < dn.size(); e++)
        ratio = std::max(ratio, xn[e] > 0 ? dn[e] / xn[e] : (dn[e] > 0 ? HUGE_VAL : 0.0));
      const bool update = ratio <= 1e-3 && refine_target_ >= 1e-12;
      if (update) {
        launch_step_update(xw_.p, zbasis_.p, ld, ntot_, nrhs, stream_);
        matvec_device(zbasis_.p, ld, nrhs, V, ld);
        launch_multi_axpy(V, vstride, 1, dcoef, rbuf_.p, ntot_, ld, nrhs, -1.0, 0, stream_);
      } else {
        launch_multi_axpy(zbasis_.p, vstride, 1, dcoef, xw_.p, ntot_, ld, nrhs, 1.0, 0, stream_);
        residual_dd(Beff, ldbeff, xw_.p, ld, nrhs, rbuf_.p, ld);
      }
      total++;
      segnorms(rbuf_.p, rs);
      note_best(rs);
      double now = 0.0;
      for (size_t e = 0; e < rs.size(); e++) now = std::max(now, rs[e] / bn[e]);
      if (trace_)
        std::fprintf(stderr, "[refine] step %d true residual %.3e (%s, step/iterate %.1e)\n", total,
                     now, update ? "FP64 update" : "double-double", ratio);
      if (!std::isfinite(now)) throw std::runtime_error("refinement produced a non-finite residual");
      const bool gained = now <= prev / 8;
      prev = now;
      if (!gained) break;
    }
  }
  for (;;) {
    segnorms(rbuf_.p, rs);
    note_best(rs);
    res = 0.0;
    std::vector<char> active(nrhs, 0);
    for (int c = 0; c < nrhs; c++) {
      double s2 = 0.0;
      for (int q = 0; q < nseg; q++) {
        const size_t e = (size_t)q + (size_t)c * nseg;
        const double rc = rs[e] / bn[e];
        if (!(rc <= res)) res = rc;
        if (rc > refine_target_) active[c] = 1;
        s2 += rs[e] * rs[e];
      }
      rn[c] = std::sqrt(s2);
    }
    if (trace_) std::fprintf(stderr, "[gmres] steps %d true residual %.3e\n", total, res);
    if (!std::isfinite(res)) throw std::runtime_error("refinement produced a non-finite residual");
    if (res <= refine_target_ || total >= refine_max_) break;
    // V_0 = r / ||r||; converged columns get a zero basis and stay put
    std::vector<double> inv(nrhs);
    for (int c = 0; c < nrhs; c++) {
      active[c] = active[c] && rn[c] > 0;
      inv[c] = active[c] ? 1.0 / rn[c] : 0.0;
      g[(size_t)c * (m + 1)] = active[c] ? rn[c] : 0.0;
      for (int i = 1; i <= m; i++) g[(size_t)c * (m + 1) + i] = 0.0;
    }
    HODLR_HIP_CHECK(hipMemcpyAsync(V, rbuf_.p, N * nrhs * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    upload(inv, dnorm);
    launch_col_scale(V, ld, ntot_, nrhs, dnorm, stream_);
    int jend = 0;
    for (int j = 0; j < m; j++) {
      double* Vj = V + (size_t)j * vstride;
      double* Vn = V + (size_t)(j + 1) * vstride;
      double* Zj = zbasis_.p + (size_t)j * vstride;
      precond(Vj, Zj);
      matvec_device(Zj, ld, nrhs, Vn, ld);
      // two passes of classical Gram-Schmidt against V_0..V_j
      std::vector<double> h((size_t)(j + 1) * nrhs, 0.0);
      for (int pass = 0; pass < 2; pass++) {
        launch_multi_dot(V, vstride, j + 1, Vn, ntot_, ld, nrhs, dcoef, stream_);
        HODLR_HIP_CHECK(hipMemcpyAsync(coef.data(), dcoef, (size_t)(j + 1) * nrhs * sizeof(double),
                                       hipMemcpyDeviceToHost, stream_));
        HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
        launch_multi_axpy(V, vstride, j + 1, dcoef, Vn, ntot_, ld, nrhs, -1.0, 0, stream_);
        for (size_t e = 0; e < h.size(); e++) h[e] += coef[e];
      }
      colnorms(Vn, hn);
      for (int c = 0; c < nrhs; c++) {
        inv[c] = (active[c] && hn[c] > 0) ? 1.0 / hn[c] : 0.0;
        double* Hj = Hc(c) + (size_t)j * (m + 1);  // column j of H, m+1 rows
        for (int i = 0; i <= j; i++) Hj[i] = active[c] ? h[i + (size_t)c * (j + 1)] : 0.0;
        Hj[j + 1] = active[c] ? hn[c] : 0.0;
        if (!active[c]) {
          cs[(size_t)c * m + j] = 1.0;
          sn[(size_t)c * m + j] = 0.0;
          Hj[j] = 1.0;  // keeps the triangular solve regular; g stays 0
          continue;
        }
        for (int i = 0; i < j; i++) {  // previous rotations
          const double a = Hj[i], b = Hj[i + 1];
          const double cc = cs[(size_t)c * m + i], ss = sn[(size_t)c * m + i];
          Hj[i] = cc * a + ss * b;
          Hj[i + 1] = -ss * a + cc * b;
        }
        const double a = Hj[j], b = Hj[j + 1];
        const double r = std::hypot(a, b);
        const double cc = r > 0 ? a / r : 1.0, ss = r > 0 ? b / r : 0.0;
        cs[(size_t)c * m + j] = cc;
        sn[(size_t)c * m + j] = ss;
        Hj[j] = r;
        Hj[j + 1] = 0.0;
        double* gc = g.data() + (size_t)c * (m + 1);
        gc[j + 1] = -ss * gc[j];
        gc[j] = cc * gc[j];
        if (!(hn[c] > 0)) active[c] = 0;  // happy breakdown: exact in this space
      }
      upload(inv, dnorm);
      launch_col_scale(Vn, ld, ntot_, nrhs, dnorm, stream_);
      total++;
      jend = j + 1;
      // Arnoldi estimate of the column residual, relative to the smallest
      // instance right-hand side of the column
      double est = 0.0;
      for (int c = 0; c < nrhs; c++) {
        double bmin = bn[(size_t)c * nseg];
        for (int q = 1; q < nseg; q++) bmin = std::min(bmin, bn[(size_t)q + (size_t)c * nseg]);
        est = std::max(est, std::fabs(g[(size_t)c * (m + 1) + j + 1]) / bmin);
      }
      if (trace_) std::fprintf(stderr, "[gmres] step %d estimate %.3e\n", total, est);
      if (est <= 0.5 * refine_target_ || total >= refine_max_) break;
    }
    // y = H^-1 g per column, x += Z y
    std::fill(coef.begin(), coef.end(), 0.0);
    for (int c = 0; c < nrhs; c++) {
      const double* Hm = Hc(c);
      const double* gc = g.data() + (size_t)c * (m + 1);
      double* y = coef.data() + (size_t)c * jend;
      for (int i = jend - 1; i >= 0; i--) {
        double v = gc[i];
        for (int k = i + 1; k < jend; k++) v -= Hm[(size_t)k * (m + 1) + i] * y[k];
        const double d = Hm[(size_t)i * (m + 1) + i];
        y[i] = d != 0.0 ? v / d : 0.0;
      }
    }
    std::vector<double> yv(coef.begin(), coef.begin() + (size_t)jend * nrhs);
    upload(yv, dcoef);
    launch_multi_axpy(zbasis_.p, vstride, jend, dcoef, xw_.p, ntot_, ld, nrhs, 1.0, 0, stream_);
    residual_dd(Beff, ldbeff, xw_.p, ld, nrhs, rbuf_.p, ld);
  }
  {  // instance columns whose last iterate is worse than their best take the best
    bool any = false;
    res = 0.0;
    for (size_t e = 0; e < rs.size(); e++) {
      const double rc = rs[e] / bn[e];
      hmask[e] = best[e] < rc;
      any = any || hmask[e];
      const double r = std::min(rc, best[e]);
      if (!(r <= res)) res = r;
    }
    if (any) masked_copy(xbest_.p, xw_.p);
  }
  if (nseg > 1) {  // undo the right-hand side scaling
    upload(bscale, segbuf_.p);
    launch_segscale(xw_.p, ld, seg_rows_.p, nseg, nrhs, segbuf_.p, stream_);
  }
  stats_.refine_iterations = total;
  stats_.last_residual = res;
  HODLR_HIP_CHECK(hipMemcpy2DAsync(dX, (size_t)ldx * sizeof(double), xw_.p, N * sizeof(double),
                                   N * sizeof(double), nrhs, hipMemcpyDeviceToDevice, stream_));
  HODLR_HIP_CHECK(hipEventRecord(ev1_, stream_));
  HODLR_HIP_CHECK(hipEventSynchronize(ev1_));
  float ms = 0;
  HODLR_HIP_CHECK(hipEventElapsedTime(&ms, ev0_, ev1_));
  stats_.solve_time = ms * 1e-3;
  stats_.refine_time = ms * 1e-3;
  arena_.reset();
  if (!(res <= refine_target_)) {
    char msg[160];
    std::snprintf(msg, sizeof msg,
                  "refinement stopped at relative residual %.3e above the target %.3e after %d "
                  "steps (the best iterate is returned)",
                  res, refine_target_, total);
    throw NotConverged(msg);
  }
}

/* calcDeterminant (HODLR_Matrix.cpp:1671-1697): sum of log|diag LU| over the
 * leaves and the recSM capacitance matrices (perturbI.cpp:61-73). */
/* Per factor (leaves, then nodes) log|det| and product of the LU diagonal of
 * the FP64 recSM factorization; inst receives the owning instance. */
void Engine::logdet_terms(std::vector<double>& hl, std::vector<double>& hp, std::vector<int>& inst) {
  factorize_fp64(0);
  Factorization& F = fac_[0];
  wait_root(F);
  std::vector<LuDesc> ld;
  inst.clear();
  for (auto& lf : leaves_) {
    ld.push_back({Lfac_.p + lf.mat_off, nullptr, lf.s, lf.s});
    inst.push_back(lf.inst);
  }
  for (int i = 0; i < (int)nodes_.size(); i++) {
    const int K = nodes_[i].kt + nodes_[i].kb;
    ld.push_back({F.C.p + F.c_off[i], nullptr, K, K});
    inst.push_back(nodes_[i].inst);
  }
  DBuf<double> lg, pr;
  lg.alloc(ld.size());
  pr.alloc(ld.size());
  launch_logdet(arena_.push(ld, stream_), (int)ld.size(), lg.p, pr.p, stream_);
  hl.assign(ld.size(), 0.0);
  hp.assign(ld.size(), 0.0);
  HODLR_HIP_CHECK(hipMemcpyAsync(hl.data(), lg.p, hl.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
  HODLR_HIP_CHECK(hipMemcpyAsync(hp.data(), pr.p, hp.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
  HODLR_HIP_CHECK(hipStreamSynchronize(stream_));
  arena_.reset();
}

/* logAbsDeterminant (HODLR_Matrix.cpp:1666-1697): sum over leaf LUs and node
 * capacitance LUs (Sylvester's identity, perturbI.cpp:61-83).  Always from
 * the FP64 factorization. */
double Engine::logabsdet(double* det_prod) {
  std::vector<double> hl, hp;
  std::vector<int> inst;
  logdet_terms(hl, hp, inst);
  double s = 0, p = 1;
  for (size_t i = 0; i < hl.size(); i++) {
    s += hl[i];
    p *= hp[i];
  }
  if (det_prod) *det_prod = p;
  return s;
}

std::vector<double> Engine::logabsdet_instances() {
  std::vector<double> hl, hp;
  std::vector<int> inst;
  logdet_terms(hl, hp, inst);
  std::vector<double> out(specs_.size(), 0.0);
  for (size_t i = 0; i < hl.size(); i++) out[inst[i]] += hl[i];
  return out;
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

/* R = B - A_H X in double-double arithmetic, rounded once (ddres.h): the
 * leaves' products first, then per level the projections T = V^T X of every
 * node half (one launch for all levels, chunk partials reduced in order) and
 * the updates Y += U T. */
void Engine::residual_dd(const double* dB, int ldb, const double* dX, int ldx, int nrhs, double* dR,
                         int ldr) {
  if (nrhs <= 0) return;
  compress();
  const size_t N = (size_t)ntot_;
  const int ld = ntot_;
  grow(ddy_, 2 * N * nrhs);
  double* yh = ddy_.p;
  double* yl = ddy_.p + N * nrhs;
  // projections of all levels
  std::vector<DdProjDesc> pd;
  std::vector<std::pair<int, int>> pd_node;  // (node, half) of each projection
  long long poff = 0, toff = 0;
  int pwg = 0, rwg = 0;
  for (int id = 0; id < (int)nodes_.size(); id++) {
    const TreeNodeH& nd = nodes_[id];
    const int st = nd.split - nd.min_i + 1, sb = nd.max_i - nd.split;
    const int top = nd.row0, bot = nd.row0 + st;
    for (int h = 0; h < 2; h++) {
      const int k = h == 0 ? nd.kt : nd.kb;
      if (k == 0) continue;
      DdProjDesc p{};
      // top block U_t V_t^T: V_t on the bottom rows; bottom block the reverse
      p.V = V_[nd.level].p + (h == 0 ? bot : top);
      p.X = dX + (h == 0 ? bot : top);
      p.ldv = ld;
      p.ldx = ldx;
      p.rows = h == 0 ? sb : st;
      p.k = k;
      p.nrhs = nrhs;
      p.nchunks = (p.rows + DD_CHUNK - 1) / DD_CHUNK;
      p.wg0 = pwg;
      p.rwg0 = rwg;
      p.part_off = poff;
      p.t_off = toff;
      pwg += dd_proj_wgs(p);
      rwg += dd_reduce_wgs(p);
      poff += (long long)p.nchunks * k * nrhs;
      toff += (long long)k * nrhs;
      pd.push_back(p);
      pd_node.push_back({id, h});
    }
  }
  grow(ddpart_, 2 * std::max<long long>(poff, 1));
  grow(ddt_, 2 * std::max<long long>(toff, 1));
  double* ph = ddpart_.p;
  double* pl = ddpart_.p + std::max<long long>(poff, 1);
  double* th = ddt_.p;
  double* tl = ddt_.p + std::max<long long>(toff, 1);
  if (!pd.empty()) {
    const DdProjDesc* dpd = arena_.push(pd, stream_);
    launch_dd_proj(dpd, (int)pd.size(), pwg, nrhs, ph, pl, stream_);
    launch_dd_reduce(dpd, (int)pd.size(), rwg, ph, pl, th, tl, stream_);
  }
  // leaves: Y = L X
  {
    std::vector<DdApplyDesc> ad;
    int wg = 0;
    for (auto& lf : leaves_) {
      DdApplyDesc a{};
      a.A = Lorig_.p + lf.mat_off;
      a.lda = lf.s;
      a.t_hi = dX + lf.row0;
      a.t_lo = nullptr;
      a.ldt = ldx;
      a.y_hi = yh + lf.row0;
      a.y_lo = yl + lf.row0;
      a.ldy = ld;
      a.rows = lf.s;
      a.k = lf.s;
      a.nrhs = nrhs;
      a.accumulate = 0;
      a.wg0 = wg;
      wg += dd_apply_wgs(a);
      ad.push_back(a);
    }
    launch_dd_apply(arena_.push(ad, stream_), (int)ad.size(), wg, nrhs, stream_);
  }
  // low-rank updates, level by level (levels overlap in rows)
  for (int d = 0; d < nlev_; d++) {
    std::vector<DdApplyDesc> ad;
    int wg = 0;
    for (size_t q = 0; q < pd.size(); q++) {
      const TreeNodeH& nd = nodes_[pd_node[q].first];
      if (nd.level != d) continue;
      const int h = pd_node[q].second;
      const int st = nd.split - nd.min_i + 1, sb = nd.max_i - nd.split;
      const int row = h == 0 ? nd.row0 : nd.row0 + st;
      DdApplyDesc a{};
      a.A = U_[d].p + row;
      a.lda = ld;
      a.t_hi = th + pd[q].t_off;
      a.t_lo = tl + pd[q].t_off;
      a.ldt = pd[q].k;
      a.y_hi = yh + row;
      a.y_lo = yl + row;
      a.ldy = ld;
      a.rows = h == 0 ? st : sb;
      a.k = pd[q].k;
      a.nrhs = nrhs;
      a.accumulate = 1;
      a.wg0 = wg;
      wg += dd_apply_wgs(a);
      ad.push_back(a);
    }
    if (!ad.empty()) launch_dd_apply(arena_.push(ad, stream_), (int)ad.size(), wg, nrhs, stream_);
  }
  launch_dd_residual(dB, ldb, yh, yl, ld, dR, ldr, ntot_, nrhs, stream_);
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



}  // namespace hodlr
