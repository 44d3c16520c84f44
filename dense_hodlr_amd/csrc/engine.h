/*
 * Host orchestration of the HODLR engine.
 *
 * A forest of one or more independent instances (each with its own tree and
 * entry source) shares level-batched launches: every kernel processes all
 * blocks, nodes or leaves of one tree level of every instance at once.
 *
 * Device layout (column-major FP64, leading dimension Ntot = total rows):
 *  U[d], V[d]  level panels from compression.  For a node at level d, rows of
 *              its top half hold U_t (top block rows) and V_b (bottom block
 *              columns); rows of its bottom half hold U_b and V_t.
 *  W (per factorization)  Ntot x Ktot: the U^SM columns of every level side
 *              by side, level d at columns [off[d], off[d] + K[d]).  The
 *              recSM factorization overwrites W[rows(D), 0:off[d]] by
 *              D^-1 W[rows(D), 0:off[d]] for every node or leaf D of level d,
 *              bottom-up (src/HODLR_Matrix.cpp:537-592).
 *  leaves      s x s blocks, original, LU-factored and inverted copies.
 *  node C      (kt+kb)^2 capacitance (recSM) or Schur (recLU) matrices, LU.
 */
#ifndef DENSE_HODLR_ENGINE_H
#define DENSE_HODLR_ENGINE_H

#include <cstring>
#include <deque>
#include <memory>
#include <string>
#include <vector>

#include "aca.h"
#include "common.h"
#include "ddres.h"
#include "dense.h"
#include "gemm.h"
#include "pool.h"

namespace hodlr {

constexpr int LU_NB = 64;  // largest panel width of the blocked LU and triangular solves

/* GMRES restart length of the iterative solver (the reference's, HODLR_Matrix.cpp:1181). */
constexpr int ITER_RESTART = 32;

/* Size of the per-GPU device array behind Engine::matvec_noalloc. */
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;

/* Launch shapes chosen by the autotuner (defaults without it). */
struct Tuning {
  int aca_pt = 2;          // panel rows per thread of the ACA row/column kernels (1, 2, 4)
  int gemm_bm = 0;         // grouped GEMM tile rows (64, 128; 0: chosen per launch)
  int gemm_target = 2048;  // tiles a grouped GEMM aims for before splitting K
  int lu_nb = 64;          // panel width of the batched LU and triangular solves (<= 64)
  int gemm_mfma = 1;       // grouped GEMMs (FP64 and FP32) on the matrix cores
  int leaf_solve = 3;      // leaf solves of >= 64 columns (the factorization's and the solves
                           // of many right-hand sides): 0 one wave per 8-column chunk,
                           // 1 blocked (diagonal-block substitution + GEMM updates), 2 the
                           // factor shared by a workgroup's waves, 3 on the matrix cores
                           // (leaves of at most 256 rows; else as 0), 4 one wave per
                           // 16-column chunk (leaves of at most 256 rows; else as 0)
  bool tuned = false;
};

template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DBuf() { release(); }
  void alloc(size_t cnt) {
    release();
    if (cnt) p = static_cast<T*>(pool_alloc(cnt * sizeof(T)));
    n = cnt;
  }
  void release() {
    pool_release(p);
    p = nullptr;
    n = 0;
  }
  size_t bytes() const { return n * sizeof(T); }
};

/* Bump allocator for per-phase descriptor arrays (device copies of small host
 * vectors).  Host vectors are kept until reset() so asynchronous copies from
 * them stay valid. */
class DescArena {
 public:
  template <class T>
  T* push(const std::vector<T>& v, hipStream_t s);
  void reset();
  size_t bytes() const;
  /* Replay mode (for stream capture): push() copies nothing and returns, in
   * order, the device pointers of the pushes made since the last reset,
   * after checking that the contents are the same. */
  void begin_replay() {
    replay_ = true;
    next_ = 0;
  }
  void end_replay() { replay_ = false; }

 private:
  struct Chunk {
    DBuf<char> buf;
    size_t used = 0;
  };
  std::deque<Chunk> chunks_;
  std::deque<std::vector<char>> keep_;
  std::vector<char*> dev_;  // device copy of keep_[i]
  bool replay_ = false;
  size_t next_ = 0;
};

template <class T>
inline T* DescArena::push(const std::vector<T>& v, hipStream_t s) {
  const size_t bytes = v.size() * sizeof(T);
  const size_t need = (bytes + 255) & ~size_t(255);
  if (need == 0) return nullptr;
  if (replay_) {
    if (next_ >= keep_.size() || keep_[next_].size() != bytes ||
        std::memcmp(keep_[next_].data(), v.data(), bytes) != 0)
      throw std::runtime_error("captured solve does not match its recording");
    return reinterpret_cast<T*>(dev_[next_++]);
  }
  if (chunks_.empty() || chunks_.back().used + need > chunks_.back().buf.n) {
    Chunk c;
    c.buf.alloc(std::max<size_t>(need, 8u << 20));
    chunks_.push_back(std::move(c));
  }
  Chunk& c = chunks_.back();
  keep_.emplace_back(reinterpret_cast<const char*>(v.data()),
                     reinterpret_cast<const char*>(v.data()) + bytes);
  char* dst = c.buf.p + c.used;
  dev_.push_back(dst);
  HODLR_HIP_CHECK(hipMemcpyAsync(dst, keep_.back().data(), bytes, hipMemcpyHostToDevice, s));
  c.used += need;
  return reinterpret_cast<T*>(dst);
}

/* A captured hipGraph of the FP64 recSM/recLU solve for one (factorization,
 * number of right-hand sides), working in place on the engine's buffer gx_. */
struct SolveGraph {
  bool valid = false;
  int which = -1, nrhs = 0;
  unsigned long long gen = 0;  // work-buffer generation it was captured with
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  DescArena arena;  // device descriptors the captured kernels read
  void release() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    valid = false;
  }
  ~SolveGraph() { release(); }
};

/* A node of a user-defined partition (user_IndexTree::node,
 * include/user_IndexTree.hpp:10-24), children by index into the node list
 * (-1: none).  A node with split -1 and at most leaf rows is a leaf; a node
 * without children gets default subtrees below its split
 * (HODLR_Tree.cpp:207-321). */
struct UserNode {
  int split = -1;
  int min_rank[2] = {-1, -1};  // top, bottom off-diagonal block
  int max_rank[2] = {-1, -1};
  int lr = -1;                 // 0 partial-pivot ACA, 1 SVD, -1 the handle's method
  int left = -1, right = -1;
};

struct InstanceSpec {
  int n = 0;
  int leaf = 30;
  int kind = 0;  // 0 kernel, 1 dense
  HodlrKernelParams kp{};
  std::vector<double> xs, ys, zs;  // coordinates (kernel)
  std::vector<double> dense;       // n x n (dense)
  std::vector<UserNode> tree;      // user partition (root first); empty: the default tree
};

struct TreeNodeH {
  int inst, level;
  int min_i, split, max_i;  // instance-local
  int row0;                 // global row of min_i
  int kt = 0, kb = 0;
  int lr = -1;                   // low-rank method of this node (-1: the handle's)
  int min_rank[2] = {-1, -1};    // rank bounds of the top and bottom blocks
  int max_rank[2] = {-1, -1};
};

struct LeafH {
  int inst, level, min_i, max_i, row0, s;
  size_t mat_off;  // offset of the s x s block in the leaf buffers
  size_t piv_off;
};

/* Mixed-precision factorization: FP32 U^SM and V panels for the bulk GEMMs,
 * FP64 node matrices (LU) and leaf LU. */
struct MixedFactorization {
  bool done = false;
  DBuf<float> W;
  std::vector<DBuf<float>> V;
  DBuf<double> C;
  DBuf<int> piv;
  std::vector<size_t> c_off;
  std::vector<size_t> p_off;
};

struct Factorization {
  bool done = false;
  bool root_pending = false;  // the roots' capacitance LU may still run on aux_
  DBuf<double> W;
  DBuf<double> C;
  DBuf<int> piv;
  DBuf<int> perm;  // node matrices' interchanges as permutations (matrix-core solves)
  std::vector<char> perm_level;  // per level: perm holds its nodes' permutations
  std::vector<size_t> c_off;   // per node (global node order)
  std::vector<size_t> p_off;
};

class Engine {
 public:
  explicit Engine(std::vector<InstanceSpec> specs);
  ~Engine();

  // settings
  void set_lr_method(int m);
  void set_tolerance(double t);
  void set_min_pivot(double p);
  void set_precision(int p);
  void set_refinement(double target, int max_iters);
  void set_initial_rank_capacity(int cap) { init_cap_ = cap; }

  // compute
  void compress();
  void factorize(int which);
  void solve_device(int which, const double* dB, int ldb, int nrhs, double* dX, int ldx);
  void solve_host(int which, const double* B, int ldb, int nrhs, double* X, int ldx);
  double logabsdet(double* det_prod);
  void matvec_device(const double* dX, int ldx, int nrhs, double* dY, int ldy);
  /* The same product with no device allocation: temporaries in the per-GPU
   * scratch array (NOALLOC_SCRATCH_BYTES); needs a compressed operator. */
  void matvec_noalloc(const double* dX, int ldx, int nrhs, double* dY, int ldy);
  /* Looks up the current device's scratch address (once per device). */
  static void noalloc_setup();
  void residual_dd(const double* dB, int ldb, const double* dX, int ldx, int nrhs, double* dR,
                   int ldr);
  /* Y = A X with the exact operator: the dense source, or the kernel
   * evaluated on the fly (per instance). */
  void exact_apply(const double* dX, int ldx, int nrhs, double* dY, int ldy);
  /* GMRES(ITER_RESTART) on the exact operator, left-preconditioned by the
   * FP64 recSM (which 0) or recLU (1) solve, from x0 = M^-1 b, until every
   * column's preconditioned relative residual is below tol or max_iters
   * iterations are spent (then NotConverged, the iterate in dX). */
  void iterative_solve(int which, const double* dB, int ldb, int nrhs, double* dX, int ldx,
                       int max_iters, double tol);

  // inspection
  int rows() const { return ntot_; }
  int num_instances() const { return (int)specs_.size(); }
  const std::vector<int>& instance_rows() const { return inst_row0_; }
  std::vector<double> logabsdet_instances();
  int num_levels() const { return nlev_; }
  const std::vector<TreeNodeH>& nodes() const { return nodes_; }
  const std::vector<LeafH>& leaves() const { return leaves_; }
  void get_factor(int idx, int which, double* host_out);
  /* Entries of rows [i0, i0+m) and columns [j0, j0+n): which 0 the source
   * matrix (dense copy or kernel), 1 the compressed operator A_H; output
   * column-major with leading dimension ldo (device or host pointer). */
  void block_device(int which, int i0, int j0, int m, int n, double* dOut, int ldo);
  void block_host(int which, int i0, int j0, int m, int n, double* out, int ldo);
  /* The root's top (side 0) or bottom (side 1) diagonal block as a matrix
   * of its own, tree part and compressed factors kept. */
  std::unique_ptr<Engine> diagonal_block(int side);
  struct Stats {
    int aca_iterations = 0;
    // factor_time: from the factorization's first kernel to its last one on
    // either stream (the roots' LU on the side stream included);
    // factor_stream_time: up to the last one on the engine's stream, where a
    // following solve starts and overlaps the roots' LU
    double compress_time = 0, factor_time = 0, factor_stream_time = 0, solve_time = 0,
           refine_time = 0;
    int refine_iterations = 0;
    double last_residual = 0;
    double aca_row_time = 0, aca_col_time = 0;  // with profiling on
    int aca_launches = 0;
    double aca_bytes = 0;
    int graph_captures = 0, graph_launches = 0;
    int iter_iterations = 0;         // iterative solve: GMRES iterations,
    double iter_precond_residual = 0;  // final ||M^-1 (b - A x)|| / ||M^-1 b||,
    double iter_residual = 0;          // and ||b - A x|| / ||b|| (largest column)
    // ACA panel growths (a level's U/V panels reallocated and copied when a
    // block outgrew their capacity) and the panels' final bytes
    int aca_panel_grows = 0;
    double aca_panel_bytes = 0;
    int pred_blocks = 0, pred_misses = 0;  // with a rank predictor attached
    double pred_log_err = 0;               // mean |log predicted - log ACA rank|
    // with profiling on, the last FP64 factorization by kernel group
    // (FAC_GEMM, FAC_LU, FAC_LEAF_SOLVE, FAC_NODE_SOLVE): summed HIP-event
    // durations, algorithmic flops and bytes, the sum over launches of
    // max(flops / FP64 peak, bytes / HBM peak), and the launch count
    double fac_time[4] = {0, 0, 0, 0}, fac_flops[4] = {0, 0, 0, 0}, fac_bytes[4] = {0, 0, 0, 0},
           fac_bound[4] = {0, 0, 0, 0};
    int fac_launches[4] = {0, 0, 0, 0};
  };
  enum { FAC_GEMM = 0, FAC_LU = 1, FAC_LEAF_SOLVE = 2, FAC_NODE_SOLVE = 3 };
  void set_profiling(bool on) { profiling_ = on; }
  /* Times the candidate launch shapes on this problem (compression for the
   * ACA rows per thread; FP64 factorization for the GEMM tiles, split-K
   * target and LU panel width), keeps the fastest and records it for every
   * later handle of the same shape in this process. */
  void autotune();
  const Tuning& tuning() const { return tune_; }
  void set_tuning(const Tuning& t);
  // replay FP64 solves (and FP64-preconditioned refinement steps) from
  // captured hipGraphs of the level sequence
  void set_graphs(bool on);
  // rank predictor (MLP parameters, host copy); null detaches it
  void set_rank_predictor(const float* params, double margin);
  void block_features(double tol, float* host_out);
  const Stats& stats() const { return stats_; }
  /* Completes figures that wait for work still queued on the side stream
   * (the factorization time including the roots' LU). */
  void sync_stats();
  long long device_bytes() const;
  hipStream_t stream() const { return stream_; }

 private:
  void build_trees();
  void upload_sources();
  /* Profiling of the FP64 factorization's kernel groups: a span is a pair of
   * events around one group launch (outermost only) with its work. */
  struct ProfSpan {
    int cat;
    double flops, bytes;
    hipEvent_t a, b;
  };
  bool prof_fac_ = false;
  int prof_depth_ = 0;
  std::vector<ProfSpan> spans_;
  std::vector<hipEvent_t> span_events_;
  size_t span_next_ = 0;
  int prof_begin(hipStream_t s);
  void prof_end(int token, hipStream_t s, int cat, double flops, double bytes);
  void prof_collect();
  void reset_attributes();
  // each compresses the blocks whose node uses its method; ranks per block
  // (2 per node: top, bottom) into blk_rank
  void aca_compress(std::vector<int>& blk_rank);
  void svd_compress(std::vector<int>& blk_rank);
  int node_method(const TreeNodeH& nd) const { return nd.lr >= 0 ? nd.lr : lr_method_; }
  void block_features_device(double tol, DBuf<float>& out);
  std::vector<int> predicted_capacities();
  void assemble_leaves();
  void factor_leaves();
  void leaf_solves(const std::vector<int>& leaf_ids, double* B, int ldb, int ncols, bool compact);
  void run_gemm(std::vector<GemmDesc>& descs);
  template <class T>
  void run_gemm_t(std::vector<GemmDescT<T>>& descs);
  template <class T>
  void run_gemm_tiled(std::vector<GemmDescT<T>>& descs);
  void ensure_ws(size_t doubles);
  void grow(DBuf<double>& b, size_t doubles);
  void lu_batch(std::vector<LuDesc>& descs);
  void lu_batch(std::vector<LuDesc>& descs, hipStream_t s, DescArena& arena);
  // the roots' LU, queued on aux_ by the factorization: stream_ waits for it
  // (wait_root), or the host does and the aux arena is released (finish_root)
  void wait_root(Factorization& F);
  void finish_root();
  void factorize_fp64(int which);
  void logdet_terms(std::vector<double>& hl, std::vector<double>& hp, std::vector<int>& inst);
  void factorize_mixed(int which);
  void solve_fp64_inplace(int which, double* X, int nrhs);
  /* X (n x nrhs, ld n) <- A^-1 X through the captured graph */
  void solve_fp64_graph(int which, double* X, int nrhs);
  void solve_mixed_once(int which, float* X, int nrhs);
  void solve_gmres(int which, bool mixed, const double* dB, int ldb, int nrhs, double* dX, int ldx);
  template <class T>
  void grow_t(DBuf<T>& b, size_t n);
  template <class T>
  long long count_nonfinite(const T* x, size_t n);
  // mode 0: by size (one launch up to 2048 rows, else blocked); 1: the
  // permutation / block substitution / GEMM path whatever the size; 2: the
  // factor shared by a workgroup's waves (systems of at most 512 rows)
  void lu_solve_batch(std::vector<SolveDesc>& descs, int mode = 0);

  std::vector<InstanceSpec> specs_;
  int ntot_ = 0, nlev_ = 0;
  std::vector<int> inst_row0_;
  std::vector<TreeNodeH> nodes_;               // level order, left to right
  std::vector<std::vector<int>> level_nodes_;  // node ids per level
  std::vector<LeafH> leaves_;
  std::vector<std::vector<int>> level_leaves_;

  int lr_method_ = 0;
  double tol_ = 1e-6;   // HODLR_Matrix.cpp:7
  int min_pivot_ = 0;
  int precision_ = 0;
  double refine_target_ = 1e-10;
  int refine_max_ = 30;
  int init_cap_ = 0;
  bool profiling_ = false;
  bool use_graphs_ = false;
  Tuning tune_;
  std::string shape_key_;
  unsigned long long buf_gen_ = 0;  // bumped whenever a work buffer is reallocated
  SolveGraph graph_;
  DBuf<double> gx_;
  std::vector<float> pred_params_;
  double pred_margin_ = 0.25;
  std::vector<double> pred_rank_;
  bool trace_ = false;
  int solve_small_max_ = LU_SOLVE_SMALL_MAX;  // one-launch LU solves up to this size
  int solve_wave_max_ = LU_SOLVE_WAVE_MAX;    // of which one wave per system up to this size
  bool solve_blk_ = true;                     // and in blocks of 64 rows above 512
  std::vector<hipEvent_t> prof_ev_;

  hipStream_t stream_ = nullptr;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  DescArena arena_;
  // The roots' capacitance LU (one system per instance, the largest of the
  // tree, a long chain of small launches) runs on aux_ while the solve that
  // follows the factorization works through the lower levels on stream_.
  hipStream_t aux_ = nullptr;
  hipEvent_t ev_cap_ = nullptr, ev_root_ = nullptr;
  hipEvent_t ev_aca_ = nullptr;  // the ACA state snapshot's copy
  // the factorization's copy of the panels into W on aux_, beside the
  // leaves' LU on stream_: its start and its end
  hipEvent_t ev_w0_ = nullptr, ev_w1_ = nullptr;
  // timing of the factorization: its start on stream_, its end on stream_
  // and the end of the roots' LU on aux_ (read by sync_stats)
  hipEvent_t ev_f0_ = nullptr, ev_f1_ = nullptr, ev_froot_ = nullptr;
  bool factor_time_pending_ = false;
  DescArena aux_arena_;
  bool overlap_root_ = true;
  DBuf<EntrySource> d_srcs_;
  std::vector<DBuf<double>> d_coords_;
  std::vector<DBuf<double>> d_dense_;
  std::vector<EntrySource> h_srcs_;

  bool compressed_ = false;
  bool leaves_factored_ = false;
  std::vector<DBuf<double>> U_, V_;
  std::vector<int> K_, off_;
  int ktot_ = 0;
  DBuf<double> Lorig_, Lfac_;
  DBuf<int> Lpiv_;
  DBuf<int> Lperm_;  // the leaves' interchanges as permutations (leaves up to LU_PERM_MAX rows)
  bool adaptive_aca_pt_ = true;  // fewer ACA rows per thread when few workgroups remain
  bool stream_update_ = true;  // small-rank updates by the streaming kernel
  static constexpr int aca_min_wg_ = 256;  // ... fewer row-pass workgroups than this
  DBuf<int> lu_swp_;  // the batched LU's recorded panel interchanges (LU_SWP_INTS per matrix)
  size_t leaf_elems_ = 0, leaf_piv_elems_ = 0;
  Factorization fac_[2];
  MixedFactorization mfac_[2];
  DBuf<float> t32a_, t32b_, x32_;
  DBuf<double> rbuf_, ybuf_, zbuf_, nbuf_, basis_, zbasis_;
  int gmres_restart_ = 20;
  DBuf<double> ws_, tmp_, tmp2_;
  DBuf<double> xw_;
  DBuf<double> xbest_;  // refinement: best iterate per instance and column so far
  DBuf<int> bestmask_;  // which instance columns to copy into (or back from) xbest_
  DBuf<double> ddy_, ddpart_, ddt_;  // double-double residual: Y, partials, projections
  DBuf<int> seg_rows_;               // instance row offsets (+ total) on the device
  DBuf<double> segbuf_, bscaled_;    // per-instance norms / scales, scaled batch RHS
  Stats stats_;
};

}  // namespace hodlr

#endif
