/*
 * C ABI of the HODLR engine (declarations and reference citations in
 * include/dense_hodlr.h).  Exceptions never cross the boundary: every entry
 * point maps them to a status code and keeps the message.
 */
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <string>

#include "../../include/dense_hodlr.h"
#include "engine.h"
#include "kd.h"
#include "matrix_io.h"
#include "peaks.h"
#include "model.h"

struct hodlr_handle_s {
  std::unique_ptr<hodlr::Engine> eng;
  std::string err;
};

struct hodlr_rank_model_s {
  std::unique_ptr<hodlr::RankModel> m;
};

namespace {

thread_local std::string g_create_err;

int fail(hodlr_handle h, int code, const std::string& msg) {
  if (h)
    h->err = msg;
  else
    g_create_err = msg;
  return code;
}

template <class F>
int guard(hodlr_handle h, F&& f) {
  if (!h) return fail(nullptr, HODLR_ERR_ARG, "null handle");
  try {
    f();
    h->err.clear();
    return HODLR_OK;
  } catch (const std::invalid_argument& e) {
    return fail(h, HODLR_ERR_ARG, e.what());
  } catch (const hodlr::NotConverged& e) {
    return fail(h, HODLR_ERR_NOT_CONVERGED, e.what());
  } catch (const hodlr::HipError& e) {
    const std::string m = e.what();
    if (m.find("hipMalloc") != std::string::npos || m.find("out of memory") != std::string::npos)
      return fail(h, HODLR_ERR_NOMEM, m);
    return fail(h, HODLR_ERR_DEVICE, m);
  } catch (const std::bad_alloc&) {
    return fail(h, HODLR_ERR_NOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(h, HODLR_ERR_STATE, e.what());
  }
}

template <class F>
int guard_model(hodlr_rank_model m, F&& f) {
  if (!m) return fail(nullptr, HODLR_ERR_ARG, "null model");
  try {
    f();
    return HODLR_OK;
  } catch (const std::invalid_argument& e) {
    return fail(nullptr, HODLR_ERR_ARG, e.what());
  } catch (const hodlr::HipError& e) {
    return fail(nullptr, HODLR_ERR_DEVICE, e.what());
  } catch (const std::bad_alloc&) {
    return fail(nullptr, HODLR_ERR_NOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(nullptr, HODLR_ERR_STATE, e.what());
  }
}

/* Host-only calls (matrix files): errors go to hodlr_last_error(NULL). */
template <class F>
int guard_io(F&& f) {
  try {
    f();
    return HODLR_OK;
  } catch (const std::invalid_argument& e) {
    return fail(nullptr, HODLR_ERR_ARG, e.what());
  } catch (const std::bad_alloc&) {
    return fail(nullptr, HODLR_ERR_NOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(nullptr, HODLR_ERR_STATE, e.what());
  }
}

int check_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return -1;
  return 0;
}

int make_handle(hodlr_handle* h, std::vector<hodlr::InstanceSpec>&& specs) {
  if (check_device() != 0) return fail(nullptr, HODLR_ERR_DEVICE, "no HIP device available");
  auto* hh = new (std::nothrow) hodlr_handle_s();
  if (!hh) return fail(nullptr, HODLR_ERR_NOMEM, "host allocation failed");
  try {
    hh->eng.reset(new hodlr::Engine(std::move(specs)));
  } catch (const std::invalid_argument& e) {  // (a user partition that does not fit)
    delete hh;
    return fail(nullptr, HODLR_ERR_ARG, e.what());
  } catch (const std::exception& e) {
    delete hh;
    return fail(nullptr, HODLR_ERR_DEVICE, e.what());
  }
  *h = hh;
  return HODLR_OK;
}

/* Validates a typed kernel descriptor and fills one instance; returns the
 * error message, empty on success. */
std::string kernel_spec(int n, const hodlr_kernel_desc* k, int leaf, hodlr::InstanceSpec& s) {
  if (n <= 0 || leaf <= 0 || !k || !k->points) return "n, leaf must be positive and points non-null";
  if (k->dim < 1 || k->dim > 3) return "dim must be 1, 2 or 3";
  if (k->family < 0 || k->family > HODLR_KERNEL_EXPONENTIAL) return "unknown kernel family";
  if (!(k->length_scale > 0.0)) return "length_scale must be > 0";
  if (k->diag_mode != HODLR_DIAG_ADD_SHIFT && k->diag_mode != HODLR_DIAG_REPLACE_VALUE)
    return "unknown diag_mode";
  s.n = n;
  s.leaf = leaf;
  s.kind = 0;
  s.xs.resize(n);
  if (k->dim > 1) s.ys.resize(n);
  if (k->dim > 2) s.zs.resize(n);
  for (int i = 0; i < n; i++) {
    s.xs[i] = k->points[(size_t)i * k->dim];
    if (k->dim > 1) s.ys[i] = k->points[(size_t)i * k->dim + 1];
    if (k->dim > 2) s.zs[i] = k->points[(size_t)i * k->dim + 2];
  }
  hodlr_kernel_params_init(&s.kp, k->family, k->dim, k->length_scale, k->diag, k->diag_mode,
                           nullptr, nullptr, nullptr);
  return std::string();
}

/* A user partition from the C node array. */
std::vector<hodlr::UserNode> user_tree(const hodlr_tree_node* nodes, int count) {
  std::vector<hodlr::UserNode> t;
  if (count < 0 || (count > 0 && !nodes)) throw std::invalid_argument("bad tree arguments");
  for (int i = 0; i < count; i++) {
    hodlr::UserNode u;
    u.split = nodes[i].split;
    u.min_rank[0] = nodes[i].top_min_rank;
    u.min_rank[1] = nodes[i].bottom_min_rank;
    u.max_rank[0] = nodes[i].top_max_rank;
    u.max_rank[1] = nodes[i].bottom_max_rank;
    u.lr = nodes[i].lr_method;
    u.left = nodes[i].left;
    u.right = nodes[i].right;
    t.push_back(u);
  }
  return t;
}

int wrap_engine(hodlr_handle* out, std::unique_ptr<hodlr::Engine> e) {
  auto* hh = new (std::nothrow) hodlr_handle_s();
  if (!hh) return fail(nullptr, HODLR_ERR_NOMEM, "host allocation failed");
  hh->eng = std::move(e);
  *out = hh;
  return HODLR_OK;
}

}  // namespace

extern "C" {

int hodlr_create_kernel(hodlr_handle* h, int n, const hodlr_kernel_desc* k, int leaf) {
  if (!h) return fail(nullptr, HODLR_ERR_ARG, "null handle pointer");
  *h = nullptr;
  std::vector<hodlr::InstanceSpec> specs(1);
  const std::string err = kernel_spec(n, k, leaf, specs[0]);
  if (!err.empty()) return fail(nullptr, HODLR_ERR_ARG, err);
  return make_handle(h, std::move(specs));
}

int hodlr_create_kernel_tree(hodlr_handle* h, int n, const hodlr_kernel_desc* k, int leaf,
                             const hodlr_tree_node* nodes, int count) {
  if (!h) return fail(nullptr, HODLR_ERR_ARG, "null handle pointer");
  *h = nullptr;
  std::vector<hodlr::InstanceSpec> specs(1);
  const std::string err = kernel_spec(n, k, leaf, specs[0]);
  if (!err.empty()) return fail(nullptr, HODLR_ERR_ARG, err);
  int rc = guard_io([&] { specs[0].tree = user_tree(nodes, count); });
  if (rc != HODLR_OK) return rc;
  return make_handle(h, std::move(specs));
}

int hodlr_create_dense_tree(hodlr_handle* h, int n, const double* A, int lda, int leaf,
                            const hodlr_tree_node* nodes, int count) {
  if (!h) return fail(nullptr, HODLR_ERR_ARG, "null handle pointer");
  *h = nullptr;
  if (n <= 0 || leaf <= 0 || !A || lda < n) return fail(nullptr, HODLR_ERR_ARG, "bad arguments");
  std::vector<hodlr::InstanceSpec> specs(1);
  hodlr::InstanceSpec& s = specs[0];
  s.n = n;
  s.leaf = leaf;
  s.kind = 1;
  int rc = guard_io([&] {
    s.tree = user_tree(nodes, count);
    s.dense.resize((size_t)n * n);
    for (int j = 0; j < n; j++)
      std::memcpy(s.dense.data() + (size_t)j * n, A + (size_t)j * lda, (size_t)n * sizeof(double));
  });
  if (rc != HODLR_OK) return rc;
  return make_handle(h, std::move(specs));
}

int hodlr_kd_sort(const double* points, int npoints, int dim, int* perm) {
  return guard_io([&] {
    if (!perm) throw std::invalid_argument("null output");
    hodlr::kd_sort(points, npoints, dim, perm);
  });
}

int hodlr_kd_tree(const double* points, int npoints, int dim, int unknowns_per_point, int max_rank,
                  int lr_method, int* perm, hodlr_tree_node* nodes, int* count) {
  return guard_io([&] {
    if (!perm || !nodes || !count) throw std::invalid_argument("null output");
    const std::vector<hodlr::UserNode> t =
        hodlr::kd_tree(points, npoints, dim, unknowns_per_point, max_rank, lr_method, perm);
    for (size_t i = 0; i < t.size(); i++) {
      nodes[i].split = t[i].split;
      nodes[i].top_min_rank = t[i].min_rank[0];
      nodes[i].bottom_min_rank = t[i].min_rank[1];
      nodes[i].top_max_rank = t[i].max_rank[0];
      nodes[i].bottom_max_rank = t[i].max_rank[1];
      nodes[i].lr_method = t[i].lr;
      nodes[i].left = t[i].left;
      nodes[i].right = t[i].right;
    }
    *count = (int)t.size();
  });
}

int hodlr_diag_block(hodlr_handle h, int side, hodlr_handle* out) {
  if (!out) return fail(h, HODLR_ERR_ARG, "null output handle");
  *out = nullptr;
  std::unique_ptr<hodlr::Engine> e;
  const int rc = guard(h, [&] { e = h->eng->diagonal_block(side); });
  if (rc != HODLR_OK) return rc;
  return wrap_engine(out, std::move(e));
}

int hodlr_split_at_top(hodlr_handle h, hodlr_handle* top, hodlr_handle* bottom) {
  if (!top || !bottom) return fail(h, HODLR_ERR_ARG, "null output handle");
  *top = *bottom = nullptr;
  std::unique_ptr<hodlr::Engine> t, b;
  const int rc = guard(h, [&] {
    t = h->eng->diagonal_block(0);
    b = h->eng->diagonal_block(1);
  });
  if (rc != HODLR_OK) return rc;
  int r = wrap_engine(top, std::move(t));
  if (r == HODLR_OK) r = wrap_engine(bottom, std::move(b));
  if (r != HODLR_OK) {
    hodlr_destroy(*top);
    *top = nullptr;
  }
  return r;
}

int hodlr_create_batch(hodlr_handle* h, int count, const int* n, const hodlr_kernel_desc* kernels,
                       int leaf) {
  if (!h) return fail(nullptr, HODLR_ERR_ARG, "null handle pointer");
  *h = nullptr;
  if (count <= 0 || !n || !kernels) return fail(nullptr, HODLR_ERR_ARG, "bad arguments");
  std::vector<hodlr::InstanceSpec> specs(count);
  long long tot = 0;
  for (int q = 0; q < count; q++) {
    const std::string err = kernel_spec(n[q], &kernels[q], leaf, specs[q]);
    if (!err.empty()) return fail(nullptr, HODLR_ERR_ARG, "instance " + std::to_string(q) + ": " + err);
    tot += n[q];
  }
  if (tot > 0x7fffffffLL) return fail(nullptr, HODLR_ERR_ARG, "total size exceeds 2^31 - 1 rows");
  return make_handle(h, std::move(specs));
}

int hodlr_batch_count(hodlr_handle h) {
  if (!h) return HODLR_ERR_ARG;
  return h->eng->num_instances();
}

int hodlr_batch_offsets(hodlr_handle h, int* row0) {
  return guard(h, [&] {
    if (!row0) throw std::invalid_argument("null output");
    const auto& r = h->eng->instance_rows();
    for (size_t q = 0; q < r.size(); q++) row0[q] = r[q];
    row0[r.size()] = h->eng->rows();
  });
}

int hodlr_logabsdet_batch(hodlr_handle h, double* out) {
  return guard(h, [&] {
    if (!out) throw std::invalid_argument("null output");
    std::vector<double> v = h->eng->logabsdet_instances();
    for (size_t q = 0; q < v.size(); q++) out[q] = v[q];
  });
}

int hodlr_create_kernel_callback(hodlr_handle* h, int n, double (*fn)(int, int, void*), void* data,
                                 int leaf) {
  if (!h) return fail(nullptr, HODLR_ERR_ARG, "null handle pointer");
  *h = nullptr;
  if (n <= 0 || leaf <= 0 || !fn) return fail(nullptr, HODLR_ERR_ARG, "bad arguments");
  std::vector<hodlr::InstanceSpec> specs(1);
  hodlr::InstanceSpec& s = specs[0];
  s.n = n;
  s.leaf = leaf;
  s.kind = 1;
  try {
    s.dense.resize((size_t)n * n);
  } catch (const std::bad_alloc&) {
    return fail(nullptr, HODLR_ERR_NOMEM, "host allocation of the dense copy failed");
  }
  for (int j = 0; j < n; j++)  // column-major, as kernelMatrix::block (kernel.cpp:19-25)
    for (int i = 0; i < n; i++) s.dense[(size_t)i + (size_t)j * n] = fn(i, j, data);
  return make_handle(h, std::move(specs));
}

int hodlr_create_dense(hodlr_handle* h, int n, const double* A, int lda, int leaf, int lr_method) {
  if (!h) return fail(nullptr, HODLR_ERR_ARG, "null handle pointer");
  *h = nullptr;
  if (n <= 0 || leaf <= 0 || !A || lda < n) return fail(nullptr, HODLR_ERR_ARG, "bad arguments");
  if (lr_method != HODLR_LR_ACA && lr_method != HODLR_LR_SVD)
    return fail(nullptr, HODLR_ERR_ARG, "unknown low-rank method");
  std::vector<hodlr::InstanceSpec> specs(1);
  hodlr::InstanceSpec& s = specs[0];
  s.n = n;
  s.leaf = leaf;
  s.kind = 1;
  try {
    s.dense.resize((size_t)n * n);
  } catch (const std::bad_alloc&) {
    return fail(nullptr, HODLR_ERR_NOMEM, "host allocation of the dense copy failed");
  }
  for (int j = 0; j < n; j++)
    std::memcpy(s.dense.data() + (size_t)j * n, A + (size_t)j * lda, (size_t)n * sizeof(double));
  int rc = make_handle(h, std::move(specs));
  if (rc != HODLR_OK) return rc;
  rc = hodlr_set_lr_method(*h, lr_method);
  if (rc != HODLR_OK) {
    g_create_err = (*h)->err;
    hodlr_destroy(*h);
    *h = nullptr;
  }
  return rc;
}

/* ---- matrix files ---- */
int hodlr_read_binary_matrix(const char* path, int* rows, int* cols, double* out, int ld) {
  return guard_io([&] {
    if (!path || !rows || !cols) throw std::invalid_argument("null argument");
    hodlr::binary_matrix_dims(path, rows, cols);
    if (out) hodlr::read_binary_matrix(path, *rows, *cols, out, ld);
  });
}

int hodlr_write_binary_matrix(const char* path, int rows, int cols, const double* A, int lda) {
  return guard_io([&] {
    if (!path || (!A && rows > 0 && cols > 0)) throw std::invalid_argument("null argument");
    hodlr::write_binary_matrix(path, rows, cols, A, lda);
  });
}

int hodlr_read_text_matrix(const char* path, int* rows, int* cols, double* out, int ld) {
  return guard_io([&] {
    if (!path || !rows || !cols) throw std::invalid_argument("null argument");
    hodlr::text_matrix_dims(path, rows, cols);
    if (out) hodlr::read_text_matrix(path, *rows, *cols, out, ld);
  });
}

int hodlr_create_dense_from_file(hodlr_handle* h, const char* path, int leaf, int lr_method) {
  if (!h) return fail(nullptr, HODLR_ERR_ARG, "null handle pointer");
  *h = nullptr;
  if (!path || leaf <= 0) return fail(nullptr, HODLR_ERR_ARG, "bad arguments");
  if (lr_method != HODLR_LR_ACA && lr_method != HODLR_LR_SVD)
    return fail(nullptr, HODLR_ERR_ARG, "unknown low-rank method");
  std::vector<hodlr::InstanceSpec> specs(1);
  hodlr::InstanceSpec& s = specs[0];
  int rows = 0, cols = 0;
  int rc = guard_io([&] {
    hodlr::binary_matrix_dims(path, &rows, &cols);
    // HODLR_Matrix.cpp:92: assert(isSquareMatrix == true)
    if (rows != cols || rows <= 0) throw std::invalid_argument("the HODLR matrix must be square");
    s.dense.resize((size_t)rows * rows);
    hodlr::read_binary_matrix(path, rows, cols, s.dense.data(), rows);
  });
  if (rc != HODLR_OK) return rc;
  s.n = rows;
  s.leaf = leaf;
  s.kind = 1;
  rc = make_handle(h, std::move(specs));
  if (rc != HODLR_OK) return rc;
  rc = hodlr_set_lr_method(*h, lr_method);
  if (rc != HODLR_OK) {
    g_create_err = (*h)->err;
    hodlr_destroy(*h);
    *h = nullptr;
  }
  return rc;
}

int hodlr_destroy(hodlr_handle h) {
  if (!h) return HODLR_ERR_ARG;
  delete h;
  return HODLR_OK;
}

int hodlr_set_lr_method(hodlr_handle h, int m) {
  return guard(h, [&] { h->eng->set_lr_method(m); });
}
int hodlr_set_tolerance(hodlr_handle h, double t) {
  return guard(h, [&] { h->eng->set_tolerance(t); });
}
int hodlr_set_min_pivot(hodlr_handle h, double p) {
  return guard(h, [&] { h->eng->set_min_pivot(p); });
}
int hodlr_set_precision(hodlr_handle h, int p) {
  return guard(h, [&] { h->eng->set_precision(p); });
}
int hodlr_set_refinement(hodlr_handle h, double target, int max_iters) {
  return guard(h, [&] { h->eng->set_refinement(target, max_iters); });
}

int hodlr_set_launch_shapes(hodlr_handle h, int aca_rows_per_thread, int gemm_tile_m,
                            int gemm_split_target, int lu_panel, int gemm_mfma,
                            int leaf_solve) {
  return guard(h, [&] {
    hodlr::Tuning t;
    t.aca_pt = aca_rows_per_thread;
    t.gemm_bm = gemm_tile_m;
    t.gemm_target = gemm_split_target;
    t.lu_nb = lu_panel;
    t.gemm_mfma = gemm_mfma;
    t.leaf_solve = leaf_solve;
    h->eng->set_tuning(t);
  });
}

int hodlr_autotune(hodlr_handle h) {
  return guard(h, [&] { h->eng->autotune(); });
}

int hodlr_set_graphs(hodlr_handle h, int on) {
  return guard(h, [&] { h->eng->set_graphs(on != 0); });
}

int hodlr_set_profiling(hodlr_handle h, int on) {
  return guard(h, [&] { h->eng->set_profiling(on != 0); });
}

int hodlr_compress(hodlr_handle h) {
  return guard(h, [&] { h->eng->compress(); });
}
int hodlr_factorize(hodlr_handle h, int which) {
  return guard(h, [&] { h->eng->factorize(which); });
}
int hodlr_solve_sm(hodlr_handle h, const double* B, int ldb, int nrhs, double* X, int ldx) {
  return guard(h, [&] {
    if (!B || !X || nrhs < 0) throw std::invalid_argument("bad arguments");
    h->eng->solve_host(HODLR_FACTOR_SM, B, ldb, nrhs, X, ldx);
  });
}
int hodlr_solve_lu(hodlr_handle h, const double* B, int ldb, int nrhs, double* X, int ldx) {
  return guard(h, [&] {
    if (!B || !X || nrhs < 0) throw std::invalid_argument("bad arguments");
    h->eng->solve_host(HODLR_FACTOR_LU, B, ldb, nrhs, X, ldx);
  });
}
int hodlr_solve_device(hodlr_handle h, int which, const double* dB, int ldb, int nrhs, double* dX,
                       int ldx) {
  return guard(h, [&] {
    if (!dB || !dX || nrhs < 0) throw std::invalid_argument("bad arguments");
    h->eng->solve_device(which, dB, ldb, nrhs, dX, ldx);
  });
}
namespace {
void iterative_prepare(hodlr_handle h, int max_iterations, double stop_tolerance,
                       double init_lr_tolerance, int lr_method, int direct_method) {
  if (max_iterations < 0 || !(stop_tolerance >= 0.0))
    throw std::invalid_argument("bad iteration settings");
  if (direct_method != HODLR_FACTOR_SM && direct_method != HODLR_FACTOR_LU)
    throw std::invalid_argument("unknown direct solver");
  // HODLR_Matrix.cpp:1141-1145: the handle keeps these settings afterwards
  h->eng->set_tolerance(init_lr_tolerance);
  h->eng->set_lr_method(lr_method);
}
}  // namespace

int hodlr_iterative_solve(hodlr_handle h, const double* B, int ldb, int nrhs, double* X, int ldx,
                          int max_iterations, double stop_tolerance, double init_lr_tolerance,
                          int lr_method, int direct_method) {
  return guard(h, [&] {
    const int n = h->eng->rows();
    if (!B || !X || nrhs < 0 || ldb < n || ldx < n) throw std::invalid_argument("bad arguments");
    iterative_prepare(h, max_iterations, stop_tolerance, init_lr_tolerance, lr_method,
                      direct_method);
    if (nrhs == 0) return;
    hodlr::DBuf<double> db, dx;
    db.alloc((size_t)n * nrhs);
    dx.alloc((size_t)n * nrhs);
    HODLR_HIP_CHECK(hipMemcpy2D(db.p, (size_t)n * 8, B, (size_t)ldb * 8, (size_t)n * 8, nrhs,
                                hipMemcpyHostToDevice));
    auto back = [&] {
      HODLR_HIP_CHECK(hipMemcpy2D(X, (size_t)ldx * 8, dx.p, (size_t)n * 8, (size_t)n * 8, nrhs,
                                  hipMemcpyDeviceToHost));
    };
    try {
      h->eng->iterative_solve(direct_method, db.p, n, nrhs, dx.p, n, max_iterations,
                              stop_tolerance);
    } catch (const hodlr::NotConverged&) {
      back();
      throw;
    }
    back();
  });
}

int hodlr_iterative_solve_device(hodlr_handle h, const double* dB, int ldb, int nrhs, double* dX,
                                 int ldx, int max_iterations, double stop_tolerance,
                                 double init_lr_tolerance, int lr_method, int direct_method) {
  return guard(h, [&] {
    if (!dB || !dX || nrhs < 0) throw std::invalid_argument("bad arguments");
    iterative_prepare(h, max_iterations, stop_tolerance, init_lr_tolerance, lr_method,
                      direct_method);
    h->eng->iterative_solve(direct_method, dB, ldb, nrhs, dX, ldx, max_iterations, stop_tolerance);
  });
}

int hodlr_exact_apply_device(hodlr_handle h, const double* dX, int ldx, int nrhs, double* dY,
                             int ldy) {
  return guard(h, [&] {
    const int n = h->eng->rows();
    if (!dX || !dY || nrhs < 0 || ldx < n || ldy < n) throw std::invalid_argument("bad arguments");
    h->eng->exact_apply(dX, ldx, nrhs, dY, ldy);
    HODLR_HIP_CHECK(hipStreamSynchronize(h->eng->stream()));
  });
}

int hodlr_logabsdet(hodlr_handle h, double* out) {
  return guard(h, [&] {
    if (!out) throw std::invalid_argument("null output");
    *out = h->eng->logabsdet(nullptr);
  });
}
int hodlr_determinant(hodlr_handle h, double* out) {
  return guard(h, [&] {
    if (!out) throw std::invalid_argument("null output");
    double p = 0;
    h->eng->logabsdet(&p);
    *out = p;
  });
}
int hodlr_matvec(hodlr_handle h, const double* X, int ldx, int nrhs, double* Y, int ldy) {
  return guard(h, [&] {
    const int n = h->eng->rows();
    if (!X || !Y || nrhs < 0 || ldx < n || ldy < n) throw std::invalid_argument("bad arguments");
    if (nrhs == 0) return;
    hodlr::DBuf<double> dx, dy;
    dx.alloc((size_t)n * nrhs);
    dy.alloc((size_t)n * nrhs);
    HODLR_HIP_CHECK(hipMemcpy2D(dx.p, (size_t)n * 8, X, (size_t)ldx * 8, (size_t)n * 8, nrhs,
                                hipMemcpyHostToDevice));
    h->eng->matvec_device(dx.p, n, nrhs, dy.p, n);
    HODLR_HIP_CHECK(hipMemcpy2D(Y, (size_t)ldy * 8, dy.p, (size_t)n * 8, (size_t)n * 8, nrhs,
                                hipMemcpyDeviceToHost));
  });
}

int hodlr_matvec_device_noalloc(hodlr_handle h, const double* dX, int ldx, int nrhs, double* dY,
                                int ldy) {
  return guard(h, [&] {
    const int n = h->eng->rows();
    if (!dX || !dY || nrhs < 0 || ldx < n || ldy < n) throw std::invalid_argument("bad arguments");
    h->eng->matvec_noalloc(dX, ldx, nrhs, dY, ldy);
  });
}

int hodlr_residual(hodlr_handle h, const double* B, int ldb, int nrhs, const double* X, int ldx,
                   double* R, int ldr) {
  return guard(h, [&] {
    const int n = h->eng->rows();
    if (!B || !X || !R || nrhs < 0 || ldb < n || ldx < n || ldr < n)
      throw std::invalid_argument("bad arguments");
    if (nrhs == 0) return;
    hodlr::DBuf<double> db, dx, dr;
    db.alloc((size_t)n * nrhs);
    dx.alloc((size_t)n * nrhs);
    dr.alloc((size_t)n * nrhs);
    HODLR_HIP_CHECK(hipMemcpy2D(db.p, (size_t)n * 8, B, (size_t)ldb * 8, (size_t)n * 8, nrhs,
                                hipMemcpyHostToDevice));
    HODLR_HIP_CHECK(hipMemcpy2D(dx.p, (size_t)n * 8, X, (size_t)ldx * 8, (size_t)n * 8, nrhs,
                                hipMemcpyHostToDevice));
    h->eng->residual_dd(db.p, n, dx.p, n, nrhs, dr.p, n);
    HODLR_HIP_CHECK(hipMemcpy2D(R, (size_t)ldr * 8, dr.p, (size_t)n * 8, (size_t)n * 8, nrhs,
                                hipMemcpyDeviceToHost));
  });
}

int hodlr_rank_model_create(hodlr_rank_model* m, unsigned long long seed) {
  if (!m) return fail(nullptr, HODLR_ERR_ARG, "null model pointer");
  *m = nullptr;
  if (check_device() != 0) return fail(nullptr, HODLR_ERR_DEVICE, "no HIP device available");
  auto* mm = new (std::nothrow) hodlr_rank_model_s();
  if (!mm) return fail(nullptr, HODLR_ERR_NOMEM, "host allocation failed");
  try {
    mm->m.reset(new hodlr::RankModel(seed));
  } catch (const std::exception& e) {
    delete mm;
    return fail(nullptr, HODLR_ERR_DEVICE, e.what());
  }
  *m = mm;
  return HODLR_OK;
}

int hodlr_rank_model_destroy(hodlr_rank_model m) {
  if (!m) return fail(nullptr, HODLR_ERR_ARG, "null model");
  delete m;
  return HODLR_OK;
}

int hodlr_rank_model_num_params(void) { return hodlr::MLP_NPARAM; }

int hodlr_rank_model_get_params(hodlr_rank_model m, float* out) {
  return guard_model(m, [&] {
    if (!out) throw std::invalid_argument("null output");
    m->m->get_params(out);
  });
}

int hodlr_rank_model_set_params(hodlr_rank_model m, const float* in) {
  return guard_model(m, [&] {
    if (!in) throw std::invalid_argument("null input");
    m->m->set_params(in);
  });
}

int hodlr_rank_model_forward(hodlr_rank_model m, const float* X, int n, float* y) {
  return guard_model(m, [&] {
    if (n < 0 || (n > 0 && (!X || !y))) throw std::invalid_argument("bad arguments");
    m->m->forward(X, n, y);
  });
}

int hodlr_rank_model_predict_bf16(hodlr_rank_model m, const float* X, int n, float* y) {
  return guard_model(m, [&] {
    if (n < 0 || (n > 0 && (!X || !y))) throw std::invalid_argument("bad arguments");
    m->m->predict_bf16(X, n, y);
  });
}

int hodlr_rank_model_gradient(hodlr_rank_model m, const float* X, const float* t, int n,
                              float* grad, float* loss) {
  return guard_model(m, [&] {
    if (n <= 0 || !X || !t || !grad) throw std::invalid_argument("bad arguments");
    const float l = m->m->gradient(X, t, n, grad);
    if (loss) *loss = l;
  });
}

int hodlr_rank_model_train(hodlr_rank_model m, const float* X, const float* t, int n, int epochs,
                           int batch, float lr, unsigned long long seed, float* epoch_loss) {
  return guard_model(m, [&] {
    if (!X || !t) throw std::invalid_argument("null data");
    m->m->train(X, t, n, epochs, batch, lr, seed, epoch_loss);
  });
}

int hodlr_block_features(hodlr_handle h, double tol, float* out) {
  return guard(h, [&] {
    if (!out || !(tol > 0.0)) throw std::invalid_argument("bad arguments");
    h->eng->block_features(tol, out);
  });
}

int hodlr_set_rank_predictor(hodlr_handle h, hodlr_rank_model m, double margin) {
  return guard(h, [&] {
    if (m)
      h->eng->set_rank_predictor(m->m->host_params().data(), margin);
    else
      h->eng->set_rank_predictor(nullptr, margin);
  });
}

int hodlr_get_stats(hodlr_handle h, hodlr_stats* out) {
  return guard(h, [&] {
    if (!out) throw std::invalid_argument("null output");
    std::memset(out, 0, sizeof(*out));
    h->eng->sync_stats();
    const hodlr::Engine& e = *h->eng;
    out->num_levels = e.num_levels();
    out->num_nodes = (int)e.nodes().size();
    out->num_leaves = (int)e.leaves().size();
    const auto& s = e.stats();
    out->aca_iterations = s.aca_iterations;
    const hodlr::Tuning& tu = e.tuning();
    out->tune_aca_rows_per_thread = tu.aca_pt;
    out->tune_gemm_tile_m = tu.gemm_bm;
    out->tune_gemm_split_target = tu.gemm_target;
    out->tune_lu_panel = tu.lu_nb;
    out->tune_gemm_mfma = tu.gemm_mfma;
    out->tune_leaf_solve = tu.leaf_solve;
    out->tuned = tu.tuned ? 1 : 0;
    out->graph_captures = s.graph_captures;
    out->graph_launches = s.graph_launches;
    out->pred_blocks = s.pred_blocks;
    out->pred_misses = s.pred_misses;
    out->pred_log_err = s.pred_log_err;
    out->compress_time = s.compress_time;
    out->factor_time = s.factor_time;
    out->factor_stream_time = s.factor_stream_time;
    out->iter_iterations = s.iter_iterations;
    out->iter_precond_residual = s.iter_precond_residual;
    out->iter_residual = s.iter_residual;
    out->solve_time = s.solve_time;
    out->refine_time = s.refine_time;
    out->refine_iterations = s.refine_iterations;
    out->last_residual = s.last_residual;
    std::vector<int> cnt(64, 0);
    for (auto& nd : e.nodes()) {
      if (nd.level >= 64) continue;
      out->max_rank[nd.level] = std::max(out->max_rank[nd.level], std::max(nd.kt, nd.kb));
      out->mean_rank[nd.level] += nd.kt + nd.kb;
      cnt[nd.level] += 2;
    }
    for (int l = 0; l < 64; l++)
      if (cnt[l]) out->mean_rank[l] /= cnt[l];
    out->device_bytes = e.device_bytes();
    out->aca_row_time = s.aca_row_time;
    out->aca_col_time = s.aca_col_time;
    out->aca_launches = s.aca_launches;
    out->aca_bytes = s.aca_bytes;
    for (int c = 0; c < 4; c++) {
      out->fac_time[c] = s.fac_time[c];
      out->fac_flops[c] = s.fac_flops[c];
      out->fac_bytes[c] = s.fac_bytes[c];
      out->fac_bound_time[c] = s.fac_bound[c];
      out->fac_launches[c] = s.fac_launches[c];
    }
    out->aca_panel_grows = s.aca_panel_grows;
    out->aca_panel_bytes = s.aca_panel_bytes;
  });
}

int hodlr_block(hodlr_handle h, int which, int i0, int j0, int m, int n, double* out, int ldo) {
  return guard(h, [&] {
    if (!out && m > 0 && n > 0) throw std::invalid_argument("null output");
    h->eng->block_host(which, i0, j0, m, n, out, ldo);
  });
}

int hodlr_block_device(hodlr_handle h, int which, int i0, int j0, int m, int n, double* dout,
                       int ldo) {
  return guard(h, [&] {
    if (!dout && m > 0 && n > 0) throw std::invalid_argument("null output");
    h->eng->block_device(which, i0, j0, m, n, dout, ldo);
  });
}

int hodlr_rows(hodlr_handle h) { return h ? h->eng->rows() : HODLR_ERR_ARG; }

int hodlr_num_nodes(hodlr_handle h) { return h ? (int)h->eng->nodes().size() : HODLR_ERR_ARG; }

int hodlr_get_nodes(hodlr_handle h, int* level, int* min_i, int* split, int* max_i, int* kt,
                    int* kb) {
  return guard(h, [&] {
    const auto& nd = h->eng->nodes();
    for (size_t i = 0; i < nd.size(); i++) {
      if (level) level[i] = nd[i].level;
      if (min_i) min_i[i] = nd[i].min_i;
      if (split) split[i] = nd[i].split;
      if (max_i) max_i[i] = nd[i].max_i;
      if (kt) kt[i] = nd[i].kt;
      if (kb) kb[i] = nd[i].kb;
    }
  });
}

int hodlr_get_factor(hodlr_handle h, int idx, int which, double* out) {
  int rank = -1;
  int rc = guard(h, [&] {
    if (!out) throw std::invalid_argument("null output");
    h->eng->get_factor(idx, which, out);
    const auto& nd = h->eng->nodes()[idx];
    rank = which < 2 ? nd.kt : nd.kb;
  });
  return rc == HODLR_OK ? rank : rc;
}

const char* hodlr_last_error(hodlr_handle h) {
  return h ? h->err.c_str() : g_create_err.c_str();
}

const char* hodlr_version(void) { return "dense_hodlr_amd 0.1 (gfx950)"; }

int hodlr_measure_device_peaks(hodlr_device_peaks* out) {
  if (!out) return fail(nullptr, HODLR_ERR_ARG, "null peaks pointer");
  if (check_device() != 0) return fail(nullptr, HODLR_ERR_DEVICE, "no HIP device available");
  try {
    const hodlr::DevicePeaks p = hodlr::measure_device_peaks();
    out->hbm_read_gbs = p.hbm_read_gbs;
    out->hbm_copy_gbs = p.hbm_copy_gbs;
    out->fp64_mfma_tflops = p.fp64_mfma_tflops;
    out->fp64_valu_tflops = p.fp64_valu_tflops;
    out->fp32_mfma_tflops = p.fp32_mfma_tflops;
  } catch (const hodlr::HipError& e) {
    const std::string m = e.what();
    if (m.find("hipMalloc") != std::string::npos || m.find("out of memory") != std::string::npos)
      return fail(nullptr, HODLR_ERR_NOMEM, m);
    return fail(nullptr, HODLR_ERR_DEVICE, m);
  } catch (const std::exception& e) {
    return fail(nullptr, HODLR_ERR_DEVICE, e.what());
  }
  return HODLR_OK;
}

int hodlr_memory_get_stats(hodlr_memory_stats* out) {
  if (!out) return fail(nullptr, HODLR_ERR_ARG, "null stats pointer");
  if (check_device() != 0) return fail(nullptr, HODLR_ERR_DEVICE, "no HIP device available");
  try {
    const hodlr::PoolStats p = hodlr::pool_stats();
    out->live_bytes = p.live_bytes;
    out->cached_bytes = p.cached_bytes;
    out->hits = p.hits;
    out->misses = p.misses;
    out->malloc_seconds = p.malloc_s;
    out->free_seconds = p.free_s;
  } catch (const std::exception& e) {
    return fail(nullptr, HODLR_ERR_DEVICE, e.what());
  }
  return HODLR_OK;
}

int hodlr_memory_trim(void) {
  if (check_device() != 0) return fail(nullptr, HODLR_ERR_DEVICE, "no HIP device available");
  try {
    hodlr::pool_trim();
  } catch (const std::exception& e) {
    return fail(nullptr, HODLR_ERR_DEVICE, e.what());
  }
  return HODLR_OK;
}

int hodlr_memory_cache(int on) {
  if (check_device() != 0) return fail(nullptr, HODLR_ERR_DEVICE, "no HIP device available");
  hodlr::pool_set_enabled(on != 0);
  return HODLR_OK;
}

}  // extern "C"
