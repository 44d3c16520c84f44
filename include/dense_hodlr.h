/*
 * dense_hodlr.h -- C ABI of the MI355X HODLR engine.
 *
 * Drop-in for the HODLR_Matrix path of the reference (amiraa127/Dense_HODLR,
 * include/HODLR_Matrix.hpp:38-413): kernel and dense constructors, the choice
 * of partial-pivot ACA or SVD compression, recSM_Solve, recLU_Solve and the
 * log-determinant.  Plain pointers and sizes only.  Matrices are column-major
 * FP64 with an explicit leading dimension, as Eigen::MatrixXd is.
 *
 * Conventions
 *  - Every call returns HODLR_OK (0) or a negative status; the message of the
 *    last failure is available from hodlr_last_error().  The reference prints
 *    "Error! ..." and calls exit(EXIT_FAILURE) (src/HODLR_Matrix.cpp:314-315,
 *    :418-421, :1249-1252) or asserts (:92, :806-807); this ABI returns codes.
 *  - Host pointers are read or written before the call returns.  The handle
 *    owns its device memory and one HIP stream; calls on one handle must not
 *    overlap (the reference object is not reentrant either).
 *  - Compression and factorization are lazy and cached, and setters that
 *    change them drop the cache, as in the reference (HODLR_Matrix.cpp:257-272,
 *    :520-534, :1232-1286).
 */
#ifndef DENSE_HODLR_H
#define DENSE_HODLR_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hodlr_handle_s* hodlr_handle;

enum {
  HODLR_OK = 0,
  HODLR_ERR_ARG = -1,      /* invalid argument (reference: assert)            */
  HODLR_ERR_STATE = -2,    /* call not valid in this state (reference: exit)  */
  HODLR_ERR_DEVICE = -3,   /* HIP runtime failure or no usable GPU            */
  HODLR_ERR_NOMEM = -4,    /* device allocation failed                        */
  HODLR_ERR_NUMERIC = -5,  /* non-finite result                               */
  HODLR_ERR_UNSUPPORTED = -6,
  /* refinement (MIXED, FP64_REFINED) ended above its target: the solution
   * output holds, per column (per instance column of a batch), the iterate
   * with the lowest true residual seen, and hodlr_stats.last_residual the
   * largest of those residuals */
  HODLR_ERR_NOT_CONVERGED = -7
};

/* Kernel families (K(i,j) = f(|x_i - x_j|) [+ diagonal term]). */
enum {
  HODLR_KERNEL_GAUSSIAN = 0,    /* exp(-r^2 / l^2)                         */
  HODLR_KERNEL_MATERN32 = 1,    /* (1 + sqrt3 r/l) exp(-sqrt3 r/l)         */
  HODLR_KERNEL_SE = 2,          /* exp(-r^2 / (2 l^2))                     */
  HODLR_KERNEL_IMQ = 3,         /* 1 / sqrt(1 + r^2/l^2)                   */
  HODLR_KERNEL_QUADRATIC = 4,   /* 1 + r^2/l^2                             */
  HODLR_KERNEL_EXPONENTIAL = 5  /* exp(-r/l)                               */
};

enum { HODLR_DIAG_ADD_SHIFT = 0, HODLR_DIAG_REPLACE_VALUE = 1 };

/* Low-rank method (reference strings "partialPiv_ACA" and "SVD",
 * src/HODLR_Matrix.cpp:332, :356). */
enum { HODLR_LR_ACA = 0, HODLR_LR_SVD = 1 };

/* Arithmetic of the factorization and solve.
 *  FP64          the reference's recSM / recLU in FP64 (no refinement);
 *  MIXED         FP32 bulk (U^SM, V panels, their GEMMs), FP64 leaf LU,
 *                node matrices and residuals, GMRES-based iterative
 *                refinement against the FP64 HODLR operator;
 *  FP64_REFINED  the FP64 factorization used as the preconditioner of the
 *                same refinement. */
enum { HODLR_PREC_FP64 = 0, HODLR_PREC_MIXED = 1, HODLR_PREC_FP64_REFINED = 2 };

/* Which factorization (recSM: Woodbury perturbI nodes; recLU: block LU). */
enum { HODLR_FACTOR_SM = 0, HODLR_FACTOR_LU = 1 };

/* Typed kernel descriptor: replaces the host callback
 * double(*)(int,int,void*) of include/kernel.hpp:27, which cannot run on the
 * GPU.  points is n*dim doubles, point-major (x0,y0,x1,y1,...), copied at
 * creation. */
typedef struct hodlr_kernel_desc {
  int family;
  int dim;              /* 1, 2 or 3 */
  const double* points;
  double length_scale;
  double diag;          /* shift added to K(i,i), or its value (diag_mode) */
  int diag_mode;        /* HODLR_DIAG_ADD_SHIFT or HODLR_DIAG_REPLACE_VALUE */
} hodlr_kernel_desc;

/* Per-call statistics (replaces the reference's clock() timers and
 * per-level rank averages, HODLR_Matrix.cpp:260-264, :424-427, :819-823).
 * Times are wall-clock seconds measured with HIP events. */
typedef struct hodlr_stats {
  int num_levels;
  int num_nodes;          /* non-leaf nodes */
  int num_leaves;
  int aca_iterations;     /* ACA iterations of the batched compression */
  double compress_time;   /* LR_ComputationTime */
  double factor_time;     /* recSM/recLU factorization time: first to last kernel of the
                             factorization, the roots' capacitance LU included (it runs on a
                             second stream and a following solve overlaps it) */
  double solve_time;      /* last solve */
  double refine_time;     /* mixed precision: refinement part of the solve */
  int refine_iterations;
  double last_residual;   /* mixed precision: final relative residual */
  int max_rank[64];       /* per level */
  double mean_rank[64];   /* per level */
  long long device_bytes; /* device memory held by the handle */
  /* with hodlr_set_profiling(h, 1): summed HIP-event durations of the ACA
   * row-phase and column-phase launches, their count (iterations), and the
   * algorithmic bytes of those residual passes, 8 (m+n) k(k+1)/2 per block */
  double aca_row_time;
  double aca_col_time;
  int aca_launches;
  double aca_bytes;
  /* with a rank predictor attached (hodlr_set_rank_predictor): blocks sized
   * by it, blocks whose ACA rank exceeded the predicted capacity (the
   * adaptive growth took over), mean |log predicted - log ACA rank| */
  int pred_blocks;
  int pred_misses;
  double pred_log_err;
  /* with hodlr_set_graphs(h, 1): solve graphs captured and graph launches */
  int graph_captures;
  int graph_launches;
  /* launch shapes in use (hodlr_autotune): ACA panel rows per thread, grouped
   * GEMM tile rows (0: per launch), split-K tile target, LU panel width, and
   * whether they come from the autotuner */
  int tune_aca_rows_per_thread;
  int tune_gemm_tile_m;
  int tune_gemm_split_target;
  int tune_lu_panel;
  int tuned;
  int tune_gemm_mfma;     /* grouped GEMMs on the matrix cores (MFMA) */
  double factor_stream_time; /* the factorization up to its last kernel on the engine's
                                stream, where a following solve starts */
  int tune_leaf_solve;    /* leaf solves of 64 or more columns: 0 one wave per 8 columns,
                             1 blocked, 2 factor shared per workgroup, 3 matrix cores
                             (the default), 4 one wave per 16 columns */
  /* the last iterative solve: GMRES iterations, the final preconditioned
   * relative residual ||M^-1 (b - A x)|| / ||M^-1 b|| it stopped on, and the
   * true one ||b - A x|| / ||b|| (largest column, FP64) */
  int iter_iterations;
  double iter_precond_residual;
  double iter_residual;
  /* with hodlr_set_profiling(h, 1), the last FP64 factorization by kernel
   * group, index 0: the grouped GEMMs of the node steps (V^T U^SM products
   * and U^SM -= U^SM' Y updates), 1: the batched LU factorizations (leaves
   * and node matrices), 2: the leaf solves of the ancestors' U^SM columns,
   * 3: the node matrices' solves.  Summed HIP-event durations per group
   * launch, algorithmic flops (2 m n k per product, (2/3) n^3 per LU, 2 n^2
   * per solved column) and bytes (operands once, an updated output read and
   * written), the sum over launches of max(flops / 78.6 TF/s, bytes / 8 TB/s)
   * (the time at MI355X's FP64 and HBM peaks), and the launch count */
  double fac_time[4];
  double fac_flops[4];
  double fac_bytes[4];
  double fac_bound_time[4];
  int fac_launches[4];
  /* the last ACA compression: panel growths (a level's U/V panels
   * reallocated and copied because a block outgrew their capacity: the
   * capacity starts at 32 columns, or at the rank predictor's sizes) and the
   * bytes of the U/V panels it ended with */
  int aca_panel_grows;
  double aca_panel_bytes;
} hodlr_stats;

/* One node of a user-defined partition: user_IndexTree::node
 * (include/user_IndexTree.hpp:10-24) with its children as indices into the
 * node array (node 0 is the root).  split is the last row of the node's top
 * (left) part; a node with split -1 and at most leaf rows is a leaf; a node
 * without children (left = right = -1) gets default subtrees below its
 * split, as the reference's createFromUsrTree does (HODLR_Tree.cpp:207-321);
 * a node must have both children or none.  Rank bounds of the node's top
 * and bottom off-diagonal blocks (-1: none; a maximum of 0 stores a zero
 * block) and its method (HODLR_LR_ACA, HODLR_LR_SVD, or -1 for the handle's
 * method) correspond to topOffDiag_minRank .. LR_Method. */
typedef struct hodlr_tree_node {
  int split;
  int top_min_rank, bottom_min_rank;
  int top_max_rank, bottom_max_rank;
  int lr_method;
  int left, right;
} hodlr_tree_node;

/* ---- construction ------------------------------------------------------ */

/* HODLR_Matrix(int nRows, int nCols, double(*)(int,int,void*), void*, int leaf)
 * (include/HODLR_Matrix.hpp:162, src/HODLR_Matrix.cpp:184-192) with a typed
 * kernel.  leaf is the size threshold (default 30 in the reference). */
int hodlr_create_kernel(hodlr_handle* h, int n, const hodlr_kernel_desc* kernel, int leaf);

/* The same constructor with the reference's host callback: entries are
 * evaluated on the host (n*n calls, as kernelMatrix::block does,
 * src/kernel.cpp:19-25) into a dense copy that is uploaded. */
int hodlr_create_kernel_callback(hodlr_handle* h, int n, double (*fn)(int, int, void*),
                                 void* data, int leaf);

/* HODLR_Matrix(Eigen::MatrixXd&, int leaf = 30, std::string LR = "partialPiv_ACA")
 * (include/HODLR_Matrix.hpp:76, src/HODLR_Matrix.cpp:89-98): copies A. */
int hodlr_create_dense(hodlr_handle* h, int n, const double* A, int lda, int leaf,
                       int lr_method);

/* HODLR_Matrix(readBinaryIntoMatrixXd(path), leaf, LR) -- how the
 * reference's drivers build from a file (src/helperFunctions.cpp:370-371):
 * the dense constructor on a matrix read from the binary format of
 * hodlr_read_binary_matrix, streamed into the handle's copy. */
int hodlr_create_dense_from_file(hodlr_handle* h, const char* path, int leaf, int lr_method);

/* The kernel and dense constructors with a user partition:
 * HODLR_Matrix(int, int, double(*)(int,int,void*), void*, int,
 * user_IndexTree&) (include/HODLR_Matrix.hpp:184) and
 * HODLR_Matrix(Eigen::MatrixXd&, int, user_IndexTree&) (:85,
 * HODLR_Matrix.cpp:100-108).  count == 0 gives the default tree. */
int hodlr_create_kernel_tree(hodlr_handle* h, int n, const hodlr_kernel_desc* kernel, int leaf,
                             const hodlr_tree_node* nodes, int count);
int hodlr_create_dense_tree(hodlr_handle* h, int n, const double* A, int lda, int leaf,
                            const hodlr_tree_node* nodes, int count);

/* kd order of npoints points of dim coordinates (point-major), the
 * reference's get_KDTree_Sorted(locations, 0) (rigidBody/MobFunctions.h:
 * 334-369: merge sort by coordinate depth % dim, halves of n/2 and the rest,
 * equal keys taking the right list first): perm[i] is the point at position
 * i.  Host only. */
int hodlr_kd_sort(const double* points, int npoints, int dim, int* perm);
/* The same order and its partition, get_KDTree_Sorted(locations,
 * pointsPerSphere, rankUpperBound, LR_Method) (MobFunctions.h:371-430): one
 * node per segment (2 npoints - 1 of them, root first) with split =
 * unknowns_per_point * (first point of the right part) - 1, both blocks'
 * maximum rank max_rank, method lr_method; single points are leaves.  nodes
 * holds 2 npoints - 1 entries; *count receives the number written.  Host
 * only. */
int hodlr_kd_tree(const double* points, int npoints, int dim, int unknowns_per_point, int max_rank,
                  int lr_method, int* perm, hodlr_tree_node* nodes, int* count);

/* ---- tree surgery (HODLR_Matrix.hpp:256-261) ---------------------------- */

/* topDiag() (side 0) and bottDiag() (side 1) (HODLR_Matrix.cpp:1551-1603):
 * a new handle for the root's top-left or bottom-right diagonal block, with
 * that part of the tree and of the compressed factors (compressing h first
 * if needed; nothing is recompressed), the block's source data (the dense
 * block, or the kernel on its points) and h's settings.  h is unchanged.
 * A root that is a leaf returns HODLR_ERR_STATE (the reference returns a
 * copy or an empty matrix with a warning); a batch HODLR_ERR_ARG. */
int hodlr_diag_block(hodlr_handle h, int side, hodlr_handle* out);
/* splitAtTop(self, topHODLR, bottHODLR) (HODLR_Matrix.hpp:261,
 * HODLR_Matrix.cpp:1507-1549): both diagonal blocks at once.  Unlike the
 * reference, which deletes self's tree, h stays usable. */
int hodlr_split_at_top(hodlr_handle h, hodlr_handle* top, hodlr_handle* bottom);

/* Releases all resources (~HODLR_Matrix). */
int hodlr_destroy(hodlr_handle h);

/* ---- settings ---------------------------------------------------------- */

/* set_LRMethod (HODLR_Matrix.hpp:225, HODLR_Matrix.cpp:1277-1286). */
int hodlr_set_lr_method(hodlr_handle h, int lr_method);
/* set_LRTolerance (HODLR_Matrix.hpp:222, HODLR_Matrix.cpp:1248-1257); default 1e-6. */
int hodlr_set_tolerance(hodlr_handle h, double tol);
/* set_minPivot (HODLR_Matrix.hpp:223, HODLR_Matrix.cpp:1259-1268).  The ACA
 * parameter is an int in the reference (include/lowRank.hpp:25): the value is
 * truncated toward zero. */
int hodlr_set_min_pivot(hodlr_handle h, double min_pivot);
/* Factorization arithmetic (HODLR_PREC_*).  No reference counterpart. */
int hodlr_set_precision(hodlr_handle h, int precision);
/* Refinement (MIXED, FP64_REFINED): relative residual ||b - A_H x|| / ||b||
 * every column must reach (default 1e-10) and the maximum number of steps
 * (default 30); a solve that ends above the target returns
 * HODLR_ERR_NOT_CONVERGED with the best iterate in its output.  Residuals are exact to rounding: evaluated in double-double
 * arithmetic, as hodlr_residual does, or, after a step at most 1e-3 of the
 * iterate (targets >= 1e-12), updated by the FP64 product of the exact step
 * (TwoSum).  FP64_REFINED takes steps x += M^-1 r while each gains 8x and
 * then flexible GMRES (restart length min(20, max_iters)); MIXED uses
 * flexible GMRES throughout. */
int hodlr_set_refinement(hodlr_handle h, double target, int max_iters);

/* Autotuner: times candidate launch shapes on this handle's own problem
 * (ACA panel rows per thread 1/2/4 on the compression; grouped GEMM tile
 * rows, split-K target, vector or matrix-core (MFMA) FP64 kernel, LU panel
 * width and the leaf-solve variant on the FP64 factorization), keeps
 * the fastest, and applies them to this handle and to every handle of the
 * same shape (instance sizes, leaf sizes) created later in the process.
 * Drops cached compression and factorization.  No reference counterpart. */
int hodlr_autotune(hodlr_handle h);

/* Fixes the launch shapes instead of the autotuner's or the defaults (2, 0,
 * 2048, 64, 1, 3): ACA panel rows per thread (1, 2, 4), grouped GEMM tile rows
 * (0: per launch, 64, 128), split-K target tiles, LU panel width (1..64),
 * grouped GEMMs (FP64, and the mixed path's FP32) on the matrix cores (1) or
 * the vector units (0), and the
 * variant of the leaf solves of 64 or more columns (the factorization's leaf
 * steps, solves of many right-hand sides): 0 one launch, one wave per column
 * chunk, 1 blocked (diagonal blocks by substitution, the rest by GEMM
 * updates), 2 one launch with the factor shared by the waves of a workgroup,
 * 3 (the default) one launch on the matrix cores, the factor shared by a
 * workgroup (leaves of at most 256 rows; others as 0), 4 as 0 with 16
 * columns per wave (leaves of at most 256 rows; others as 0; bit-identical
 * to 0).  Drops cached compression and factorization.
 * No reference counterpart. */
int hodlr_set_launch_shapes(hodlr_handle h, int aca_rows_per_thread, int gemm_tile_m,
                            int gemm_split_target, int lu_panel, int gemm_mfma, int leaf_solve);

/* Replay FP64 solves from captured hipGraphs (off by default).  The level
 * sequence of a recSM/recLU solve with a given number of right-hand sides
 * (every kernel launch of every level) is captured once per factorization
 * and replayed by one graph launch; the FP64-preconditioned refinement
 * (HODLR_PREC_FP64_REFINED) replays it for every preconditioner
 * application.  Results are identical to the eager solve.  No reference
 * counterpart. */
int hodlr_set_graphs(hodlr_handle h, int on);

/* Record HIP events around the ACA residual launches (for roofline figures);
 * off by default.  No reference counterpart. */
int hodlr_set_profiling(hodlr_handle h, int on);

/* ---- compute ----------------------------------------------------------- */

/* storeLRinTree (HODLR_Matrix.hpp:210, HODLR_Matrix.cpp:257-272). */
int hodlr_compress(hodlr_handle h);
/* recSM_Factorize / recLU_Compute (HODLR_Matrix.hpp:217, .cpp:505-534, :794-803). */
int hodlr_factorize(hodlr_handle h, int which);
/* recSM_Solve (HODLR_Matrix.hpp:215, HODLR_Matrix.cpp:805-841): X = A^-1 B,
 * B is n x nrhs (ldb), X is n x nrhs (ldx); host pointers. */
int hodlr_solve_sm(hodlr_handle h, const double* B, int ldb, int nrhs, double* X, int ldx);
/* recLU_Solve (HODLR_Matrix.hpp:214, HODLR_Matrix.cpp:757-792). */
int hodlr_solve_lu(hodlr_handle h, const double* B, int ldb, int nrhs, double* X, int ldx);
/* recSM/recLU solve with device pointers (which = HODLR_FACTOR_SM / _LU). */
int hodlr_solve_device(hodlr_handle h, int which, const double* dB, int ldb, int nrhs,
                       double* dX, int ldx);
/* iterative_Solve(input_RHS, maxIterations, stop_tolerance, init_LRTolerance,
 * input_LRMethod, directSolve_Method) (HODLR_Matrix.hpp:219,
 * HODLR_Matrix.cpp:1138-1213): sets the handle's tolerance to
 * init_lr_tolerance and its method to lr_method (kept afterwards, as in the
 * reference), compresses and factors (direct_method HODLR_FACTOR_LU, the
 * reference's "recLU", or HODLR_FACTOR_SM), then runs GMRES with restart 32
 * on the exact operator -- the dense matrix, or for kernel handles the
 * kernel evaluated on the device -- left-preconditioned by that HODLR solve
 * (IML/include/gmres.h:80-151), from x0 = M^-1 b, until every column's
 * ||M^-1 (b - A x)|| / ||M^-1 b|| is below stop_tolerance as gmres.h
 * decides it: by the Arnoldi estimate within a cycle (:123-129) or by the
 * residual at a restart (:135-140).  The reference ignores its maxIterations
 * and stop_tolerance arguments and always uses 1000 and 1e-10; here they are
 * honoured (pass 1000 and 1e-10 for its behaviour).  A column that has not
 * converged after max_iterations returns HODLR_ERR_NOT_CONVERGED with the
 * last iterate in X (the reference returns the iterate without a status).
 * hodlr_get_stats reports the iterations, the preconditioned residual
 * evaluated at the end and the true residual ||b - A x|| / ||b|| (both can
 * sit above a tolerance near the rounding floor of their FP64 evaluation,
 * eps ||A|| ||x|| / ||b||).  Host pointers; _device takes device pointers. */
int hodlr_iterative_solve(hodlr_handle h, const double* B, int ldb, int nrhs, double* X, int ldx,
                          int max_iterations, double stop_tolerance, double init_lr_tolerance,
                          int lr_method, int direct_method);
int hodlr_iterative_solve_device(hodlr_handle h, const double* dB, int ldb, int nrhs, double* dX,
                                 int ldx, int max_iterations, double stop_tolerance,
                                 double init_lr_tolerance, int lr_method, int direct_method);
/* Y = A X with the exact operator the iterative solver uses (the dense
 * matrix, or the kernel evaluated on the fly), device pointers.  No
 * reference counterpart (kernelMatrix::operator*, src/kernel.cpp:27-35, on
 * the host). */
int hodlr_exact_apply_device(hodlr_handle h, const double* dX, int ldx, int nrhs, double* dY,
                             int ldy);
/* logAbsDeterminant (HODLR_Matrix.hpp:249, HODLR_Matrix.cpp:1666-1697). */
int hodlr_logabsdet(hodlr_handle h, double* out);
/* determinant (HODLR_Matrix.hpp:248): the product of the LU diagonals without
 * the permutation sign, as the reference computes it (.cpp:1684-1688). */
int hodlr_determinant(hodlr_handle h, double* out);
/* Y = A_H X with the compressed FP64 operator (leaves plus U V^T, the
 * operator that fill_Block reconstructs, HODLR_Matrix.cpp:1404-1480). */
int hodlr_matvec(hodlr_handle h, const double* X, int ldx, int nrhs, double* Y, int ldy);

/* Y = A_H X on device pointers with no device allocation: no counterpart in
 * the reference; for callers that cannot allocate (code under stream
 * capture, serving loops, harnesses that forbid hipMalloc).  The result is
 * bit-identical to hodlr_matvec's.  Every temporary (projections V^T X of a
 * level, split-K partials, launch descriptors) lives in one 64 MiB device
 * array per GPU, held until the process ends: calls whose temporaries need
 * more (many right-hand sides on a large operator) return HODLR_ERR_ARG.
 * The operator must be compressed already (hodlr_compress), else
 * HODLR_ERR_STATE.  Calls that overlap on one GPU share that array and race:
 * serialize them. */
int hodlr_matvec_device_noalloc(hodlr_handle h, const double* dX, int ldx, int nrhs, double* dY,
                                int ldy);

/* R = B - A_H X evaluated in double-double arithmetic (error-free products
 * and sums) and rounded once: the residual that refinement targets are
 * checked with.  No reference counterpart. */
int hodlr_residual(hodlr_handle h, const double* B, int ldb, int nrhs, const double* X, int ldx,
                   double* R, int ldr);

/* ---- matrix files (src/matrixIO.cpp) ------------------------------------ */

/* Binary format of saveMatrixXdToBinary / readBinaryIntoMatrixXd
 * (include/matrixIO.hpp:42, :51; src/matrixIO.cpp:165-208): int32 rows, int32
 * cols, then rows x cols doubles in row-major order, native byte order.  With
 * out == NULL only rows and cols are returned; else out (column-major,
 * ld >= rows) receives the matrix.  Failures (file missing, short or of
 * another size) return HODLR_ERR_ARG with the message in
 * hodlr_last_error(NULL); the reference prints "Error! File ... could not be
 * opened" and exits. */
int hodlr_read_binary_matrix(const char* path, int* rows, int* cols, double* out, int ld);
int hodlr_write_binary_matrix(const char* path, int rows, int cols, const double* A, int lda);
/* Text format of readTxtIntoMatrix (matrixIO.hpp:20, matrixIO.cpp:11-50): a
 * line "rows cols", then lines "i j value" (0-based); entries not listed are
 * zero (the reference leaves them uninitialized).  Same calling convention
 * as hodlr_read_binary_matrix. */
int hodlr_read_text_matrix(const char* path, int* rows, int* cols, double* out, int ld);

/* ---- batched instances (no reference counterpart) ----------------------- */

/* count independent kernel systems in one handle, for many small or uneven
 * problems (e.g. Gaussian-process covariance systems): instance q has n[q]
 * points and its own kernel descriptor; all share the leaf size.  They form
 * one block-diagonal system of order sum n[q] whose rows are the instances'
 * rows in order, and every call on the handle (compress, factorize, solves,
 * refinement, matvec, residual) processes all instances level-synchronously:
 * one launch per kernel and tree level serves the whole batch.  Right-hand
 * sides and solutions are the stacked instance vectors; with refinement each
 * instance reaches the target on its own. */
int hodlr_create_batch(hodlr_handle* h, int count, const int* n, const hodlr_kernel_desc* kernels,
                       int leaf);
/* Number of instances (1 for the single-matrix constructors). */
int hodlr_batch_count(hodlr_handle h);
/* First row of each instance in the stacked layout, plus the total
 * (count + 1 entries). */
int hodlr_batch_offsets(hodlr_handle h, int* row0);
/* log|det| of each instance (count entries), as hodlr_logabsdet computes it
 * for one matrix. */
int hodlr_logabsdet_batch(hodlr_handle h, double* out);

/* ---- learned rank predictor (no reference counterpart) ------------------ */

/* A small MLP, 8 -> 64 -> 64 -> 1 with GELU, that predicts the log of a
 * block's ACA rank from its geometry (the 8 features of
 * hodlr_block_features).  FP32 parameters on the GPU; training (Adam on the
 * mean squared error) and inference run in the library's own HIP kernels,
 * inference in BF16.  Errors of the model calls are reported by
 * hodlr_last_error(NULL). */
typedef struct hodlr_rank_model_s* hodlr_rank_model;

/* Glorot-uniform weights, zero biases, from a seeded generator. */
int hodlr_rank_model_create(hodlr_rank_model* m, unsigned long long seed);
int hodlr_rank_model_destroy(hodlr_rank_model m);
/* Parameter count (4801): W1[64][8], b1[64], W2[64][64], b2[64], W3[64], b3. */
int hodlr_rank_model_num_params(void);
int hodlr_rank_model_get_params(hodlr_rank_model m, float* out);
/* Sets the parameters and clears the optimizer state. */
int hodlr_rank_model_set_params(hodlr_rank_model m, const float* in);
/* y[i] = f(X[i, :]) for n samples of 8 features (row-major), FP32. */
int hodlr_rank_model_forward(hodlr_rank_model m, const float* X, int n, float* y);
/* The same in BF16 arithmetic (parameters and activations rounded to
 * bfloat16, FP32 accumulation): the inference used to size compression. */
int hodlr_rank_model_predict_bf16(hodlr_rank_model m, const float* X, int n, float* y);
/* Gradient of mean((f(X) - t)^2) over the n samples (num_params floats) and
 * the loss. */
int hodlr_rank_model_gradient(hodlr_rank_model m, const float* X, const float* t, int n,
                              float* grad, float* loss);
/* Adam (lr, beta 0.9 / 0.999) on mini-batches of a seeded permutation per
 * epoch; epoch_loss (epochs entries, may be NULL) gets the full-data loss. */
int hodlr_rank_model_train(hodlr_rank_model m, const float* X, const float* t, int n, int epochs,
                           int batch, float lr, unsigned long long seed, float* epoch_loss);
/* Features of every off-diagonal block of a kernel handle, 2 per node in
 * hodlr_get_nodes order (top block, then bottom), 8 floats each:
 * log(m + n), centroid distance / (radius_rows + radius_cols),
 * radius_rows / l, radius_cols / l, log(l), log10(tol), level, family. */
int hodlr_block_features(hodlr_handle h, double tol, float* out);
/* Attach a copy of the model's current parameters (NULL detaches): the next
 * compression sizes every level's U/V panels from the predicted ranks times
 * (1 + margin) before ACA starts; ACA still decides every rank, and blocks
 * whose rank exceeds the prediction grow the panels as without a predictor.
 * Drops cached compression, like the other setters. */
int hodlr_set_rank_predictor(hodlr_handle h, hodlr_rank_model m, double margin);

/* ---- inspection -------------------------------------------------------- */

int hodlr_get_stats(hodlr_handle h, hodlr_stats* out);
/* rows() / cols() (HODLR_Matrix.hpp:246-247). */
int hodlr_rows(hodlr_handle h);
/* Number of non-leaf nodes, and per node (level order, left to right) its
 * level, index range and the two off-diagonal ranks. */
int hodlr_num_nodes(hodlr_handle h);
int hodlr_get_nodes(hodlr_handle h, int* level, int* min_i, int* split, int* max_i, int* rank_top,
                    int* rank_bottom);
/* Copy of one low-rank factor of node idx: which 0 top U, 1 top V, 2 bottom
 * U, 3 bottom V (topOffDiagU/V, bottOffDiagU/V, HODLR_Tree.hpp:41-44); out is
 * rows x rank column-major.  Returns the rank or a negative status. */
int hodlr_get_factor(hodlr_handle h, int idx, int which, double* out);
/* Entry access: block(min_i, min_j, numRows, numCols), row(i) = block(i, 0,
 * 1, n) and col(j) = block(0, j, n, 1) (HODLR_Matrix.hpp:252-254,
 * HODLR_Matrix.cpp:1396-1480).  which = HODLR_BLOCK_SOURCE returns entries of
 * the matrix the handle was built from (its dense copy, or the kernel
 * evaluated on the device), as the reference's block() does when it holds
 * that data; HODLR_BLOCK_COMPRESSED returns the compressed operator A_H
 * (leaves plus U V^T, the reference's fill_Block), compressing first if
 * needed.  out is m x n column-major with leading dimension ldo >= m: host
 * memory (hodlr_block) or device memory (hodlr_block_device).  For a batch
 * the matrix is the block-diagonal system of the instances. */
enum { HODLR_BLOCK_SOURCE = 0, HODLR_BLOCK_COMPRESSED = 1 };
int hodlr_block(hodlr_handle h, int which, int i0, int j0, int m, int n, double* out, int ldo);
int hodlr_block_device(hodlr_handle h, int which, int i0, int j0, int m, int n, double* dout,
                       int ldo);
/* Message of the last failure on this handle (or of the last failed create
 * when h is NULL). */
const char* hodlr_last_error(hodlr_handle h);
/* Library version string. */
const char* hodlr_version(void);

/* ---- device ceilings ---------------------------------------------------- */

/* Sustained throughput of the current device, measured by probe kernels:
 * a streaming read of 1 GiB (the access pattern of the ACA panel passes), a
 * 512 MiB device copy (bytes read plus written), and back-to-back
 * v_mfma_f64_16x16x4_f64, v_fma_f64 and v_mfma_f32_16x16x4_f32 from
 * registers; best of five launches each.  These are the practical ceilings
 * the engine's kernels are measured against beside the datasheet peaks (HBM
 * 8 TB/s, FP64 78.6 TFLOP/s, FP32 157.3 TFLOP/s).  Takes about half a second
 * and allocates 1 GiB for its duration.  No reference counterpart. */
typedef struct hodlr_device_peaks {
  double hbm_read_gbs;
  double hbm_copy_gbs;
  double fp64_mfma_tflops;
  double fp64_valu_tflops;
  double fp32_mfma_tflops;
} hodlr_device_peaks;

int hodlr_measure_device_peaks(hodlr_device_peaks* out);

/* ---- device memory cache ----------------------------------------------- */

/* Device blocks released by handles (destroyed, grown, refactored) are kept
 * per device and reused for later requests they fit, so repeated handles of
 * one shape stop calling hipMalloc/hipFree; the cache holds at most half of
 * the device's memory and is given back when an allocation fails.
 * HODLR_NO_MEMORY_CACHE in the environment starts with the cache off.  No
 * reference counterpart. */
typedef struct hodlr_memory_stats {
  long long live_bytes;   /* in use by handles */
  long long cached_bytes; /* released, kept for reuse */
  long long hits;         /* requests served from the cache */
  long long misses;       /* requests that called hipMalloc */
  double malloc_seconds;  /* host time spent in hipMalloc */
  double free_seconds;    /* host time spent in hipFree */
} hodlr_memory_stats;

/* Counters of the current device. */
int hodlr_memory_get_stats(hodlr_memory_stats* out);
/* Frees every cached block of the current device (after it is idle). */
int hodlr_memory_trim(void);
/* on 0: releases free at once and the caches of all devices are emptied;
 * on 1: releases are cached again. */
int hodlr_memory_cache(int on);

#ifdef __cplusplus
}
#endif

#endif /* DENSE_HODLR_H */
