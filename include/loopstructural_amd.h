/*
 * loopstructural_amd: MI355X-native discrete implicit interpolation.
 *
 * C ABI of libloopstructural_amd.so.  Plain pointers and sizes only.  Every
 * array argument is DEVICE memory unless stated, every call takes the HIP
 * stream to run on (NULL = default stream) and returns an int status
 * (LSAMD_OK = 0).  The message of the last failing call on the calling thread
 * is returned by lsamd_last_error().  The caller owns every buffer it passes;
 * workspace lives behind an explicit create/destroy handle (lsamd_fd_system,
 * lsamd_batch).  Node ids are int64 at the ABI and F-order (x fastest):
 * g = i + nx*j + nx*ny*k.
 *
 * Reference interfaces replaced (paths relative to the LoopStructural
 * checkout, LI = packages/loop_interpolation/src/loop_interpolation,
 * LC = packages/loop_common/src/loop_common):
 *   lsamd_fd_locate        LC/supports/_3d_base_structured.py:251-293,386-412
 *                          LC/supports/_3d_structured_grid.py:90-162,393-454
 *   lsamd_fd_build_rows    LI/_finite_difference_interpolator.py:236-269,342-476,478-539
 *                          LI/_discrete_interpolator.py:316-389 (row normalisation)
 *                          LC/math/_maths.py:276-332 (strike/dip basis)
 *   lsamd_fd_system_*      LI/_finite_difference_interpolator.py:643-692,1052-1164 (regulariser)
 *                          LI/_discrete_interpolator.py:684-755,815-833 (build_matrix, scaling)
 *                          LI/_solver_pipeline.py:86-124 (ridge, column scaling)
 *                          LI/_solver_strategy.py:48-71 (cg on A.T@A)
 *   lsamd_fd_system_lsmr   LI/_solver_strategy.py:108-136 (lsmr on A S)
 *   lsamd_fd_evaluate      LC/supports/_3d_structured_grid.py:295-391
 *                          LI/_discrete_interpolator.py:1342-1379
 *   lsamd_rows_adjoint     no reference counterpart (gradient of a held-out
 *                          misfit w.r.t. constraint values, weights, normals)
 */
#ifndef LOOPSTRUCTURAL_AMD_H
#define LOOPSTRUCTURAL_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LSAMD_OK 0
#define LSAMD_ERR_ARG 1        /* invalid argument / shape */
#define LSAMD_ERR_HIP 2        /* HIP runtime error */
#define LSAMD_ERR_ALLOC 3      /* device allocation failed */
#define LSAMD_ERR_INJECTED 9   /* fault injected through LSAMD_FAULT_INJECT */

/* Precision of the stencil/row operator. */
#define LSAMD_FP64 0
#define LSAMD_FP32 1

/* Row families for lsamd_fd_build_rows. */
#define LSAMD_ROWS_VALUE 0    /* 1 row / point: trilinear N, b = v            */
#define LSAMD_ROWS_NORMAL 1   /* 3 rows / point: T_d, b = n_d (d = 0,1,2)     */
#define LSAMD_ROWS_GRADIENT 2 /* 2 rows / point: strike.T, dip.T, b = 0       */
#define LSAMD_ROWS_TANGENT 3  /* 1 row / point: t_hat.T, b = 0                */

/* Structured grid: nnodes = nsteps_cells + 1 per axis, node i at
 * origin + i * step.  With axis_nodes[0..2] all set the grid is rectilinear
 * (LC/supports/_3d_rectilinear_grid.py): node i of axis a is the DEVICE value
 * axis_nodes[a][i] (nnodes[a] strictly increasing values), origin and step
 * are not used by point location, row building and evaluation, points are
 * inside only strictly between the first and last node, local coordinates
 * are not clipped, and the regulariser of an FD system on it comes as rows
 * (its stencil weights zero, the border families excepted).  P1 meshes and
 * batches take uniform grids only (axis_nodes all NULL). */
typedef struct lsamd_grid {
  double origin[3];
  double step[3];
  int64_t nnodes[3];
  const double* axis_nodes[3];
} lsamd_grid;

/* Regulariser weights, all already squared: w2_xx = dxx^2, w2_xy = (dxy/4)^2,
 * w2_bx = dx^2 (border families, 0 when dx == 0 or nnodes[0] < 3),
 * ridge2 = ridge_factor^2 (0 when ridge is off). */
typedef struct lsamd_fd_reg {
  double w2_xx, w2_yy, w2_zz;
  double w2_xy, w2_xz, w2_yz;
  double w2_bx, w2_by, w2_bz;
  double ridge2;
} lsamd_fd_reg;

/* PCG options.  The stopping test is scipy's on the column-scaled system:
 * |S r| < max(atol, rtol |S rhs|), S = diag(N)^-1/2 (Jacobi), checked before
 * each iteration; maxiter <= 0 means 10 * dof. */
typedef struct lsamd_solve_opts {
  double rtol;
  double atol;
  int64_t maxiter;
  int32_t precision;    /* LSAMD_FP64: plain FP64 PCG.
                           LSAMD_FP32: FP32 operator/vectors, FP64 dots, FP64
                           residual recomputation + outer refinement. */
  int32_t check_every;  /* iterations per captured graph chunk (<=0: 64)   */
  double inner_rtol;    /* FP32 inner solve tolerance (<=0: 1e-4)          */
  int64_t max_outer;    /* refinement sweeps (<=0: 50)                     */
  int32_t jacobi;       /* 1: Jacobi preconditioner = the reference's column
                           scaling (apply_scaling_matrix=True); 0: plain CG
                           with the 2-norm stopping test; 2: geometric
                           multigrid preconditioned flexible CG (FD systems
                           only; FP32 V-cycle with an additive-Schwarz
                           Chebyshev smoother, FP64 outer iteration, the same
                           Jacobi-scaled stopping test; precision, inner_rtol
                           and max_outer unused; check_every capped at 8.
                           Environment: LSAMD_MG_SMOOTHER=jacobi for point
                           Jacobi, LSAMD_MG_DEGREE=1..8 for the Chebyshev
                           degree of the coarse levels (4),
                           LSAMD_MG_DEGREE0 for the fine level's (2, at most
                           the coarse one), LSAMD_MR_DEGREE=1..8 for the
                           coarse levels of the multi-rhs V-cycle (6),
                           LSAMD_MG_TIMING=1 prints set-up
                           times, LSAMD_MR_COV_LANES=1|2|4|8 sets the
                           right-hand sides per thread of the multi-rhs
                           covered-node smoothing (2)).
                           No reference counterpart: the reference solves
                           with Jacobi-scaled CG; the minimiser is the same. */
  int32_t pad;
} lsamd_solve_opts;

typedef struct lsamd_solve_info {
  int64_t iterations;   /* total (inner) CG iterations                      */
  int64_t outer;        /* refinement sweeps (0 for LSAMD_FP64)             */
  int32_t converged;    /* 1 if the stopping test was met                   */
  int32_t pad;
  double residual;      /* final |S r| (FP64)                               */
  double threshold;     /* max(atol, rtol |S rhs|)                          */
} lsamd_solve_info;

typedef struct lsamd_fd_system lsamd_fd_system;

const char* lsamd_last_error(void);
int lsamd_version(void);
int lsamd_device_count(int* count);

/* A2-A4: locate points in the grid.  pts (n,3) row-major FP64.  Outputs
 * (any may be NULL): idc (n,8) int64 corner node ids (-1 outside), w (n,8)
 * trilinear weights, T (n,3,8) shape-function gradients, inside (n) uint8.
 * Bit-identical to the reference's numpy arithmetic. */
int lsamd_fd_locate(const lsamd_grid* g, const double* pts, int64_t n, int64_t* idc,
                    double* w, double* T, uint8_t* inside, void* stream);

/* A5-A7: least-squares rows for n points that are INSIDE the grid.
 * pts (n, ld) FP64 row-major: x y z | value or vector ... | weight.
 * Writes rows_per_point*n rows of 8 entries, family-major (row = r*n + i):
 * cols (rows,8) int32, coef (rows,8) FP64 normalised, b (rows) normalised,
 * w (rows) = wscale * (use_point_weight ? pts[i, weight_col] : 1),
 * norm (rows) = |a| before normalisation.  GRADIENT/TANGENT normalise the
 * vector columns of pts IN PLACE (the reference does the same). */
int lsamd_fd_build_rows(const lsamd_grid* g, int32_t kind, double* pts, int64_t n, int64_t ld,
                        int32_t weight_col, double wscale, int32_t* cols, double* coef,
                        double* b, double* w, double* norm, void* stream);

/* A6: row normalisation of nrows rows of width k (k <= 128) in place:
 * a /= |a|, b /= |a| where |a| > 0, with |a| summed in numpy's pairwise
 * order (LI/_discrete_interpolator.py:361-366); norm (nullable) gets |a|.
 * Used for interface rows (k = 16) and caller-supplied rows. */
int lsamd_rows_normalise(int64_t nrows, int32_t k, double* coef, double* b, double* norm,
                         void* stream);

/* A8-A12: the normal-equations system N = A^T W^2 A + R^T W R + ridge2 I of
 * one structured grid.  scale2 (nnodes, nullable) multiplies every
 * regulariser row weight^2 at its centre node. */
int lsamd_fd_system_create(const lsamd_grid* g, const lsamd_fd_reg* reg, const double* scale2,
                           lsamd_fd_system** out);
int lsamd_fd_system_destroy(lsamd_fd_system* s);

/* Data rows in CSR form: row_ptr (nrows+1) int64, cols int32, coef FP64,
 * b (nrows), w (nrows) (weights, not squared).  Copied into the handle;
 * builds the node-major transpose, rhs and Jacobi diagonal. */
int lsamd_fd_system_set_rows(lsamd_fd_system* s, int64_t nrows, int64_t nnz,
                             const int64_t* row_ptr, const int32_t* cols, const double* coef,
                             const double* b, const double* w, void* stream);

/* y = N x in FP64 (x, y: nnodes FP64) or FP32 (x, y: nnodes float). */
int lsamd_fd_system_apply(lsamd_fd_system* s, int32_t precision, const void* x, void* y,
                          void* stream);
/* rhs = A^T W^2 b ; diag = diag(N). */
int lsamd_fd_system_rhs(lsamd_fd_system* s, double* rhs, void* stream);
int lsamd_fd_system_diag(lsamd_fd_system* s, double* diag, void* stream);
/* Solve N c = rhs (rhs NULL: the system's own).  c (nnodes) FP64, output. */
int lsamd_fd_system_solve(lsamd_fd_system* s, const lsamd_solve_opts* opts, const double* rhs,
                          double* c, lsamd_solve_info* info, void* stream);

/* Several right-hand sides of one FD system, multigrid (opts->jacobi must be
 * 2): rhs and c are (nrhs, nnodes) FP64, one vector after the other; infos
 * (nrhs) host memory.  Realisations that share the constraint locations and
 * differ in observed values have one matrix and nrhs right-hand sides; up to
 * 64 of them are solved per launch sequence (their vectors one after the
 * other), with the single-system hierarchy and stopping test per rhs.  No
 * reference counterpart (the reference solves realisations one by one). */
int lsamd_fd_system_solve_multi(lsamd_fd_system* s, const lsamd_solve_opts* opts, int32_t nrhs,
                                const double* rhs, double* c, lsamd_solve_info* infos,
                                void* stream);
/* rhs (nnodes) = A^T W^2 b for per-row values b (nrows) of the handle's
 * rows (their coefficients and weights). */
int lsamd_fd_system_rhs_rows(lsamd_fd_system* s, const double* b, double* rhs, void* stream);

/* A13: LSMR on the rectangular least-squares system of an FD handle
 * (replaces solve_with_lsmr, LI/_solver_strategy.py:108-136, i.e.
 * scipy.sparse.linalg.lsmr(A S, b, atol, btol, conlim, maxiter) on the
 * ridge-augmented, column-scaled system of LI/_solver_pipeline.py:86-124,
 * then c = S y).  Matrix-free: data rows from set_rows, regulariser rows
 * from the stencil definition (rows of one mask centred at one node merged,
 * which leaves every LSMR iterate unchanged), ridge rows.  FP64 throughout.
 * maxiter <= 0 means the number of nodes (scipy's min(m, n)); conlim <= 0
 * disables the condition test.  FD handles only (not P1 or batches). */
typedef struct lsamd_lsmr_opts {
  double atol;
  double btol;
  double conlim;
  int64_t maxiter;
  int32_t check_every;  /* iterations per captured graph chunk (<=0: 64)   */
  int32_t scaling;      /* 1: column scaling S (apply_scaling_matrix=True);
                           0: S = I                                        */
} lsamd_lsmr_opts;

typedef struct lsamd_lsmr_info {
  int64_t iterations;
  int32_t istop;        /* scipy's codes: 0 b = 0, 1 |r| small, 2 LS optimal,
                           3 cond > conlim, 4-6 the same at machine
                           precision, 7 maxiter                            */
  int32_t pad;
  double normr, normar, normA, condA, normx;
} lsamd_lsmr_info;

int lsamd_fd_system_lsmr(lsamd_fd_system* s, const lsamd_lsmr_opts* opts, double* c,
                         lsamd_lsmr_info* info, void* stream);

/* Measurement: launch the regulariser stencil kernel alone `iters` times
 * back to back (FP64 or FP32, on the handle's work vectors, after one warm
 * launch) between two HIP events on `stream`; *avg_ms = elapsed / iters. */
int lsamd_fd_system_time_stencil(lsamd_fd_system* s, int32_t precision, int32_t iters,
                                 double* avg_ms, void* stream);

/* Builds, or brings up to date with the current rows, the multigrid hierarchy
 * of an FD system (what the first multigrid solve after set_rows does), so
 * that callers can prepare many systems concurrently, one stream each,
 * before a batched solve.  No reference counterpart. */
int lsamd_fd_system_mg_build(lsamd_fd_system* s, void* stream);

/* A12 for an uneven ensemble: the multigrid-preconditioned solve of nsys
 * independent FD systems (distinct handles, rows set) in one launch sequence
 * per iteration for all of them; each system keeps its own hierarchy,
 * stopping test and iteration count (opts as lsamd_fd_system_solve with the
 * multigrid preconditioner; precision and jacobi are ignored).  c[i]: device
 * output of system i (its node count); infos[i]: host.  Replaces a loop of
 * the reference's solve_system('cg') over the ensemble
 * (LI/_discrete_interpolator.py:1221-1318, LI/_solver_strategy.py:48-71). */
int lsamd_fd_systems_solve_mg(lsamd_fd_system* const* systems, int32_t nsys,
                              const lsamd_solve_opts* opts, double* const* c,
                              lsamd_solve_info* infos, void* stream);

/* Times one kernel of the multi-right-hand-side multigrid on the fine level,
 * reusing the work of this handle's widest lsamd_fd_system_solve_multi group
 * (64 lanes, else 32; LSAMD_ERR_ARG without one) and leaving its contents
 * undefined: kernel 0 = a later Chebyshev smoothing step over every node
 * (k_mr_cheb_all; the split step of the uncovered nodes, k_mr_cheb, on grids
 * without a covered list), 1 = the update of the FP64 x and the FP32 r
 * (k_mr_upd), 2 = the regulariser stencil over the lanes.  One warm launch, then
 * `iters` launches between two HIP events on `stream`: *avg_ms per launch,
 * *bytes = algorithmic bytes per launch (every input and output element
 * once).  A measurement hook with no reference counterpart. */
int lsamd_fd_system_time_multi(lsamd_fd_system* s, int32_t kernel, int32_t iters,
                               double* avg_ms, double* bytes, void* stream);

/* Autotuner result: form of the plane-streaming stencil chosen for the
 * precision: the z-chunk 8, 16 or 32 with planes staged through registers on
 * 64 x 8 tiles, 1008 / 1016 / 1032 for those z-chunks with planes copied
 * straight into LDS, or 2008 / 2016 / 2032 for register-staged planes on
 * 32 x 16 tiles (0 = exact per-node kernel: an axis has < 7 nodes or a
 * per-node weight scale is set; also 0 before the first apply/solve), and the
 * timed forms in that order (8, 16, 32, 1008, 1016, 1032, 2008, 2016, 2032)
 * in microseconds (tune_us[9], nullable, -1 = skipped). */
int lsamd_fd_system_stencil_info(lsamd_fd_system* s, int32_t precision, int32_t* zc,
                                 float* tune_us);

/* A14: evaluate c at n points.  values (n) or grads (n,3), either NULL.
 * Values: outside -> NaN, NaN coordinate -> 0.  Gradients: outside -> 0. */
int lsamd_fd_evaluate(const lsamd_grid* g, const double* c, const double* pts, int64_t n,
                      double* values, double* grads, void* stream);

/* A7 (use_regularisation_weight_scale): squared distance from every node
 * (np.linspace coordinates) to the nearest of m points (m,3); d2 (nnodes).
 * Replaces the KDTree query of LI/_finite_difference_interpolator.py:66-88. */
int lsamd_fd_nearest_d2(const lsamd_grid* g, const double* pts, int64_t m, double* d2,
                        void* stream);

/* A20: out (nnodes) = E^T gv, E the (n x nnodes) trilinear evaluation matrix
 * of the points (outside points contribute nothing); deterministic. */
int lsamd_fd_evaluate_adjoint(const lsamd_grid* g, const double* pts, int64_t n,
                              const double* gv, double* out, void* stream);

/* Allocation-free variant of lsamd_fd_evaluate_adjoint (same output, bit for
 * bit).  No reference counterpart: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  It calls no hipMalloc/hipFree and never synchronises:
 * its entry arrays and sort storage live in one static 24 MiB scratch array
 * per GPU, held until the process ends.  At most LSAMD_NOALLOC_MAX_POINTS
 * points (LSAMD_ERR_ARG above that).  Calls on one GPU share the array, so
 * calls that overlap on one GPU (other streams, other threads) race. */
#define LSAMD_NOALLOC_MAX_POINTS 65536
int lsamd_fd_evaluate_adjoint_noalloc(const lsamd_grid* g, const double* pts, int64_t n,
                                      const double* gv, double* out, void* stream);

/* A20 on a system handle: lsamd_fd_evaluate_adjoint on the handle's grid (same
 * output, bit for bit) with its temporaries in a buffer cached on the handle
 * and grown when a call needs more, so repeated calls (one per autograd
 * backward pass) neither allocate nor synchronise after the first.  Like
 * every call on a handle, calls on one handle must not overlap. */
int lsamd_fd_system_evaluate_adjoint(lsamd_fd_system* s, const double* pts, int64_t n,
                                     const double* gv, double* out, void* stream);

/* A20: per data row r of the system (set_rows order), with u = lambda . a_r
 * and the row's normalised coefficients a_r, rhs b_r, weight w_r:
 *   g_b[r] = w_r^2 u            (d loss / d b_hat_r)
 *   g_w[r] = 2 w_r u (b_r - a_r . c)   (d loss / d w_r)
 * for the loss L with N lambda = dL/dc.  Either output may be NULL. */
int lsamd_rows_adjoint(lsamd_fd_system* s, const double* c, const double* lambda, double* g_b,
                       double* g_w, void* stream);

/* ------------------------------------------------------------------ P1 ----
 * Structured 5-tet mesh on the grid g (nnodes = cells + 1 per axis): tet id
 * = local + 5 * cell, cells x fastest; node ids as above.
 * LC/supports/_3d_structured_tetra.py, LI/_p1interpolator.py. */

/* A16: get_element_for_location (_3d_structured_tetra.py:274-365, strict
 * inside :264-272).  Outputs (any may be NULL): tet (n) int64 (-1 outside),
 * nodes (n,4) int64 tet corner ids (-1), bc (n,4) barycentric coordinates
 * (NaN), grad (n,3,4) d bc_k / d x_d (NaN; get_element_gradient_for_location
 * :502-545), inside (n) uint8.  Tet choice and bc are bit-identical to the
 * reference; grad agrees to a few ulp (adjugate vs LU inverse). */
int lsamd_p1_locate(const lsamd_grid* g, const double* pts, int64_t n, int64_t* tet,
                    int64_t* nodes, double* bc, double* grad, uint8_t* inside, void* stream);

/* A15: get_elements (:376-399) -> elements (ntet,4) and get_neighbours
 * (:626-740) -> neighbours (ntet,4), -1 = none; either may be NULL. */
int lsamd_p1_elements(const lsamd_grid* g, int64_t* elements, int64_t* neighbours, void* stream);

/* A16: get_element_gradients(tets) (:421-486) -> grads (n,3,4); size (n,
 * nullable) = element_size (:80-94).  Negative ids count from the end. */
int lsamd_p1_tet_gradients(const lsamd_grid* g, const int64_t* tets, int64_t n, double* grads,
                           double* size, void* stream);

/* A15: number of face-sharing tet pairs (host only; -1 on a bad grid). */
int64_t lsamd_p1_n_pairs(const lsamd_grid* g);

/* A15/A17: the face table (_init_face_table :120-184) and the edge-jump rows
 * (minimise_edge_jumps, LI/_p1interpolator.py:115-166), one per pair, in the
 * reference's pair order.  Any output may be NULL: pairs (P,2) (larger tet
 * first), faces (P,3) shared node ids ascending, cols (P,8) int32,
 * coef (P,8) normalised by the 8-entry norm, w (P) = cgw, or with scale (n
 * nodes, nullable) 0.5 * (mean scale over T1 + mean over T2) * cgw
 * (:150-154), norm (P) the 8-entry norm.  The face normal is taken over the
 * ascending nodes; the reference's node order may differ, which flips the
 * sign of a row and leaves its normal equations unchanged. */
int lsamd_p1_pairs(const lsamd_grid* g, int64_t* pairs, int64_t* faces, int32_t* cols,
                   double* coef, double* w, double* norm, const double* scale, double cgw,
                   void* stream);

/* A17: rows of n points INSIDE the mesh (lsamd_p1_locate), pts (n, ld) FP64:
 * x y z | value or vector.  VALUE: 1 row of 4 (bc), b = v (:98-113).
 * NORMAL: 3 rows of 4 per point, point-major (row = 3i + d), b = n_d
 * (:73-96).  TANGENT: 1 row (unit vector . grad), b = bval, weight scaled by
 * the tet volume (add_gradient_orthogonal_constraints :248-296).  Rows are
 * normalised (a/|a|, b/|a|); w = wscale (point weights are ignored, as in
 * the reference); norm (nullable) = |a|. */
int lsamd_p1_build_rows(const lsamd_grid* g, int32_t kind, const double* pts, int64_t n,
                        int64_t ld, double wscale, double bval, int32_t* cols, double* coef,
                        double* b, double* w, double* norm, void* stream);

/* A18: evaluate_value / evaluate_gradient (:208-262); NaN outside (strict)
 * for both.  values (n) and grads (n,3) nullable. */
int lsamd_p1_evaluate(const lsamd_grid* g, const double* c, const double* pts, int64_t n,
                      double* values, double* grads, void* stream);

/* A17: the P1 normal-equations system N = A^T W^2 A + cgw^2 E^T E + ridge2 I
 * with E the normalised edge-jump rows, applied matrix-free by a 21-tap table
 * stencil.  The handle works with every lsamd_fd_system_* call (set_rows,
 * apply, rhs, diag, solve, time_stencil, rows_adjoint, destroy). */
int lsamd_p1_system_create(const lsamd_grid* g, double cgw, double ridge2,
                           lsamd_fd_system** out);

/* ---------------------------------------------------------------- batch ----
 * An ensemble of FD systems with different grids (every axis >= 7 nodes) and
 * one set of regulariser weights, solved together: the node vectors are
 * concatenated (system s owns [off_s, off_s+1)), every launch covers all
 * systems, CG scalars and stopping tests are per system and converged
 * systems drop out.  No reference counterpart (the reference solves one
 * system per call, LI/_discrete_interpolator.py:1221-1319); each system's
 * result equals its single-system solve.  The handle also accepts
 * lsamd_fd_system_apply / _rhs / _diag / _time_stencil / _destroy.  The first
 * solve fixes the stencil form by timing four tile tables (64 x 8 or 32 x 16
 * tiles, z-chunk 16 or 32); LSAMD_BATCH_TILE=64|32 and LSAMD_BATCH_ZC=16|32
 * pin the tile width and the z-chunk. */
int lsamd_batch_create(int32_t nsys, const lsamd_grid* grids, const lsamd_fd_reg* reg,
                       lsamd_fd_system** out);
/* node_off (nsys+1, host): first node of each system, total last. */
int lsamd_batch_offsets(lsamd_fd_system* s, int64_t* node_off);
/* Rows of all systems as one CSR with GLOBAL node ids, system s owning rows
 * [sys_rows[s], sys_rows[s+1]) (host array, nsys+1); a row may only address
 * its own system's nodes (checked: LSAMD_ERR_ARG). */
int lsamd_batch_set_rows(lsamd_fd_system* s, const int64_t* sys_rows, int64_t nrows, int64_t nnz,
                         const int64_t* row_ptr, const int32_t* cols, const double* coef,
                         const double* b, const double* w, void* stream);
/* Solve every system (opts as lsamd_fd_system_solve, jacobi must be 1;
 * maxiter <= 0: 10 x the largest system); c (total nodes) FP64; info (nsys,
 * host, nullable) per system. */
int lsamd_batch_solve(lsamd_fd_system* s, const lsamd_solve_opts* opts, const double* rhs,
                      double* c, lsamd_solve_info* info, void* stream);

/* Introspection (host memory, no GPU work): the 21-tap coefficient table of
 * the edge-jump operator, tab[((vz*nv[1] + vy)*nv[0] + vx)*21 + q], tab of at
 * least 343*21 doubles; nv (3) the virtual axis lengths. */
int lsamd_p1_operator_table(const lsamd_grid* g, double cgw, double* tab, int32_t* nv);

#ifdef __cplusplus
}
#endif
#endif
