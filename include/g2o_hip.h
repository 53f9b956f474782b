/*
 * g2o_hip.h - C ABI of the MI355X (gfx950) implementation of g2o's
 * per-iteration Levenberg-Marquardt linear-system path for bundle adjustment
 * (BlockSolver<BlockSolverTraits<6,3>> with Schur) and 3D pose graphs
 * (6x6 blocks without Schur).
 *
 * Plain C types only: pointers to host arrays and sizes. Every call returns a
 * status code (G2OHIP_OK = 0); nothing throws across the ABI. After a failure
 * g2ohip_last_error() returns a message for that context.
 *
 * Each entry point names the g2o interface it replaces (paths relative to the
 * g2o source tree). The host-side C++ classes BlockSolverHIP / LinearSolverHIP
 * (g2o_amd/host/) implement g2o's Solver / LinearSolver contracts on top of it.
 *
 * Threading: one context is used from one host thread at a time, as g2o calls
 * its Solver (g2o/core/solver.h:44-153). Calls are stream-ordered on the
 * context's own HIP stream and synchronous on return.
 */
#ifndef G2O_HIP_H
#define G2O_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum g2ohip_status {
  G2OHIP_OK = 0,
  G2OHIP_ERR_INVALID = 1,      /* bad argument or call order */
  G2OHIP_ERR_NOT_PD = 2,       /* singular / non-positive-definite block */
  G2OHIP_ERR_NAN = 3,          /* non-finite value in the system */
  G2OHIP_ERR_OOM = 4,          /* device allocation failed */
  G2OHIP_ERR_HIP = 5,          /* HIP runtime error */
  G2OHIP_ERR_UNSUPPORTED = 6,  /* structure outside the supported limits */
  G2OHIP_ERR_NO_DEVICE = 7,    /* context created without a device */
  G2OHIP_ERR_IO = 8            /* file cannot be opened, read or written */
};

/* precision of the Hessian blocks, the Schur reduction and the PCG SpMV */
enum g2ohip_precision { G2OHIP_FP64 = 0, G2OHIP_FP32 = 1 };
/* how the reduced camera system is applied inside the PCG */
enum g2ohip_schur_mode { G2OHIP_SCHUR_EXPLICIT = 0, G2OHIP_SCHUR_IMPLICIT = 1 };

typedef struct g2ohip_ctx g2ohip_ctx;

/* Solver options. Defaults (g2ohip_default_options) follow g2o: LM tau 1e-5 and
 * maxTrialsAfterFailure 10 (optimization_algorithm_levenberg.cpp:42-56); PCG
 * tolerance 1e-6, absolute-tolerance memory on, maxIter = system size
 * (linear_solver_pcg.h:45-50). */
typedef struct g2ohip_options {
  int precision;          /* g2ohip_precision */
  int schur_mode;         /* g2ohip_schur_mode */
  double pcg_tolerance;   /* LinearSolverPCG::setTolerance */
  int pcg_max_iter;       /* LinearSolverPCG::setMaxIterations, <0: n */
  int refine_steps;       /* FP32 only: at most this many FP64 iterative-
                           * refinement corrections (default 4); they stop at
                           * the first one that finds the FP64 residual of the
                           * reduced system already within the solve's
                           * threshold. The FP32 operator keeps its diagonal
                           * blocks in FP64.
                           * If the FP32 solution's FP64 residual misses the
                           * first solve's threshold by more than 10x, or a
                           * correction stops above it, the FP32 reduced
                           * system no longer resolves its weakest modes: that
                           * solve is redone in FP64 at the same state and the
                           * rest of the optimize() call (or until the next
                           * set_options) runs in FP64 */
  int max_trials;         /* OptimizationAlgorithmLevenberg maxTrialsAfterFailure */
  double user_lambda_init;/* "initialLambda" property, <=0: tau * max diag */
  double tau;             /* OptimizationAlgorithmLevenberg::_tau */
} g2ohip_options;

/* Per-LM-iteration record, fields named after G2OBatchStatistics
 * (g2o/core/batch_stats.h:49-67) where one exists. Times are seconds. */
typedef struct g2ohip_iter_stats {
  double chi2_before;
  double chi2;                 /* after the iteration (accepted state) */
  double lambda;               /* after the iteration */
  int levenberg_iterations;    /* inner trials */
  int iterations_linear_solver;/* PCG iterations summed over trials */
  int refine_steps;            /* FP64 refinement corrections summed */
  int status;                  /* 0 OK, 1 Terminate (levenberg.cpp:149-152), 2 Fail */
  double time_residuals;
  double time_quadratic_form;
  double time_schur_complement;
  double time_linear_solver;
  double time_update;
  double time_iteration;
  int fp64_fallback;           /* 1 if an FP32 solve of this iteration was redone in FP64 */
} g2ohip_iter_stats;

void g2ohip_default_options(g2ohip_options* opt);

/* ------------------------------------------------------------ contexts */
/* device >= 0 binds a HIP device. device == -1 makes a host-only context in
 * which only g2ohip_ba_set_structure / g2ohip_ba_structure_info work (used to
 * check the structure builder without a GPU). */
int g2ohip_create(int device, g2ohip_ctx** out);
void g2ohip_destroy(g2ohip_ctx* ctx);
const char* g2ohip_last_error(const g2ohip_ctx* ctx);
/* Also resets the PCG residual memory, as LinearSolverPCG::init does when the
 * algorithm is (re)initialised. */
int g2ohip_set_options(g2ohip_ctx* ctx, const g2ohip_options* opt);
int g2ohip_get_options(const g2ohip_ctx* ctx, g2ohip_options* opt);

/* ------------------------------------------------- bundle adjustment graph */
/* SparseOptimizer::initializeOptimization + buildIndexMapping
 * (sparse_optimizer.cpp:175-321) and BlockSolver::buildStructure
 * (block_solver.hpp:110-269) for EdgeSE3ProjectXYZ edges between
 * VertexSE3Expmap cameras and marginalized VertexPointXYZ points.
 * edge_cam/edge_pt: [E] vertex indices; cam_fixed/pt_fixed: [C]/[P] 0/1 masks
 * (NULL = none fixed). An edge with both vertices fixed is inactive. */
int g2ohip_ba_set_structure(g2ohip_ctx* ctx, int num_cams, int num_pts, int num_edges, const int* edge_cam,
                            const int* edge_pt, const uint8_t* cam_fixed, const uint8_t* pt_fixed);

/* Structure summary: info[0..12] = free cameras, free points, active edges,
 * Hschur upper blocks, linearisation chunks, Schur point windows,
 * off-diagonal Schur pair products, off-diagonal Schur partial sums, max
 * slots per point, Hschur blocks in both triangles, (chunk, camera) partial
 * rows of the linearisation, points with active edges, observations with a
 * free camera and a free point (Hpl blocks). Point windows and partial sums
 * are 0 when the explicit Schur complement forms its off-diagonal blocks by
 * camera rows instead (no partial sums staged). */
int g2ohip_ba_structure_info(g2ohip_ctx* ctx, long long* info);

/* Measurements of EdgeSE3ProjectXYZ (edge_project_xyz.cpp:44-49): meas[E*2]
 * pixel (u, v); weight[E] scalar information (Omega = weight * I2, NULL = 1);
 * intr[C*4] fx fy cx cy per camera; huber_delta > 0 attaches
 * RobustKernelHuber(delta) to every edge (robust_kernel_impl.cpp:60-74). */
int g2ohip_ba_set_measurements(g2ohip_ctx* ctx, const double* meas, const double* weight, const double* intr,
                               double huber_delta);

/* BaseEdge::setInformation (g2o/core/base_edge.h:165-180) for every edge:
 * omega[E*3] = [O00 O01 O11], the upper triangle of the edge's 2x2 information
 * matrix (the three values EdgeSE3ProjectXYZ::read takes,
 * edge_project_xyz.cpp:34-42). Each must be positive definite. The edge's
 * information is weight * Omega (the weight of set_measurements scales it).
 * NULL (or all identity) resets every edge to Omega = I, which is also the
 * state after set_structure. */
int g2ohip_ba_set_information(g2ohip_ctx* ctx, const double* omega);

/* Vertex estimates: cams[C*7] = SE3Quat [tx ty tz qx qy qz qw] (world to
 * camera, se3quat.h), pts[P*3]. set_state normalizes quaternions as the
 * SE3Quat constructor does (se3quat.h:65-86). */
int g2ohip_ba_set_state(g2ohip_ctx* ctx, const double* cams, const double* pts);
int g2ohip_ba_get_state(g2ohip_ctx* ctx, double* cams, double* pts);

/* ------------------------------------ Solver contract (g2o/core/solver.h) */
/* computeActiveErrors + activeRobustChi2 (sparse_optimizer.cpp:78-130) */
int g2ohip_compute_active_errors(g2ohip_ctx* ctx, double* chi2);
/* per-edge residual/Jacobian errors: err[E*2] (zero for inactive edges) */
int g2ohip_get_errors(g2ohip_ctx* ctx, double* err);
/* BlockSolver::buildSystem (block_solver.hpp:492-565) */
int g2ohip_build_system(g2ohip_ctx* ctx);
/* BlockSolver::setLambda / restoreDiagonal (block_solver.hpp:567-606). The
 * damping is applied inside the Schur and PCG kernels; no diagonal backup. */
int g2ohip_set_lambda(g2ohip_ctx* ctx, double lambda);
int g2ohip_restore_diagonal(g2ohip_ctx* ctx);
/* BlockSolver::solve (block_solver.hpp:331-477): Schur complement, reduced
 * camera system by block-Jacobi PCG, back-substitution. *ok = 0 when the
 * system was not solvable (the LM then rejects the step). */
int g2ohip_solve(g2ohip_ctx* ctx, int* ok);
/* Solver::x() / Solver::b() (solver.h:101-108), host copies [6Cf + 3Pf] */
int g2ohip_get_x(g2ohip_ctx* ctx, double* x);
int g2ohip_get_b(g2ohip_ctx* ctx, double* b);
/* BlockSolverBase::multiplyHessian (block_solver.h:90-97, 149-151): dest =
 * Hpp src, the product with the pose-pose blocks alone, as the reference's
 * _Hpp->multiplySymmetricUpperTriangle: block diagonal for bundle adjustment
 * (no camera-camera edges), the whole Hessian of a pose graph. src and dest
 * hold the system's length ([6Cf + 3Pf], or [6Nf]); only the pose rows of
 * dest are written. Hpp is that of the last build_system, undamped
 * (set_lambda's damping lives inside the solve). */
int g2ohip_multiply_hessian(g2ohip_ctx* ctx, const double* src, double* dest);

/* SparseOptimizer::update(solver.x()) (sparse_optimizer.cpp:457-468) */
int g2ohip_update(g2ohip_ctx* ctx);
/* SparseOptimizer::push / pop / discardTop (sparse_optimizer.cpp:547-567,
 * 633-637): state stack of depth <= 4 */
int g2ohip_push(g2ohip_ctx* ctx);
int g2ohip_pop(g2ohip_ctx* ctx);
int g2ohip_discard_top(g2ohip_ctx* ctx);
/* OptimizationAlgorithmLevenberg::computeLambdaInit / computeScale
 * (optimization_algorithm_levenberg.cpp:155-176) */
int g2ohip_lambda_init(g2ohip_ctx* ctx, double* lambda);
int g2ohip_compute_scale(g2ohip_ctx* ctx, double* scale);

/* SparseOptimizer::optimize(iterations) (sparse_optimizer.cpp:393-455) with
 * the context's algorithm: OptimizationAlgorithmLevenberg
 * (optimization_algorithm_levenberg.cpp:60-153, the default) or
 * OptimizationAlgorithmGaussNewton (optimization_algorithm_gauss_newton.cpp:
 * 47-90). stats[iterations] may be NULL. *done = iterations run. */
int g2ohip_optimize(g2ohip_ctx* ctx, int iterations, g2ohip_iter_stats* stats, int* done);
/* SparseOptimizer::setAlgorithm. Gauss-Newton builds and solves the
 * undamped system each iteration (lambda 0; stats report
 * levenberg_iterations 0), updates the state and stops after the first failed
 * solve (status 1). A failed solve leaves the state unchanged, where the
 * reference applies whatever the solver left in x. Dogleg
 * (optimization_algorithm_dogleg.cpp:62-213, default properties: initialDelta
 * 1e4, maxTrialsAfterFailure 100, initialLambda 1e-7, lambdaFactor 10) uses
 * BlockSolver::multiplyHessian's product with Hpp alone (block_solver.h:
 * 149-151), whose landmark part is zero for bundle adjustment, as the
 * reference does; its stats report the trust-region radius in lambda, the
 * tries in levenberg_iterations, and status 1 Terminate / 2 Fail (damping
 * limit). */
enum g2ohip_algorithm { G2OHIP_LEVENBERG = 0, G2OHIP_GAUSS_NEWTON = 1, G2OHIP_DOGLEG = 2 };
int g2ohip_set_algorithm(g2ohip_ctx* ctx, int algorithm);
/* SparseOptimizer::setForceStopFlag (sparse_optimizer.h:179-183): the flag is
 * read between LM trials; NULL clears it. */
int g2ohip_set_force_stop_flag(g2ohip_ctx* ctx, const volatile int* flag);

/* SparseOptimizer::computeMarginals / BlockSolver::computeMarginals
 * (block_solver.hpp:479-490, g2o.cpp:620-649): the requested blocks
 * (rows[k], cols[k]) of Hpp^-1, in free-vertex (hessianIndex) numbering, into
 * blocks[nblocks*36] (row-major 6x6), from the system the last build_system
 * or optimize linearised (undamped). Bundle adjustment: Hpp is the camera
 * block of the Hessian, block diagonal for projection edges, so a diagonal
 * block is the inverse of that camera's Hpp block (exact, FP64) and every
 * other block is zero; this is the reference's result, not the covariance of
 * the reduced camera system. Pose graphs: Hpp is the whole Hessian, and each
 * requested block column is found by six FP64 block-Jacobi PCG solves
 * stopped at r.M^-1 r <= tol * r0.M^-1 r0 (tol <= 0: 1e-24), where the
 * reference factorises (marginal_covariance_cholesky.h:42-80). The
 * reference's LinearSolverPCG itself provides no marginals
 * (linear_solver.h:82-90 returns false). */
int g2ohip_compute_marginals(g2ohip_ctx* ctx, int nblocks, const int* rows, const int* cols, double* blocks,
                             double tol);

/* ------------------------------------------------- inspection (FP64 copies) */
/* Hpp[Cf*36], Hll[Pf*9], Hpl per edge [E*18] (6x3 row-major; zero for edges
 * without a block), all row-major, of the last build_system, or of the last
 * exact-Hessian evaluation (g2ohip_ba_adjoint / g2ohip_ba_newton) if that came
 * later. NULL skips. */
int g2ohip_ba_get_hessian(g2ohip_ctx* ctx, double* Hpp, double* Hll, double* Hpl_edge);
/* Hschur upper blocks of the last solve, including the damping: rows/cols
 * [nS], blocks [nS*36] row-major; bschur [6Cf]. */
int g2ohip_ba_get_schur(g2ohip_ctx* ctx, int* rows, int* cols, double* blocks, double* bschur);
/* statistics of the last g2ohip_solve: info[0] PCG iterations (summed over
 * refinement), info[1] refinement corrections, info[2] final PCG residual
 * memory (LinearSolverPCG::_residual), info[3] 1 if the solve fell back to
 * FP64 (see g2ohip_options.precision) */
int g2ohip_last_solve_info(g2ohip_ctx* ctx, double* info);
/* The LM trials of the last g2ohip_optimize with the Levenberg algorithm, one
 * record of G2OHIP_TRIAL_FIELDS doubles each (the do-loop of
 * optimization_algorithm_levenberg.cpp:97-144): lambda used, currentChi,
 * tempChi, the divisor of rho (computeScale() + 1e-3, or 1 after a failed
 * solve), rho, accepted (1/0), PCG iterations, refinement corrections, FP64
 * fallback (1/0), PCG residual memory after the solve. Copies at most max
 * records into out (may be NULL) and stores the count in *count. */
#define G2OHIP_TRIAL_FIELDS 10
int g2ohip_lm_trials(g2ohip_ctx* ctx, double* out, int max, int* count);

/* ------------------------------------------------------- kernel timing */
/* Per-kernel durations from HIP events recorded on the context's stream around
 * each launch (the device-side counterpart of the G2OBatchStatistics timers,
 * g2o/core/batch_stats.h:49-67). g2ohip_set_profiling(on) clears the counters.
 * G2OHIP_K_PCG covers one whole PCG solve (all its kernels). */
enum g2ohip_kernel {
  G2OHIP_K_LINEARIZE = 0,
  G2OHIP_K_SCHUR = 1,
  G2OHIP_K_PCG = 2,
  G2OHIP_K_REFINE = 3,
  G2OHIP_K_CHI2 = 4,
  G2OHIP_K_UPDATE = 5,
  G2OHIP_K_SCHUR_WINDOWS = 6, /* the off-diagonal Schur window kernel alone (within G2OHIP_K_SCHUR) */
  G2OHIP_K_PCG_SPMV = 7,      /* the explicit-operator iteration launch (one-launch iteration) or its
                                 direction + SpMV kernel alone (two launches), within G2OHIP_K_PCG */
  G2OHIP_K_COUNT = 8
};
int g2ohip_set_profiling(g2ohip_ctx* ctx, int on);
/* Time of the PCG iteration kernels on the operator of the last explicit
 * bundle-adjustment solve: `reps` direction + SpMV launches back to back
 * between two HIP events on the context's stream give *us_spmv (microseconds
 * per launch); `reps` whole iterations (SpMV + update launches) give
 * *us_iteration. With the one-launch iteration (the default up to 2,048
 * cameras; not with G2OHIP_PCG_FUSED=0) both are the time of one such launch. The launches
 * continue that solve's recurrences on its scratch vectors (the solution x
 * was already copied out; the next solve starts afresh). */
int g2ohip_replay_pcg(g2ohip_ctx* ctx, int reps, double* us_spmv, double* us_iteration);
int g2ohip_get_profile(g2ohip_ctx* ctx, int kernel, double* total_ms, int* launches);

/* ------------------------- LinearSolver contract (g2o/core/linear_solver.h) */
/* LinearSolverPCG::solve (linear_solver_pcg.hpp:87-161) on a host matrix:
 * symmetric, upper blocks only (brow <= bcol), uniform block size bd (3 or 6),
 * blocks[nblk*bd*bd] row-major. x is overwritten; residual_io carries the
 * LinearSolverPCG::_residual memory (pass -1 for a fresh solver).
 * *iterations receives the iteration count. FP64 throughout. */
int g2ohip_linear_solve_pcg(g2ohip_ctx* ctx, int nb, int bd, int nblk, const int* brow, const int* bcol,
                            const double* blocks, const double* b, double* x, double tol, int max_iter,
                            double* residual_io, int* iterations);

/* The same solve on a matrix already in device memory, allocating nothing. No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops). All pointers
 * are device pointers: full block rows of the symmetric matrix (both
 * triangles) F_ptr[nb+1], F_col[nnz] (sorted within a row), F_val[nnz*bd*bd]
 * row-major blocks, F_diag[nb] the position of block (i, i) in row i; b and x
 * [nb*bd]. The scratch (preconditioner, PCG vectors, partial sums) is a
 * per-GPU static array of 6 MiB held until the process ends, so
 * nb * bd <= G2OHIP_NOALLOC_MAX_N (larger systems: G2OHIP_ERR_UNSUPPORTED).
 * Calls that overlap on one GPU share that array and race: serialise them.
 * Same stop rule, residual memory and iteration count as
 * g2ohip_linear_solve_pcg; x is written on the device. */
#define G2OHIP_NOALLOC_MAX_N 65536
int g2ohip_linear_solve_pcg_device(g2ohip_ctx* ctx, int nb, int bd, const int* F_ptr, const int* F_col,
                                   const double* F_val, const int* F_diag, const double* b, double* x, double tol,
                                   int max_iter, double* residual_io, int* iterations);

/* ------------------------------------------------------ 3D pose graph */
/* SparseOptimizer::initializeOptimization + BlockSolver::buildStructure for
 * EdgeSE3 edges between VertexSE3 poses (edge_se3.cpp, vertex_se3.h), 6x6
 * blocks without Schur (block_solver.hpp:333-343). Replaces the context's
 * problem: afterwards the generic Solver calls above (compute_active_errors,
 * build_system, set_lambda, solve, get_x, get_b, update, push / pop /
 * discard_top, lambda_init, compute_scale, optimize) act on the pose graph;
 * the ba_* calls fail until a ba_set_structure. The solve is block-Jacobi
 * PCG in FP64 (g2ohip_options.pcg_*). fixed[N] 0/1 (NULL = none). */
int g2ohip_pg_set_structure(g2ohip_ctx* ctx, int num_poses, int num_edges, const int* edge_from, const int* edge_to,
                            const uint8_t* fixed);
/* info[0..3] = free poses, active edges, upper Hessian blocks, blocks in both triangles */
int g2ohip_pg_structure_info(g2ohip_ctx* ctx, long long* info);
/* EdgeSE3::setMeasurement / setInformation: meas[E*7] relative pose from -> to
 * as [tx ty tz qx qy qz qw], information[E*36] row-major 6x6 (error order
 * [t; qx qy qz] of toVectorMQT, isometry3d_mappings.cpp:80-85). */
int g2ohip_pg_set_measurements(g2ohip_ctx* ctx, const double* meas, const double* information);
/* VertexSE3 estimates poses[N*7] = Isometry3 [tx ty tz qx qy qz qw]
 * (x_world = R x + t); quaternions are normalised on entry. */
int g2ohip_pg_set_state(g2ohip_ctx* ctx, const double* poses);
int g2ohip_pg_get_state(g2ohip_ctx* ctx, double* poses);
/* Undamped Hessian upper blocks of the last build_system (diagonal blocks
 * first, then the pairs): rows/cols [info[2]], blocks [info[2]*36] row-major,
 * b [6 * free poses]. NULL skips. */
int g2ohip_pg_get_system(g2ohip_ctx* ctx, int* rows, int* cols, double* blocks, double* b);

/* ------------------------------------------------ BAL (9-DoF cameras) */
/* g2o's BAL example (g2o/examples/bal/bal_example.cpp:70-207, 227-341):
 * VertexCameraBAL (9 values: angle-axis rotation w, translation t, focal
 * length f, radial distortion k1 k2; oplusImpl adds the update, :86-89),
 * VertexPointBAL (3 values, +), EdgeObservationBAL (error = f r(p) p - z with
 * P = R(w) X + t, p = -P.xy / P.z, r = 1 + k1 |p|^2 + k2 |p|^4, :154-204),
 * solved by BlockSolver<BlockSolverTraits<9, 3>> with Schur and
 * LinearSolverPCG (:227-250) under OptimizationAlgorithmLevenberg. The
 * reference differentiates the error automatically (G2O_MAKE_AUTO_AD_FUNCTIONS,
 * g2o/core/auto_differentiation.h:155-302); the kernels use the closed form.
 * After g2ohip_bal_set_structure the context is a BAL context: the Solver
 * calls above (compute_active_errors, get_errors, build_system, set_lambda,
 * solve, get_x / get_b with x = [9 Cf camera values, 3 Pf point values],
 * update, push / pop / discard_top, lambda_init, compute_scale, optimize
 * with Levenberg-Marquardt or Gauss-Newton) act on it. FP64 with the explicit
 * Schur complement only (other options: G2OHIP_ERR_UNSUPPORTED); no Dogleg,
 * multiplyHessian or marginals. */

/* SparseOptimizer::initializeOptimization + BlockSolver<9, 3>::buildStructure
 * (block_solver.hpp:110-269): edge_cam / edge_pt [E], fixed masks [C] / [P]
 * (NULL = none fixed, as bal_example.cpp leaves them). A camera that observes
 * one point twice is G2OHIP_ERR_UNSUPPORTED. */
int g2ohip_bal_set_structure(g2ohip_ctx* ctx, int num_cams, int num_pts, int num_edges, const int* edge_cam,
                             const int* edge_pt, const uint8_t* cam_fixed, const uint8_t* pt_fixed);
/* info[0..6] = free cameras, free points, active edges, Hschur upper blocks,
 * Hschur blocks in both triangles, off-diagonal Schur pair products, Hpl
 * blocks */
int g2ohip_bal_structure_info(g2ohip_ctx* ctx, long long* info);
/* EdgeObservationBAL::setMeasurement / setInformation (bal_example.cpp:
 * 305-315): meas[E*2] pixels relative to the image centre, weight[E] > 0
 * (information weight * I2, NULL = the identity the example sets);
 * huber_delta > 0 attaches RobustKernelHuber(delta) to every edge. */
int g2ohip_bal_set_measurements(g2ohip_ctx* ctx, const double* meas, const double* weight, double huber_delta);
/* VertexCameraBAL / VertexPointBAL estimates: cams[C*9], pts[P*3] */
int g2ohip_bal_set_state(g2ohip_ctx* ctx, const double* cams, const double* pts);
int g2ohip_bal_get_state(g2ohip_ctx* ctx, double* cams, double* pts);
/* Blocks of the last build_system: Hpp[Cf*81], Hll[Pf*9], Hpl_edge[E*27]
 * (each row-major, optional; an edge without a block gets zeros). */
int g2ohip_bal_get_hessian(g2ohip_ctx* ctx, double* Hpp, double* Hll, double* Hpl_edge);
/* The reduced camera system of the last solve (damped by its lambda): upper
 * blocks rows[], cols[], blocks[81 each] in (row, col) order, count info[3]
 * of g2ohip_bal_structure_info, and bschur[9 Cf]. */
int g2ohip_bal_get_schur(g2ohip_ctx* ctx, int* rows, int* cols, double* blocks, double* bschur);

/* ------------------------------------------------------- g2o file format */
/* Host-side graph container for the path's vertex and edge types, read from
 * and written to g2o's text format (OptimizableGraph::load / save,
 * optimizable_graph.cpp:344-567, 736-780). One line per element: a tag, the
 * vertex id or the edge's vertex ids, then the values:
 *   VERTEX_SE3:EXPMAP id  tx ty tz qx qy qz qw   camera-to-world pose (the file
 *                         holds the inverse of the world-to-camera estimate,
 *                         vertex_se3_expmap.cpp:35-44)
 *   VERTEX_TRACKXYZ id  x y z                    (vertex_pointxyz.cpp:40-46)
 *   EDGE_SE3_PROJECT_XYZ:EXPMAP point camera  u v  O00 O01 O11
 *                         (edge_project_xyz.cpp:34-42; the information matrix
 *                         as its upper triangle, base_edge.h:165-180)
 *   VERTEX_SE3:QUAT id  tx ty tz qx qy qz qw    (vertex_se3.cpp:46-55)
 *   EDGE_SE3:QUAT from to  tx ty tz qx qy qz qw  21 upper-triangle values
 *                         (edge_se3.cpp:43-58; the quaternion is normalised)
 *   PARAMS_CAMERAPARAMETERS id  f cx cy baseline
 *                         (parameter_cameraparameters.cpp:43-55)
 *   EDGE_PROJECT_XYZ2UV:EXPMAP point camera param  u v  O00 O01 O11
 *                         (edge_project_xyz2uv.cpp:37-47: the projection
 *                         of EDGE_SE3_PROJECT_XYZ with fx = fy = f and (cx, cy)
 *                         from the parameter, which must precede the edge)
 *   FIX id ...            fixes vertices defined earlier in the file
 * Empty lines and lines whose first token starts with '#' are skipped. As in
 * the reference, a vertex with a negative or already used id, an edge whose
 * vertices are not defined yet or repeat, and an unknown tag are reported in
 * the graph's log and skipped. Stricter than the reference: a line with too
 * few or malformed values, an edge joining vertices of the wrong types or
 * naming an unassigned (-1) vertex is skipped too (g2o keeps such elements
 * partly read). EdgeSE3ProjectXYZ's intrinsics (fx fy cx cy) are not part of
 * the format; the caller supplies them. */
typedef struct g2ohip_graph g2ohip_graph;

int g2ohip_graph_create(g2ohip_graph** out);
void g2ohip_graph_destroy(g2ohip_graph* g);
/* message of the graph's last failed call */
const char* g2ohip_graph_last_error(const g2ohip_graph* g);
/* reports of skipped lines and elements, one per line, as the reference logs
 * them (G2O_ERROR / G2O_WARN in optimizable_graph.cpp:344-521) */
const char* g2ohip_graph_log(const g2ohip_graph* g);

/* OptimizableGraph::load(const char*) (optimizable_graph.cpp:344-530): appends
 * the file's elements to the graph. The text is parsed by nthreads host
 * threads (<= 0: hardware concurrency, at most 16) and applied in line
 * order, so the result does not depend on nthreads. */
int g2ohip_graph_load(g2ohip_graph* g, const char* filename, int nthreads);

/* info[0..11] = cameras (VERTEX_SE3:EXPMAP), points (VERTEX_TRACKXYZ),
 * projection edges (both tags), poses (VERTEX_SE3:QUAT), pose edges, fixed
 * vertices, lines read, lines with an unknown tag, skipped elements, BAL
 * observations, EDGE_PROJECT_XYZ2UV edges, camera parameters */
int g2ohip_graph_info(const g2ohip_graph* g, long long* info);

/* Bundle-adjustment part in the layout g2ohip_ba_set_structure /
 * set_measurements / set_state take: cameras and points each sorted by vertex
 * id (the order of SparseOptimizer's active vertices,
 * sparse_optimizer.cpp:175-200), edges in file order with indices into those
 * arrays. cams[C*7] world-to-camera SE3Quat [t q]; omega[E*3] = O00 O01 O11.
 * NULL skips an output. */
int g2ohip_graph_get_ba(const g2ohip_graph* g, int* cam_id, double* cams, uint8_t* cam_fixed, int* pt_id,
                        double* pts, uint8_t* pt_fixed, int* edge_cam, int* edge_pt, double* meas, double* omega);
/* Camera parameters sorted by id: param_id[np], params[np*4] = f cx cy
 * baseline; edge_param[E] = each projection edge's parameter (index into
 * that order) or -1 for EDGE_SE3_PROJECT_XYZ. NULL skips. */
int g2ohip_graph_get_ba_params(const g2ohip_graph* g, int* param_id, double* params, int* edge_param);
/* Pose-graph part: poses sorted by id, poses[N*7] Isometry3 [t q], edges in
 * file order, meas[E*7], info[E*36] full symmetric. NULL skips. */
int g2ohip_graph_get_pg(const g2ohip_graph* g, int* pose_id, double* poses, uint8_t* fixed, int* edge_from,
                        int* edge_to, double* meas, double* info);
/* Replace the estimates (same order as the getters), as
 * SparseOptimizer::optimize leaves them in the vertices before save. */
int g2ohip_graph_set_ba_state(g2ohip_graph* g, const double* cams, const double* pts);
int g2ohip_graph_set_pg_state(g2ohip_graph* g, const double* poses);

/* Add elements from arrays (OptimizableGraph::addVertex / addEdge,
 * optimizable_graph.cpp:179-246), e.g. to write a problem built in memory.
 * cam_id / pt_id NULL: ids C0.., then points after the cameras, where C0 is
 * one past the graph's largest id (0 for an empty graph); omega NULL:
 * identity. A duplicate id fails the call and adds nothing. */
int g2ohip_graph_add_ba(g2ohip_graph* g, int num_cams, int num_pts, int num_edges, const int* cam_id,
                        const double* cams, const uint8_t* cam_fixed, const int* pt_id, const double* pts,
                        const uint8_t* pt_fixed, const int* edge_cam, const int* edge_pt, const double* meas,
                        const double* omega);
/* pose_id NULL: ids C0..; info NULL: identity */
int g2ohip_graph_add_pg(g2ohip_graph* g, int num_poses, int num_edges, const int* pose_id, const double* poses,
                        const uint8_t* fixed, const int* edge_from, const int* edge_to, const double* meas,
                        const double* info);

/* OptimizableGraph::save(const char*, level 0) (optimizable_graph.cpp:531-567):
 * the vertices of the edges, sorted by id, each followed by "FIX id" if
 * fixed, then the edges in insertion order; each value printed as an ostream
 * prints a double at the given precision (1..17; g2o's streams use 6). */
int g2ohip_graph_save(const g2ohip_graph* g, const char* filename, int precision);

/* BAL text format as bal_example.cpp:265-335 parses it: "C P E", E lines
 * "camera point x y", C cameras of 9 values (angle-axis rotation, translation,
 * f, k1, k2), P points of 3 values. Replaces the graph's BAL part. */
int g2ohip_graph_load_bal(g2ohip_graph* g, const char* filename, int nthreads);
/* info: [0] cameras, [1] points, [2] observations; NULL skips an output */
int g2ohip_graph_get_bal(const g2ohip_graph* g, long long* info, double* cams9, double* pts, int* edge_cam,
                         int* edge_pt, double* meas);

/* ---------------------------------------------- adjoint (differentiation) */
/* Gradient of a scalar loss L(x*) at the converged state with respect to the
 * measurements and the per-edge information weights, by implicit
 * differentiation of the optimality condition G(x*, z, w) = grad(F/2) = 0,
 * F = sum_e rho(w_e |e_e|^2):
 *   H lambda = dL/dx,  dL/dz = -(dG/dz)^T lambda,  dL/dw = -(dG/dw)^T lambda,
 * H = dG/dx the exact Hessian (Gauss-Newton term plus the second-order terms
 * of the projection, of the exp-map update and of the Huber kernel), solved
 * once by the explicit Schur complement and PCG in FP64 to the relative
 * tolerance tol (<0 defaults to 1e-24 on r.M^-1 r; max_iter < 0: system size).
 * dL_dcam[C*6] is the loss gradient with respect to the left-multiplied camera
 * update [omega; upsilon] (rows of fixed cameras ignored), dL_dpt[P*3] (NULL =
 * zero). Outputs dL_dmeas[E*2] and dL_dweight[E] (NULL skips; zero for
 * inactive edges). Requires a stationary state (see g2ohip_ba_newton).
 * No counterpart in the reference, which stops at the solution. */
int g2ohip_ba_adjoint(g2ohip_ctx* ctx, const double* dL_dcam, const double* dL_dpt, double* dL_dmeas,
                      double* dL_dweight, double tol, int max_iter, int* iterations);
/* Newton steps x <- x + (H + mu I)^-1 (-grad(F/2)) with the exact Hessian
 * above (FP64, PCG tolerance tol as for g2ohip_ba_adjoint), damped as
 * Levenberg-Marquardt on that Hessian (mu = 0 while the robust chi2 does not
 * increase; a PCG that meets non-positive curvature counts as a failed
 * factorisation and raises mu) with a trust region: a step whose largest
 * component exceeds max_step (> 0; <= 0: unbounded) is rejected. Drives an LM
 * solution, whose accuracy LM's chi2-based step acceptance limits, to the
 * stationary point that g2ohip_ba_adjoint differentiates. *ok = 0 if a step
 * found no acceptable damping. No counterpart in the reference. */
int g2ohip_ba_newton(g2ohip_ctx* ctx, int steps, double tol, double max_step, int* ok);

#ifdef __cplusplus
}
#endif

#endif /* G2O_HIP_H */
