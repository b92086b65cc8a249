/*
 * sophus_gpu.h — C ABI of the MI355X (gfx950) Schur-complement Levenberg–Marquardt
 * NLLS library.  Every entry point takes plain pointers and sizes; no C++ or torch
 * types cross this boundary.
 *
 * The surface mirrors sophus_opt's NLLS API:
 *   optimize_nlls            crates/sophus_opt/src/nlls.rs:73-80
 *   Optimizer::new/step      crates/sophus_opt/src/nlls/optimizer.rs:71-120,345-377
 *   Optimizer::build_solution crates/sophus_opt/src/nlls/optimizer.rs:609-645
 *   OptParams (+ defaults)   crates/sophus_opt/src/nlls/optimizer_types.rs:161-240
 *   VarKind                  crates/sophus_opt/src/variables.rs:17-24
 *   VarFamily::new_with_const_ids crates/sophus_opt/src/variables/var_family.rs:26-46
 *   CostFn / CostTerms       crates/sophus_opt/src/nlls/cost/cost_fn.rs:105-151,
 *                            crates/sophus_opt/src/nlls/cost/cost_term.rs:74-105
 *   RobustKernel::Huber      crates/sophus_opt/src/robust_kernel.rs:28-57
 *   TerminationReason        crates/sophus_opt/src/nlls/optimizer_types.rs:93-122
 *   NllsError                crates/sophus_opt/src/nlls/optimizer_types.rs:422-504
 *
 * Conventions (identical to the reference):
 *   - SE(3) members are 7 doubles (qw, qx, qy, qz, tx, ty, tz) = world_from_camera;
 *     tangent order [omega; nu] (rotation first); update T <- exp(delta) * T.
 *   - Vec3 members are 3 doubles; pinhole members are 4 doubles (fx, fy, cx, cy); EUCM
 *     members are 6 doubles (fx, fy, cx, cy, alpha, beta).
 *   - Family partitions: Free families in lexicographic name order, then
 *     Marginalized families; Conditioned families and const members get no blocks.
 *   - The cost has no 1/2 factor: cost = sum rho(|r|) (Huber) or |r|^2.
 *
 * Threading: an sgpu_optimizer handle owns its device buffers and one HIP stream;
 * calls on one handle must not overlap; distinct handles are independent and may be
 * used from different threads at once; sgpu_last_error() is per calling thread.
 */
#ifndef SOPHUS_GPU_H
#define SOPHUS_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SGPU_ABI_VERSION 11

/* Status codes.  One code per NllsError variant the NLLS path can raise
 * (crates/sophus_opt/src/nlls/optimizer_types.rs:422-504) plus device errors. */
typedef enum sgpu_status {
  SGPU_OK = 0,
  SGPU_ERR_INVALID_ARGUMENT = 1,
  /* NllsError::SchurMissingFreeOrMarg  (optimizer.rs:680-697) */
  SGPU_ERR_SCHUR_MISSING_FREE_OR_MARG = 2,
  /* NllsError::SchurCostConnectsTwoMargFamilies (optimizer.rs:699-714) */
  SGPU_ERR_SCHUR_COST_CONNECTS_TWO_MARG_FAMILIES = 3,
  /* VarFamilyError::UnknownFamily (var_families.rs) */
  SGPU_ERR_UNKNOWN_FAMILY = 4,
  /* VarFamilyError::TypeError */
  SGPU_ERR_FAMILY_TYPE = 5,
  /* an entity index is >= the family's member count */
  SGPU_ERR_INDEX_OUT_OF_RANGE = 6,
  /* a layout this library does not implement (see sgpu_last_error) */
  SGPU_ERR_UNSUPPORTED = 7,
  SGPU_ERR_HIP = 8,
  SGPU_ERR_NO_DEVICE = 9,
  SGPU_ERR_OUT_OF_MEMORY = 10,
  /* a non-finite residual or pivot (LdltDecompositionError::NonFinitePivot) */
  SGPU_ERR_NON_FINITE = 11,
  /* the handle has already terminated */
  SGPU_ERR_TERMINATED = 12,
  /* NllsError::SchurConstraintOnMarginalizedFamily (optimizer.rs:664-678) */
  SGPU_ERR_SCHUR_CONSTRAINT_ON_MARGINALIZED_FAMILY = 13
} sgpu_status;

/* VarKind (crates/sophus_opt/src/variables.rs:17-24) */
typedef enum sgpu_var_kind {
  SGPU_VAR_FREE = 0,
  SGPU_VAR_CONDITIONED = 1,
  SGPU_VAR_MARGINALIZED = 2
} sgpu_var_kind;

/* Variable types on the Schur BA path. */
typedef enum sgpu_var_type {
  SGPU_VAR_ISOMETRY3 = 0, /* Isometry3F64: 7 params, 6 dof */
  SGPU_VAR_VEC3 = 1,      /* VecF64<3>:   3 params, 3 dof */
  SGPU_VAR_PINHOLE = 2,   /* PinholeCameraF64: 4 params, 4 dof (additive) */
  /* EnhancedUnifiedCameraF64: 6 params (fx, fy, cx, cy, alpha, beta), 6 dof (additive;
   * perspective_camera.rs:76-89, distortions/enhanced_unified.rs) */
  SGPU_VAR_EUCM = 3
} sgpu_var_type;

/* Cost term kinds (closed set instead of HasResidualFn trait objects). */
typedef enum sgpu_cost_kind {
  /* PinholeCameraReprojectionCostTerm: args (intrinsics, world_from_camera, point);
   * measurement uv[2].  pinhole_reprojection.rs:42-95 */
  SGPU_COST_PINHOLE_REPROJECTION = 0,
  /* PoseGraph3CostTerm: args (world_from_pose_m, world_from_pose_n);
   * measurement pose_m_from_pose_n[7].  pose_graph3.rs:33-79 */
  SGPU_COST_POSE_GRAPH3 = 1,
  /* EucmCameraReprojectionCostTerm: args (intrinsics, world_from_camera, point);
   * measurement uv[2].  example_problems/ba_problem.rs:45-110.  The intrinsics family may
   * be Conditioned, or Free with one active member (one calibration shared by the
   * observations: its block couples to every camera, so it is ordered last in the RCS) */
  SGPU_COST_EUCM_REPROJECTION = 2,
  /* Isometry3PriorCostTerm: arg (world_from_pose) of the pose family; measurement the prior
   * mean [7] then the precision W [36, row-major]: r = log(T E^-1), cost r~^T W r~ with
   * r~ = w(|r|) r (functor_library/costs/isometry3_prior.rs:46-88, make with a precision,
   * cost/evaluated_term.rs:286-327) */
  SGPU_COST_ISOMETRY3_PRIOR = 3
} sgpu_cost_kind;

/* Equality-constraint kinds (closed set instead of HasEqConstraintResidualFn objects). */
typedef enum sgpu_eq_kind {
  /* TranslationNormConstraint: c(T) = |t(T)| - r on one SE(3) argument, residual dim 1,
   * one parameter r per term.  example_problems/ba_scale_constraint.rs:44-99 */
  SGPU_EQ_TRANSLATION_NORM = 0
} sgpu_eq_kind;

typedef enum sgpu_robust_kernel {
  SGPU_ROBUST_NONE = 0,
  SGPU_ROBUST_HUBER = 1 /* RobustKernel::Huber(delta), robust_kernel.rs:28-35 */
} sgpu_robust_kernel;

/* LinearSolverEnum subset (crates/sophus_solver/src/lib.rs:295-318) plus the
 * iterative RCS solvers of this library. */
typedef enum sgpu_solver {
  /* SchurBlockSparseLdlt: Schur complement + block-sparse LDLt of the RCS */
  SGPU_SOLVER_SCHUR_BLOCK_SPARSE_LDLT = 0,
  /* Schur complement + block-Jacobi preconditioned CG on the RCS */
  SGPU_SOLVER_SCHUR_PCG_JACOBI = 1,
  /* Schur complement + PCG with the learned (MLP-corrected) block-Jacobi */
  SGPU_SOLVER_SCHUR_PCG_LEARNED = 2,
  /* SchurSparseLdlt: Schur complement + scalar sparse LDLt of the RCS (SparseLdlt,
   * sparse_ldlt.rs:571-650): the same factorization, in the same block order, with the
   * scalar pivot rule |d_jj| <= max(|d_jj|, 1) * 1e-12 -> 0 (sparse_ldlt.rs:399-445) */
  SGPU_SOLVER_SCHUR_SPARSE_LDLT = 3
} sgpu_solver;

typedef enum sgpu_precision {
  /* FP64 everywhere (parity path) */
  SGPU_PRECISION_F64 = 0,
  /* FP32 residuals/Jacobians/assembly/Schur products; FP64 RCS, pivots, cost
   * sums, LM control */
  SGPU_PRECISION_MIXED = 1
} sgpu_precision;

/* TerminationReason (optimizer_types.rs:93-122) */
typedef enum sgpu_termination {
  SGPU_TERM_NONE = 0,
  SGPU_TERM_MAX_ITERATIONS = 1,
  SGPU_TERM_FUNCTION_TOLERANCE = 2,
  SGPU_TERM_GRADIENT_TOLERANCE = 3,
  SGPU_TERM_PARAMETER_TOLERANCE = 4,
  SGPU_TERM_MAX_CONSECUTIVE_REJECTS = 5
} sgpu_termination;

/* VarFamily (var_family.rs:15-46).  `members` is num_members x params AoS. */
typedef struct sgpu_var_family {
  const char* name;
  int32_t type; /* sgpu_var_type */
  int32_t kind; /* sgpu_var_kind */
  int64_t num_members;
  const double* members;
  int64_t num_const_ids;
  const int64_t* const_ids; /* members held constant (VarFamily::new_with_const_ids) */
} sgpu_var_family;

/* CostFn::new_boxed(global_constants, CostTerms::new(family_names, terms), kernel)
 * (cost_fn.rs:105-151).  entity_indices: num_terms x arity; measurements:
 * num_terms x (2 | 7). */
typedef struct sgpu_cost_fn {
  int32_t kind;          /* sgpu_cost_kind */
  int32_t robust_kernel; /* sgpu_robust_kernel */
  double huber_delta;
  const char* family_names[3]; /* arity 3 (reprojection) or 2 (pose graph) */
  int64_t num_terms;
  const int64_t* entity_indices;
  const double* measurements;
} sgpu_cost_fn;

/* EqConstraintFn::new_boxed((), EqConstraints::new(family_names, terms))
 * (crates/sophus_opt/src/nlls/constraint/eq_constraint_fn.rs:61-140).  All terms of one set
 * share one constraint block and one Lagrange multiplier block, as in the reference's
 * EqSystem (one partition of one block per set, eq_system.rs:33-37, 207-262): their
 * residuals and Jacobians add.  entity_indices: num_terms x arity; params: num_terms x
 * (1: the radius r). */
typedef struct sgpu_eq_fn {
  int32_t kind; /* sgpu_eq_kind */
  int32_t reserved0;
  const char* family_names[1];
  int64_t num_terms;
  const int64_t* entity_indices;
  const double* params;
} sgpu_eq_fn;

/* OptParams (optimizer_types.rs:161-189) + build-only fields. */
typedef struct sgpu_opt_params {
  int64_t num_iterations;        /* default 100 */
  double initial_lm_damping;     /* default 10.0 */
  int32_t parallelize;           /* default 1 (accepted, no effect on the GPU) */
  int32_t solver;                /* sgpu_solver, default SCHUR_BLOCK_SPARSE_LDLT */
  int32_t skip_final_hessian;    /* default 0 */
  int32_t precision;             /* sgpu_precision, default F64 */
  double function_tolerance;     /* default 1e-10 */
  double gradient_tolerance;     /* default 1e-14 */
  double parameter_tolerance;    /* default 1e-12 */
  int64_t max_consecutive_rejects; /* default 20 */
  double pcg_relative_tolerance; /* default 1e-8 (PCG solvers only) */
  int64_t pcg_max_iterations;    /* default 500 */
  int32_t device;                /* HIP device ordinal, default 0 */
  /* 0 (default): the term sort, the S pattern and the pair lists are built on the device;
   * 1: by the host analysis (the same arrays) */
  int32_t host_symbolic;
  /* SCHUR_PCG_LEARNED only: SGPU_PRECOND_WEIGHTS FP32 weights of the block-Jacobi
   * correction (copied at create; see sgpu_precond_train).  Default NULL. */
  const float* precond_weights;
  /* PCG solvers: forcing-sequence CG tolerance (an inexact-Newton choice; no reference
   * counterpart).  0 (default): every LM step solves to pcg_relative_tolerance.  eta > 0:
   * the step's relative CG tolerance is max(pcg_relative_tolerance, eta * min(1,
   * sqrt(|g|_inf / |g0|_inf))), g the gradient of the last linearization whose scalars the
   * host has read (the previous step's), g0 the first one's: loose while the gradient is
   * large, tightening as it falls. */
  double pcg_forcing;
} sgpu_opt_params;

/* StepInfo (optimizer_types.rs:67-83) plus per-step diagnostics. */
typedef struct sgpu_step_info {
  double merit;
  double smooth_cost;
  int64_t inner_step;
  double lm_damping;     /* damping after the step */
  int32_t termination;   /* sgpu_termination, NONE while running */
  int32_t accepted;      /* 1 if the step was accepted */
  double new_merit;      /* merit of the proposed variables (NaN if not computed) */
  double gain_ratio;     /* NaN if not computed */
  double max_abs_neg_grad;
  double max_abs_delta;
  int64_t pcg_iterations; /* 0 for the direct solver */
  /* per-phase device milliseconds (HIP events on the handle's stream) */
  double ms_linearize;   /* residual/Jacobian + assembly + Schur elimination */
  double ms_factor;      /* RCS factorization or PCG */
  double ms_solve;       /* triangular solves + back-substitution + update */
  double ms_cost;        /* merit of the proposed variables */
  double ms_total;       /* wall time of the step on the host */
} sgpu_step_info;

/* OptimizationSolution (optimizer_types.rs:243-257), scalar part. */
typedef struct sgpu_solution {
  double initial_cost;
  double final_cost;
  int64_t num_steps;
  int32_t termination; /* sgpu_termination */
  int32_t reserved0;
  double final_max_abs_neg_grad; /* from the final (nu = 0) linearization */
} sgpu_solution;

typedef struct sgpu_optimizer sgpu_optimizer;

/* ---- parameters / diagnostics ------------------------------------------------ */

/* OptParams::default() (optimizer_types.rs:225-240). */
void sgpu_opt_params_default(sgpu_opt_params* params);

/* Message of the last error raised on this thread ("" if none). */
const char* sgpu_last_error(void);

int32_t sgpu_abi_version(void);

/* ---- full solve ----------------------------------------------------------------
 * optimize_nlls(variables, cost_fns, [], params) (nlls.rs:73-80).
 * out_members[i] (may be NULL) receives family i's final members
 * (num_members x params).  Solver failures inside LM are rejects, not errors
 * (optimizer_helpers.rs:176-185). */
sgpu_status sgpu_optimize_nlls(const sgpu_var_family* families, int64_t num_families,
                               const sgpu_cost_fn* cost_fns, int64_t num_cost_fns,
                               const sgpu_opt_params* params, double* const* out_members,
                               sgpu_solution* solution);

/* ---- stateful optimizer --------------------------------------------------------
 * Optimizer::new (optimizer.rs:71-120): validates, sorts terms, uploads, computes
 * the initial merit and the symbolic structure. */
sgpu_status sgpu_optimizer_create(const sgpu_var_family* families, int64_t num_families,
                                  const sgpu_cost_fn* cost_fns, int64_t num_cost_fns,
                                  const sgpu_opt_params* params, sgpu_optimizer** out);

/* optimize_nlls(variables, cost_fns, eq_fns, params) (nlls.rs:73-80) and Optimizer::new with
 * equality constraints (optimizer.rs:71-120): the constrained LM step of step_nlls_inner
 * (optimizer_step.rs:366-443) with the range-space KKT solve of SchurFactor
 * (schur_ldlt.rs:450-480, 674-730, 762-772): G^T lambda in the free rows of the right-hand
 * side, -c in the constraint rows (eq_system.rs:162-262); Y = S^-1 G^T, M = G Y,
 * dlambda = M^-1 (G u + c), dx_f = u - Y dlambda with u = S^-1 rhs; lambda <- lambda - dlambda
 * with an accepted step (update_lambdas, eq_system.rs:149-159); the gain ratio over the whole
 * step, multiplier rows included; no gradient-tolerance test (optimizer.rs:527-535).  The
 * merit is the cost.  Constraint sets must name the pose family (Free; a Marginalized one is
 * SGPU_ERR_SCHUR_CONSTRAINT_ON_MARGINALIZED_FAMILY, optimizer.rs:664-678); the direct solver
 * in FP64 only; at most 16 sets.  A singular M rejects the step (SingularKktConstraint). */
sgpu_status sgpu_optimizer_create_eq(const sgpu_var_family* families, int64_t num_families,
                                     const sgpu_cost_fn* cost_fns, int64_t num_cost_fns,
                                     const sgpu_eq_fn* eq_fns, int64_t num_eq_fns,
                                     const sgpu_opt_params* params, sgpu_optimizer** out);
sgpu_status sgpu_optimize_nlls_eq(const sgpu_var_family* families, int64_t num_families,
                                  const sgpu_cost_fn* cost_fns, int64_t num_cost_fns,
                                  const sgpu_eq_fn* eq_fns, int64_t num_eq_fns,
                                  const sgpu_opt_params* params, double* const* out_members,
                                  sgpu_solution* solution);
/* The multipliers (EqSystem::lambda) of the current variables and the constraint residuals
 * c at them, one per non-empty set in the order given. */
sgpu_status sgpu_optimizer_eq_state(sgpu_optimizer* opt, double* lambda, double* residual,
                                    int64_t capacity);

/* Optimizer::step (optimizer.rs:345-377 + optimizer_step.rs:366-443). */
sgpu_status sgpu_optimizer_step(sgpu_optimizer* opt, sgpu_step_info* info);

/* Optimizer::variables(): copy family `name`'s members to `out` (capacity doubles). */
sgpu_status sgpu_optimizer_get_members(sgpu_optimizer* opt, const char* name, double* out,
                                       int64_t capacity);

/* Optimizer::merit() / lm_damping() */
double sgpu_optimizer_merit(const sgpu_optimizer* opt);
double sgpu_optimizer_lm_damping(const sgpu_optimizer* opt);

/* Optimizer::eval_cost_at (optimizer.rs:570): merit of the handle's cost terms at
 * caller-provided family values (same family order/shape as at create). */
sgpu_status sgpu_optimizer_eval_cost_at(sgpu_optimizer* opt, const double* const* members,
                                        double* cost);

/* Optimizer::build_solution (optimizer.rs:609-645): one more linearization with
 * nu = 0 unless skip_final_hessian. */
sgpu_status sgpu_optimizer_build_solution(sgpu_optimizer* opt, sgpu_solution* solution);

/* Block (row, col) of the inverse of the damped Hessian of the handle's final
 * linearization, by family name and member index: SchurFactor::inverse_block
 * (crates/sophus_solver/src/ldlt/schur_ldlt.rs:814-958) as DampedHessian::
 * inverse_block_with_damping reaches it (crates/sophus_opt/src/nlls/damped_hessian.rs:98).
 * The matrix is H + nu N at the current variables with build_solution's nu (0, or the
 * current nu with skip_final_hessian); it is linearized and factored on the device at the
 * first query after a step and reused by later queries.  free x free: S^+ (the min-norm
 * pseudo-inverse where S is singular, e.g. along a gauge direction); free x marg:
 * -(S^+ H_fm H_mm^-1); marg x free: its transpose; marg x marg:
 * H_mm^-1 delta + H_mm^-1 H_mf S^+ H_fm H_mm^-1.  out receives the block row-major
 * (6 or 3 rows by 6 or 3 columns).  Conditioned families and constant members have no block
 * (SGPU_ERR_INVALID_ARGUMENT; the reference returns None).  Direct solver only. */
sgpu_status sgpu_optimizer_inverse_block(sgpu_optimizer* opt, const char* row_family,
                                         int64_t row_member, const char* col_family,
                                         int64_t col_member, double* out, int64_t capacity);

/* OptimizationSolution::covariance_block (crates/sophus_opt/src/nlls/optimizer_types.rs:
 * 335-373): a block of the min-norm inverse of the undamped final Hessian H.  Where H has
 * full rank this is H^-1, which the Schur formulas of sgpu_optimizer_inverse_block give
 * exactly (at nu = 0).  A rank-deficient H (an unfixed gauge: a non-positive pivot of S, or
 * a pivot below 1e-9 of S's diagonal entry in its column) returns SGPU_ERR_UNSUPPORTED:
 * the reference then takes the min-norm inverse of the whole of H, which differs from the
 * Schur-complement blocks.  Requires skip_final_hessian = 0. */
sgpu_status sgpu_optimizer_covariance_block(sgpu_optimizer* opt, const char* row_family,
                                            int64_t row_member, const char* col_family,
                                            int64_t col_member, double* out, int64_t capacity);

/* Structure of the reduced camera system of this problem. */
typedef struct sgpu_rcs_stats {
  int64_t num_free_blocks;   /* RCS block rows (6x6 blocks) */
  int64_t num_marg_blocks;   /* eliminated point blocks */
  int64_t num_observations;  /* reduced reprojection terms */
  int64_t num_pose_terms;    /* reduced pose-graph terms */
  int64_t nnz_s_blocks;      /* lower-triangular S blocks incl. diagonal */
  int64_t nnz_l_blocks;      /* strictly-lower L blocks after ordering */
  int64_t num_schur_pairs;   /* (point, cam_i, cam_j<=cam_i) contributions */
  int64_t num_levels;        /* levels of the supernodal elimination tree */
  double factor_flops;       /* FP64 flops of one RCS factorization */
  int64_t num_supernodes;    /* supernodes (dense fronts) of the factorization */
  int64_t max_front_blocks;  /* largest front, in 6x6 blocks */
  int64_t front_bytes;       /* device bytes of all fronts */
  double mf_step_flops;      /* algorithmic flops of the trailing updates of all rounds of panel
                                steps (k_mf_step launches, one panel step of every ready
                                supernode each) */
  double mf_step_bytes;      /* their algorithmic HBM bytes: read + write of the lower trailing
                                fronts, one read of the panel columns below, L written */
  int64_t mf_step_launches;  /* rounds of panel steps per factorization */
} sgpu_rcs_stats;

sgpu_status sgpu_optimizer_rcs_stats(const sgpu_optimizer* opt, sgpu_rcs_stats* stats);

/* Host-only symbolic analysis of a problem (validation, layout, RCS ordering and
 * supernodes) without touching a device; same errors as sgpu_optimizer_create. */
sgpu_status sgpu_analyze(const sgpu_var_family* families, int64_t num_families,
                         const sgpu_cost_fn* cost_fns, int64_t num_cost_fns,
                         sgpu_rcs_stats* stats);

/* Elimination order of the reduced camera system (no reference counterpart as an entry
 * point; it replaces the AMD permutation that BlockSparseLdlt computes internally,
 * crates/sophus_solver/src/ldlt.rs:49-67): perm[new] = old over the free camera blocks,
 * numbered in the variables' layout order.  Host-only, like sgpu_analyze.  *num_blocks
 * receives the block count; perm (may be NULL) needs that capacity. */
sgpu_status sgpu_analyze_rcs_order(const sgpu_var_family* families, int64_t num_families,
                                   const sgpu_cost_fn* cost_fns, int64_t num_cost_fns,
                                   int64_t* perm, int64_t capacity, int64_t* num_blocks);
/* The symbolic analysis both ways on `device`: the term sort, the S pattern and the pair
 * lists by the host (what sgpu_analyze runs) and on the device (what an optimizer runs
 * unless host_symbolic is set), compared element by element over every array the numeric
 * phase reads (the sorted terms, the pattern, the pair lists, the camera-major lists, the
 * ordering and the supernodal structure).  No reference counterpart: a self-check of the
 * device pipeline.  out[0] = mismatched elements, out[1] = Schur pairs, out[2] = S blocks;
 * ms[0] / ms[1] = host / device wall time of the analysis (NULL to skip). */
sgpu_status sgpu_symbolic_check(const sgpu_var_family* families, int64_t num_families,
                                const sgpu_cost_fn* cost_fns, int64_t num_cost_fns,
                                int32_t device, int64_t out[3], double ms[2]);
/* The same order for an optimizer handle. */
sgpu_status sgpu_optimizer_rcs_order(const sgpu_optimizer* opt, int64_t* perm, int64_t capacity,
                                     int64_t* num_blocks);
/* Pivots d of the last direct RCS factorization (BlockDiagonalLdltSystem's D_j =
 * L_jj diag(d) L_jj^T, block_diag_ldlt.rs:80-117): 6 per block in elimination order,
 * zero where the PSD rule dropped a pivot (dense_ldlt.rs:164-256).  d needs 6 x blocks. */
sgpu_status sgpu_optimizer_rcs_pivots(sgpu_optimizer* opt, double* d, int64_t capacity);

/* Per-kernel device timing recorded with HIP events on the handle's stream since the
 * last reset: names[i] (static strings), count[i] launches, total_ms[i].
 * Returns the number of kernels written (<= capacity). */
int32_t sgpu_optimizer_kernel_times(const sgpu_optimizer* opt, const char** names,
                                    int64_t* counts, double* total_ms, int32_t capacity);
void sgpu_optimizer_reset_kernel_times(sgpu_optimizer* opt);
/* Event timing on the handle's stream: 1 = one event pair around each phase kernel
 * (point_prep, camera_rows, rcs_factor, rcs_solve, backsub, pose_update, cost);
 * 2 = also one pair around every panel-step launch of the direct factorization
 * ("mf_step"); 0 disables. */
void sgpu_optimizer_set_kernel_timing(sgpu_optimizer* opt, int32_t enable);

void sgpu_optimizer_destroy(sgpu_optimizer* opt);

/* ---- stage-level entry points (parity tests of one stage at a time) -----------
 * All take host arrays and copy them to/from the device internally. */

/* Kernel (a): per-term residual, Jacobians and the term's Hessian/gradient blocks
 * for pinhole reprojection terms, evaluated on the GPU in FP64 (precision = F64) or
 * FP32 (MIXED).  Mirrors PinholeCameraReprojectionCostTerm::eval + make
 * (pinhole_reprojection.rs:60-95, evaluated_term.rs:286-327).
 * Per term t: residual_out[2t..], jac_pose_out[12t..] (2x6 row-major),
 * jac_point_out[6t..] (2x3 row-major), cost_out[t] (rho(|r|)), weight_out[t].
 * Any output pointer may be NULL. */
sgpu_status sgpu_eval_pinhole_terms(int32_t precision, int64_t num_terms,
                                    const double* intrinsics /*4 per term*/,
                                    const double* world_from_camera /*7 per term*/,
                                    const double* point_in_world /*3 per term*/,
                                    const double* uv /*2 per term*/, int32_t robust_kernel,
                                    double huber_delta, double* residual_out,
                                    double* jac_pose_out, double* jac_point_out,
                                    double* cost_out, double* weight_out);

/* Kernel (a) for EUCM reprojection terms, FP64: EucmCameraReprojectionCostTerm::eval + make
 * (example_problems/ba_problem.rs:68-110, evaluated_term.rs:286-327).  Per term t:
 * residual_out[2t..], jac_intrinsics_out[12t..] (2x6 row-major, d/d(fx, fy, cx, cy, alpha,
 * beta)), jac_pose_out[12t..], jac_point_out[6t..], cost_out[t], weight_out[t].  Any output
 * pointer may be NULL. */
sgpu_status sgpu_eval_eucm_terms(int64_t num_terms, const double* intrinsics /*6 per term*/,
                                 const double* world_from_camera /*7 per term*/,
                                 const double* point_in_world /*3 per term*/,
                                 const double* uv /*2 per term*/, int32_t robust_kernel,
                                 double huber_delta, double* residual_out,
                                 double* jac_intrinsics_out, double* jac_pose_out,
                                 double* jac_point_out, double* cost_out, double* weight_out);

/* Isometry3PriorCostTerm (isometry3_prior.rs:46-88) on the GPU (FP64): per term the residual
 * log(T E^-1) (residual_out[6t..]), its Jacobian by forward-mode duals of log(exp(x) T E^-1)
 * at x = 0 (jac_out[36t..], row-major) and the cost r~^T W r~ (cost_out[t]; W 36 per term,
 * robust_kernel / huber_delta as in sgpu_cost_fn). */
sgpu_status sgpu_eval_isometry3_prior_terms(int64_t num_terms, const double* world_from_pose,
                                            const double* prior_mean, const double* precision,
                                            int32_t robust_kernel, double huber_delta,
                                            double* residual_out, double* jac_out, double* cost_out);

/* Kernel (a'): SE(3) between factor (pose_graph3.rs:33-79) on the GPU (FP64).
 * Per term: residual_out[6t..], jac_m_out[36t..], jac_n_out[36t..] (6x6 row-major). */
sgpu_status sgpu_eval_pose_graph3_terms(int64_t num_terms, const double* world_from_pose_m,
                                        const double* world_from_pose_n,
                                        const double* pose_m_from_pose_n,
                                        double* residual_out, double* jac_m_out,
                                        double* jac_n_out);

/* Kernel (d): block-sparse LDLt solve of a symmetric system with 6x6 blocks on the
 * GPU in FP64 (BlockSparseLdlt::factorize_impl + solve_inplace,
 * block_sparse_ldlt.rs:96-190,479-535).  Lower blocks in CSC-by-block-column:
 * col_ptr[nb+1], row_idx[nnz] (row >= col, diagonal included), values[nnz*36]
 * (each block 6x6 column-major).  rhs[nb*6] -> x[nb*6]. */
sgpu_status sgpu_block_ldlt_solve(int64_t num_blocks, const int64_t* col_ptr,
                                  const int64_t* row_idx, const double* values,
                                  const double* rhs, double* x);
/* The same input solved as SparseLdlt::factorize + solve_inplace (sparse_ldlt.rs:237-330,
 * 571-650): the scalar LDLt in the block order, with its per-pivot rule (a pivot with
 * |d_jj| <= max(|d_jj|, 1) * 1e-12 is zero and its column of L is empty).  The reference's
 * solve divides by a zero pivot; that component of x is NaN here. */
sgpu_status sgpu_sparse_ldlt_solve(int64_t num_blocks, const int64_t* col_ptr,
                                   const int64_t* row_idx, const double* values,
                                   const double* rhs, double* x);

/* Allocation-free variant of sgpu_block_ldlt_solve.  It has no counterpart in the
 * reference; it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * results and errors, on the current device, but it calls no hipMalloc/hipFree: every
 * device buffer (pattern, fronts, factor, work vectors) is carved from one static array of
 * SGPU_NOALLOC_SCRATCH_BYTES per GPU, which the library holds until the process ends.  A
 * system whose buffers do not fit returns SGPU_ERR_OUT_OF_MEMORY (sgpu_last_error states
 * the bytes needed and the limit) before touching the device.  Calls on one GPU share the
 * array: they are serialised within one process; calls that overlap on one GPU from
 * several processes race. */
#define SGPU_NOALLOC_SCRATCH_BYTES (64 * 1024 * 1024)
sgpu_status sgpu_block_ldlt_solve_noalloc(int64_t num_blocks, const int64_t* col_ptr,
                                          const int64_t* row_idx, const double* values,
                                          const double* rhs, double* x);

/* Kernel (d'), no reference counterpart: block-Jacobi preconditioned conjugate
 * gradients on the same input, FP64, x0 = 0, stopping at |r| <= relative_tolerance |b|
 * or max_iterations.  *iterations receives the iterations taken (may be NULL).
 * Returns SGPU_ERR_NON_FINITE on breakdown (p.Sp <= 0, e.g. an indefinite matrix). */
sgpu_status sgpu_block_pcg_solve(int64_t num_blocks, const int64_t* col_ptr,
                                 const int64_t* row_idx, const double* values,
                                 const double* rhs, double* x, double relative_tolerance,
                                 int64_t max_iterations, int64_t* iterations);

/* ---- learned block-Jacobi correction (no reference counterpart) -------------------
 * For camera block b of the reduced camera system with D = S_bb: J = diag(D)^-1/2,
 * J D J = L L^T, B = J L^-T.  A 3-layer MLP (32 -> 128 -> 128 -> 21, tanh hidden
 * layers) maps 32 features of the block (the correlations and log-scales of D, the
 * neighbour count, norms of the scaled off-diagonal blocks) to a lower-triangular G
 * (diagonal exp(o_ii), row-major lower-triangle order) and the preconditioner block
 * is M_b^-1 = B G G^T B^T (G = I: block-Jacobi).  Inference uses bf16 weights with FP32
 * accumulation.  Weight layout (FP32): W1[128][32], b1[128], W2[128][128], b2[128],
 * W3[21][128], b3[21]. */
#define SGPU_PRECOND_FEATURES 32
#define SGPU_PRECOND_OUTPUTS 21
#define SGPU_PRECOND_WEIGHTS 23445

/* Features of every camera block of a block system given like sgpu_block_ldlt_solve's
 * input; features_out is num_blocks x 32. */
sgpu_status sgpu_precond_features(int64_t num_blocks, const int64_t* col_ptr,
                                  const int64_t* row_idx, const double* values,
                                  float* features_out);
/* MLP outputs (n x 21) for n feature rows; use_bf16 = 1 rounds the weights to bf16. */
sgpu_status sgpu_precond_forward(int64_t n, const float* features, const float* weights,
                                 int32_t use_bf16, float* outputs);
/* M_b^-1 = B G G^T B^T for n diagonal blocks (n x 36, row-major) and outputs (n x 21);
 * blocks that are not positive definite give NaN. */
sgpu_status sgpu_precond_blocks(int64_t n, const double* diag_blocks, const float* outputs,
                                double* minv_out);
/* Training loss and its gradient for n blocks: loss = sum_b |G G^T B^T u_b - w_b|^2 /
 * |w_b|^2 (FP32 network, FP64 loss), grad_out = d loss / d weights (SGPU_PRECOND_WEIGHTS). */
sgpu_status sgpu_precond_loss_grad(int64_t n, const float* features, const double* diag_blocks,
                                   const double* u, const double* w, const float* weights,
                                   double* loss_out, float* grad_out);

typedef struct sgpu_precond_train_params {
  int32_t num_problems;  /* seeded random-walk BA problems generating the snapshots */
  int32_t num_cameras;
  int64_t num_points;
  double geometric_p;
  int32_t lm_iterations; /* LM steps per problem; one RCS snapshot per step */
  int32_t epochs;
  int32_t probes;        /* probe vectors per snapshot and step */
  int32_t device;
  double learning_rate;  /* Adam */
  uint64_t seed;
  int32_t batch_snapshots;   /* snapshots per Adam step (0: 1) */
  int32_t holdout_problems;  /* further problems (the seeds after the training ones) whose
                                snapshots are held out for the validation loss */
} sgpu_precond_train_params;

/* Trains the correction on the GPU (hand-written forward, backward and Adam kernels; the
 * layer products and weight gradients on the matrix cores) from RCS snapshots the solver
 * generates on seeded problems.  Each snapshot is kept as its camera blocks' features and
 * `probes` fixed probe pairs (w, u~); an Adam step takes batch_snapshots snapshots in a
 * shuffled order.  weights_inout holds the initial weights and receives the trained ones;
 * loss_history and val_history (epochs + 1 entries each, may be NULL) receive the mean
 * training and held-out loss before training and after each epoch. */
sgpu_status sgpu_precond_train(const sgpu_precond_train_params* params, float* weights_inout,
                               double* loss_history, double* val_history, int64_t* num_snapshots);

/* ---- batched small windows (no reference counterpart) --------------------------
 * Many independent bundle-adjustment windows solved together: per window one shared
 * pinhole (Conditioned), SE(3) poses (Free, camera 0 of the window constant), points
 * (Marginalized) and pinhole reprojection terms (Huber if huber_delta > 0).  Each
 * window runs the same LM as sgpu_optimize_nlls (same rules, tolerances and
 * termination) with its reduced camera system solved densely in LDS; at most 31
 * cameras per window.  Arrays are concatenated over windows with prefix offsets. */
typedef struct sgpu_batch sgpu_batch;
typedef struct sgpu_batch_desc {
  int64_t num_problems;
  const int64_t* cam_ptr;    /* num_problems + 1 */
  const int64_t* point_ptr;  /* num_problems + 1 */
  const int64_t* obs_ptr;    /* num_problems + 1 */
  const double* intrinsics;  /* num_problems x 4 (fx, fy, cx, cy) */
  const double* poses;       /* cameras x 7, world_from_camera */
  const double* points;      /* points x 3 */
  const int32_t* obs_cam;    /* per observation: camera index within its window */
  const int32_t* obs_point;  /* per observation: point index within its window */
  const double* obs_uv;      /* observations x 2 */
  double huber_delta;        /* <= 0: no robust kernel */
} sgpu_batch_desc;

/* Uploads the windows and picks the workgroup size (128, 256 or 512 threads) with a
 * timed LM iteration of the whole batch per candidate (the autotuner).  params: the
 * LM parameters of every window; solver must be SCHUR_BLOCK_SPARSE_LDLT; precision
 * MIXED computes Jacobians and Schur products in FP32. */
sgpu_status sgpu_batch_create(const sgpu_batch_desc* desc, const sgpu_opt_params* params,
                              sgpu_batch** out);
/* Advances every unfinished window by up to `iterations` LM iterations (one launch
 * per iteration; use_graph = 1 replays them from a captured hipGraph). */
sgpu_status sgpu_batch_run(sgpu_batch* batch, int32_t iterations, int32_t use_graph,
                           double* seconds);
/* The same in one persistent launch: each workgroup (as many as are resident) takes the
 * next window from a queue, largest first, and runs up to `iterations` LM iterations of it
 * until it terminates, then takes the next; use_graph = 1 replays the launch from a
 * captured hipGraph.  Results are those of sgpu_batch_run with the same iterations. */
sgpu_status sgpu_batch_run_persistent(sgpu_batch* batch, int32_t iterations, int32_t use_graph,
                                      double* seconds);
/* Per window, from the device wall clock of the last persistent run: when a workgroup took
 * it and when its run ended, in ms after the first window was taken (num_problems each). */
sgpu_status sgpu_batch_latencies(const sgpu_batch* batch, double* start_ms, double* end_ms);
/* Restarts every window from its initial values. */
sgpu_status sgpu_batch_reset(sgpu_batch* batch);
/* Final variables (same layout as the description), costs, steps and termination
 * (sgpu_termination, NONE while running) per window; any pointer may be NULL. */
sgpu_status sgpu_batch_results(sgpu_batch* batch, double* poses_out, double* points_out,
                               double* initial_cost, double* final_cost, int32_t* num_steps,
                               int32_t* termination);
/* The autotuner's choices: the per-iteration kernel's workgroup size and its three timings
 * (ms, one LM iteration of the batch at 128, 256 and 512 threads), the LDS bytes per
 * workgroup, the persistent kernel's size, its three timings (one persistent launch of one
 * iteration per window) and its grid (the resident workgroups).  Any pointer may be NULL. */
sgpu_status sgpu_batch_info(const sgpu_batch* batch, int32_t* threads, double* tune_ms3,
                            int64_t* lds_bytes, int32_t* persistent_threads,
                            double* persistent_tune_ms3, int32_t* persistent_grid);
void sgpu_batch_destroy(sgpu_batch* batch);

/* ---- synthetic problems ---------------------------------------------------------
 * Seeded BA generator (SplitMix64 -> xoshiro256**, Box-Muller), the input of the
 * parity tests and the benchmark.  Two-call protocol: call with NULL arrays to get
 * the sizes, then with arrays of those sizes. */
typedef struct sgpu_ba_spec {
  int32_t layout;       /* 0 = ring of cameras (cube of points), 1 = random-walk */
  int32_t num_cameras;
  int64_t num_points;
  int32_t obs_per_point_ring; /* layout 0: distinct cameras per point; 0 = uniform in
                                 [min_obs, max_obs] per point */
  int32_t min_obs, max_obs;   /* layout 1: clip of the geometric draw (layout 0: range) */
  double geometric_p;         /* layout 1: success probability */
  double pixel_noise;         /* px sigma */
  double rot_noise;           /* rad sigma of the pose perturbation */
  double trans_noise;         /* m sigma of the pose perturbation */
  double point_noise;         /* m sigma of the point perturbation */
  double step_length;         /* layout 1 */
  double yaw_jitter;          /* layout 1, rad sigma */
  uint64_t seed;
} sgpu_ba_spec;

typedef struct sgpu_ba_sizes {
  int64_t num_cameras;
  int64_t num_points;
  int64_t num_obs;
} sgpu_ba_sizes;

sgpu_status sgpu_generate_ba(const sgpu_ba_spec* spec, sgpu_ba_sizes* sizes,
                             double* intrinsics /*4*/, double* poses_gt /*7 x cams*/,
                             double* poses_init, double* points_gt /*3 x pts*/,
                             double* points_init, int64_t* obs_idx /*3 x obs: cam,pose,pt*/,
                             double* obs_uv /*2 x obs*/);

/* The reference's own BaProblem geometry (example_problems/ba_problem.rs:154-345) with this
 * library's RNG: num_cameras on a ring of radius 8 looking into an open box of num_points
 * candidate points, one shared EUCM camera; points seen fewer than twice are dropped.
 * obs_idx rows are (intrinsics 0, pose, point).  Poses 0 and 1 are unperturbed (the gauge
 * the reference holds constant).  Two-call protocol as above. */
sgpu_status sgpu_generate_ba_eucm(int32_t num_cameras, int64_t num_points, uint64_t seed,
                                  sgpu_ba_sizes* sizes, double* intrinsics_gt /*6*/,
                                  double* intrinsics_init /*6*/, double* poses_gt,
                                  double* poses_init, double* points_gt, double* points_init,
                                  int64_t* obs_idx, double* obs_uv);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* SOPHUS_GPU_H */
