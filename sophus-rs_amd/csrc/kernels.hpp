// Device kernels of the Schur LM iteration (launch wrappers; all asynchronous on `stream`).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "problem.hpp"

namespace sgpu {

// Per-iteration device scalars (one struct in device memory, copied back once per step).
struct DevScalars {
  double cost;            // merit of the evaluated variables
  double dd, dg;          // delta.delta, delta.neg_grad
  unsigned long long max_delta_bits;  // max |delta| as the bits of a non-negative double
  unsigned long long max_grad_bits;   // max |neg_grad|
  int32_t singular_marg;  // an H_mm block had det == 0
  int32_t bad_pivot;      // non-finite RCS pivot
  int32_t pcg_iterations;
  int32_t pad;
};

// Read-only problem arrays on the device.
struct DevProblem {
  int32_t nb, npt, nobs, npose, nintr;
  const int32_t* obs_pose;
  const int32_t* obs_point;
  const int32_t* obs_intr;
  const double* obs_uv;
  const int32_t* pt_ptr;
  const int32_t* pt_nu;
  const int32_t* pt_block;  // point member -> active index or -1
  const int32_t* cam_nu;
  const int32_t* pose_block;
  const int32_t* cam_member;
  const int32_t* cam_ptr;
  const int32_t* cam_obs;
  const int32_t* s_ptr;
  const int32_t* s_col;
  const int32_t* s_row;     // slot -> row camera
  const int64_t* pair_ptr;  // slot -> cross-pair range
  const int32_t* pairs;     // (q_i, q_j) per cross pair: positions in cam_obs
  const int32_t* own_chunks;     // (camera, q0, q1) per own-term chunk
  // camera-major copies of the observations' point, intrinsics and uv (positions q of cam_obs)
  const int32_t* cm_point;
  const int32_t* cm_intr;
  const double* cm_uv;
  const int32_t* cam_chunk_ptr;  // nb + 1
  int32_t n_own_chunks;
  int64_t nslot;
  const int64_t* s_off;
  const int32_t* s_ld;
  int robust;
  double delta;
  int32_t cam_model;     // 0 pinhole, 1 EUCM
  // free shared intrinsics: their free block (after the poses; -1 when conditioned), their
  // member, the S slot of (intrinsics, pose block) per pose block (-1: none) and of
  // (intrinsics, intrinsics)
  int32_t intr_block, intr_member;
  const int32_t* intr_slot;
  int64_t intr_diag_slot;
  int32_t n_eq;  // equality-constraint sets (their KKT block adds one step-sum partial)
};

// Equality-constraint sets on the device: translation-norm terms on poses, the terms of a set
// contiguous and the sets in order.
struct DevEq {
  int32_t nset, nterm;
  const int32_t* term_member;  // pose member per term
  const int32_t* set_ptr;      // nset + 1
  const double* radius;        // per term
  double* G;                   // 7 per term: dc/d(pose tangent) (6), then c
  double* c;                   // per set: the sum of its terms' residuals
  double* Y;                   // nset x 6 nb: S^-1 G_k^T (camera order)
  double* b;                   // 6 nb: G_k^T (camera order)
  double* bperm;               // 6 nb
  double* z;                   // 6 nb
  double* lambda[2];           // multipliers per set (indexed like the variables' buffers)
};

// RCS supernodal multifrontal factorization on the device.
struct DevFactor {
  int32_t nb, nsn;
  const int32_t* perm;  // perm[new] = old
  const int32_t* sn_c0;
  const int32_t* sn_nc;
  const int32_t* sn_nf;
  const int64_t* sn_off;
  const int64_t* sn_loff;
  const int32_t* sn_rows_ptr;
  const int32_t* sn_rows;
  const int32_t* sn_parent;
  const int32_t* sn_upos_ptr;
  const int32_t* sn_upos;
  const int32_t* lvl_sn;
  const int32_t* launch_tiles;  // tiles of the rounds: (node, panel, tile_row, tile_col)
  const int32_t* leaves;        // nodes without children
  const int32_t* sn_hand;       // 1: the node's last step factors its parent's first block
  const int32_t* solve_cols;  // backward solve: (node, kSolveCols-column tile) per level
  const int64_t* sn_uoff;     // right-hand side offsets (F entries per node)
  double* usolve;             // each node's right-hand side in front row order (u_elems)
  hipEvent_t* step_events;    // optional: 2 per round of panel steps, around its launches
  const int32_t* panel_items;  // (node, panel, kPanelRows-row chunk) per panel-row workgroup
  double* wbuf;                // per node F x kPW: unscaled W of the current panel (ld F)
  const int64_t* sn_woff;
  double* tinv;                // per panel kPW x kPW row-major: L_PP^-T (unit block diagonal)
  const int64_t* sn_poff;
  double* ltbuf;               // per node K x K: L_KK row-major (backward solve)
  const int64_t* sn_toff;
  double* front;   // all fronts, column-major (workspace)
  double* lbuf;    // per node F x K: the factor's L columns (column-major, ld F)
  double* piv_L;   // nb x 36 unit-lower pivot factors (row-major)
  double* piv_d;   // nb x 6
  double* piv_M;   // nb x 36: (L_jj D_j L_jj^T)^+ = L_jj^-T D_j^-1 L_jj^-1 (0 rows for zero pivots)
  int64_t front_elems;
  const Symbolic* host;  // launch schedule
  const int32_t* sn_child_ptr;  // children of each node (standalone forward solve; may be null
  const int32_t* sn_child;      // where that solve is not used)
  // pivot rule: 0 the 6x6 block rule of BlockSparseLdlt (dense_ldlt.rs:164-256), 1 the
  // scalar rule of SparseLdlt (sparse_ldlt.rs:399-445)
  int32_t scalar_pivots;
};

constexpr int kPtRec = 18;  // doubles per packed point record (DevLin::prec)

// Linearization workspace.
struct DevLin {
  double* cw;         // npose x 12: camera_from_world rotation (row-major) + translation
  double* prec;       // npt x kPtRec packed point record: x (3), H_ll^-1 upper (00, 01, 02, 11,
                      // 12, 22), y (3), damped H_ll upper (6); zeros but x for a constant point
  double* pt_s;       // npt x 3: neg-gradient of the point (sum of A^T r~)
  double* rhs_nat;    // nb x 6: reduced rhs in camera order
  double* grad_cam;   // nb x 6: neg-gradient of the cameras
  double* x_perm;     // nb x 6 solve buffer in permuted order
  double* x_aux;      // nb x 6 second solve buffer
  double* dx_cam;     // nb x 6 camera step (camera order)
  double* bsr;        // nslot x 36 lower S blocks for the iterative solvers (else null)
  void* vobs;         // 12 per camera-major observation: U_o and lift point (double, or float when mixed)
  double* own_part;   // 48 per own-term chunk: the chunk's (o, o) block, rhs and gradient sums
  double* partial;    // reduction partials
  int32_t partial_cap;
  // free shared intrinsics only (else null): per point V_b = W_b H_ll^-1 (6x3 row-major) then
  // W_b = H_cl (6x3); per point-pass workgroup the partial sums of the intrinsics rows
  double* pt_vw;
  double* intr_part;
  double* intr_cpart;  // 36 per own-term chunk: the chunk's share of its (intrinsics, camera) block
};

void launch_cam_cache(hipStream_t s, int npose, const double* poses, double* cw);
// precision: 0 = FP64, 1 = FP32 Jacobians/assembly/Schur products
void launch_point_prep(hipStream_t s, int precision, const DevProblem& p, const double* poses,
                       const double* points, const double* intr, const double* cw, double nu,
                       DevLin& w, DevScalars* sc);
void launch_camera_rows(hipStream_t s, int precision, const DevProblem& p, const double* points,
                        const double* intr, const double* cw, double nu, DevLin& w,
                        DevFactor& f, DevScalars* sc);
void launch_fill(hipStream_t s, double* ptr, int64_t n, double v);
// Zero the lower triangle of every front (before the Schur assembly).
void launch_zero_fronts(hipStream_t s, const DevFactor& f);
// Full numeric factorization (all levels and panels) of the assembled fronts, with the
// forward solve of rhs (camera order) folded in: y (nb*6) receives z = D^-1 L^-1 P rhs.
void launch_rcs_factor(hipStream_t s, const DevFactor& f, DevScalars* sc, const double* rhs_nat,
                       double* y);
// Backward solve with the factor: y (from launch_rcs_factor) -> dx in camera order.
void launch_rcs_solve(hipStream_t s, const DevFactor& f, double* y, double* dx_cam);
// Standalone forward solve with a factored RCS: z = D^+ L^-1 P b for a right-hand side b in
// camera order (node right-hand sides in f.usolve are overwritten); z in elimination order,
// ready for launch_rcs_solve.
void launch_rcs_forward(hipStream_t s, const DevFactor& f, const double* b_nat, double* b_perm,
                        double* z);
// H_{c,b} = J_cam^T J_point (6 x 3 row-major) of every observation of point b with a free
// camera, at the point's current value: out[18 q + ..] for its q-th observation, cam[q] its
// free block (-1: a constant camera); hinv (9, row-major) receives the point's H_ll^-1 from
// its packed record.  nq observations of the point.
void launch_point_coupling(hipStream_t s, const DevProblem& p, const double* points,
                           const double* intr, const double* cw, const double* prec, int b,
                           double* out, int32_t* cam, double* hinv);
void launch_backsub_points(hipStream_t s, int precision, const DevProblem& p,
                           const double* points, const double* intr, const double* cw,
                           const DevLin& w, double* points_new, DevScalars* sc);
// Pose update T <- exp(delta) T into poses_new, and the free intrinsics' additive update into
// intr_new (which holds a copy of intr).
void launch_update_poses(hipStream_t s, const DevProblem& p, const double* poses,
                         const double* intr, const DevLin& w, double* poses_new, double* intr_new,
                         DevScalars* sc);
// Free shared intrinsics (p.intr_block >= 0): V_b and W_b per point, the S blocks
// (intrinsics, pose) and (intrinsics, intrinsics), the intrinsics' reduced rhs and
// neg-gradient.  After launch_camera_rows (which leaves those slots zero but the damping).
void launch_intrinsics_rows(hipStream_t s, const DevProblem& p, const double* points,
                            const double* intr, const double* cw, DevLin& w, DevFactor& f);
// doubles of DevLin::intr_part for the problem
int intrinsics_partials(const DevProblem& p);
// Equality constraints (range-space KKT, schur_ldlt.rs:450-480, 674-730, 762-772):
// launch_eq_eval evaluates c and G at `poses` and, with add_rhs, adds G^T lambda to the free
// rows of rhs_nat and grad_cam (eq_system.rs:207-262); launch_eq_solve, after the RCS factor
// and the backward solve of u = S^-1 rhs into dx_cam, forms Y = S^-1 G^T (one standalone
// forward and backward solve per set), M = G Y, dlambda = M^-1 (G u + c), dx_cam = u - Y
// dlambda, lambda_next = lambda - dlambda, and the step sums' multiplier rows (|dlambda|^2,
// -c.dlambda, max |dlambda|).  A singular M sets the bad-pivot flag (the step is rejected).
void launch_eq_eval(hipStream_t s, const DevProblem& p, const DevEq& q, const double* poses,
                    const double* lambda, DevLin& w, bool add_rhs);
void launch_eq_solve(hipStream_t s, const DevProblem& p, const DevEq& q, const DevFactor& f,
                     const double* lambda, double* lambda_next, DevLin& w, DevScalars* sc);
void launch_cost(hipStream_t s, const DevProblem& p, const double* poses, const double* points,
                 const double* intr, double* partial, int32_t partial_cap, DevScalars* sc);
void launch_finalize_step(hipStream_t s, const DevProblem& p, const DevLin& w, DevScalars* sc);
void launch_finalize_cost(hipStream_t s, const double* partial, int nblocks, DevScalars* sc);
void launch_max_grad(hipStream_t s, const double* grad_cam, int nb, DevScalars* sc);
void launch_zero_scalars(hipStream_t s, DevScalars* sc);

// pose-graph terms (SE(3) between factor)
// Pose terms of one cost function: SE(3) between factors, or (prior) Isometry3 priors, whose
// term t has pose m[t] (nn[t] = m[t]), the prior mean in meas[7t..] and the precision W in
// prec[36t..] (row-major).
struct DevPoseGraph {
  int32_t n;
  const int32_t* m;
  const int32_t* nn;
  const double* meas;
  const uint8_t* first;
  const int32_t* dest;  // 4 per term: S slots of (m,m), (n,n), (max,min) or -1; flip
  int robust;
  double delta;
  int prior;
  const double* prec;
};
void launch_pose_graph_linearize(hipStream_t s, const DevPoseGraph& g, const DevProblem& p,
                                 const double* poses, double nu, DevLin& w, DevFactor& f,
                                 const double* centered_cw);
// Converts centred camera steps back: delta = G delta' (nu += c x omega).
void launch_uncenter(hipStream_t s, const DevProblem& p, const double* cw, double* dx_cam);
// Diagonal of the assembled S (6 nb, camera order) from the fronts, before the factorization.
void launch_gather_s_diag(hipStream_t s, const DevProblem& p, const double* front, double* out);
// Scatter S blocks given per slot (36 doubles, row-major, camera order) into the fronts.
void launch_scatter_s(hipStream_t s, const DevProblem& p, int64_t nslots, const double* blocks,
                      DevFactor& f);
void launch_pose_graph_cost(hipStream_t s, const DevPoseGraph& g, const double* poses,
                            double* partial, int32_t partial_cap, DevScalars* sc);

// stage-level evaluation kernels
void launch_eval_pinhole(hipStream_t s, int precision, int64_t n, const double* intr,
                         const double* pose, const double* pt, const double* uv, int robust,
                         double delta, double* res, double* jpose, double* jpt, double* cost,
                         double* weight);
void launch_eval_eucm(hipStream_t s, int64_t n, const double* intr, const double* pose,
                      const double* pt, const double* uv, int robust, double delta, double* res,
                      double* jintr, double* jpose, double* jpt, double* cost, double* weight);
void launch_eval_pose_graph(hipStream_t s, int64_t n, const double* m, const double* nn,
                            const double* mn, double* res, double* jm, double* jn);
void launch_eval_prior(hipStream_t s, int64_t n, const double* poses, const double* means,
                       const double* prec, int robust, double delta, double* res, double* jac,
                       double* cost);

constexpr int kCostBlocks = 1024;

}  // namespace sgpu
