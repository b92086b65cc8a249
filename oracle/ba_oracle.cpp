// ORACLE — test infrastructure only (parity checker and CPU baseline).  Nothing in
// the product library links, loads or calls this file; only tests/, the smoke check
// and the benchmark's CPU-baseline leg load liboracle.so.
//
// A CPU FP64 restatement of sophus-rs's Schur-complement LM path:
//   optimize_nlls / Optimizer::new / step / build_solution
//       crates/sophus_opt/src/nlls.rs:73-104, crates/sophus_opt/src/nlls/optimizer.rs
//   step_nlls_inner            crates/sophus_opt/src/nlls/optimizer_step.rs:366-443
//   compute_gain_ratio, calc_merit  crates/sophus_opt/src/nlls/optimizer_helpers.rs:159-173,309-321
//   CostFn::sort / eval / calc_total_cost  crates/sophus_opt/src/nlls/cost/cost_fn.rs:173-433
//   make (r~ = w r; g = J^T r~; H = J^T J; cost = r~^T r~)
//                               crates/sophus_opt/src/nlls/cost/evaluated_term.rs:286-327
//   accumulate_term (per-term nu I)  crates/sophus_opt/src/nlls/linear_system/cost_system.rs:290-376
//   VarFamilies layout          crates/sophus_opt/src/variables/var_families.rs:60-113
//   SchurFactor::factorize / solve  crates/sophus_solver/src/ldlt/schur_ldlt.rs:82-182,536-812
//   BlockSparseLdlt             crates/sophus_solver/src/ldlt/block_sparse_ldlt.rs:96-190,400-747
//   ldlt_decompose_inplace      crates/sophus_solver/src/ldlt/dense_ldlt.rs:164-256
//   SparseLdlt (SchurSparseLdlt)  crates/sophus_solver/src/ldlt/sparse_ldlt.rs:121-445,571-650
//   Isometry3PriorCostTerm      crates/sophus_opt/src/nlls/functor_library/costs/isometry3_prior.rs:46-88
//   EliminationTree             crates/sophus_solver/src/ldlt/elimination_tree.rs:20-105
//
// Documented differences (rounding-level only):
//   - S is stored with its structural non-zero blocks only; the reference also
//     stores every all-zero free-free block (schur_ldlt.rs:572-600).  Zero blocks
//     contribute exact zeros to every sum.
//   - The RCS block ordering is a deterministic minimum-degree ordering written
//     here; the reference uses faer's AMD (crates/sophus_solver/src/ldlt.rs:49-67),
//     which is not available.  Orderings change rounding only.  A caller may set the
//     ordering instead (oracle_set_rcs_order), e.g. to the one the GPU library uses, so
//     that both eliminate the blocks in the same order.
//   - Sums run sequentially in term order (the reference's rayon sums are
//     order-unspecified).
//   - Bunch-Kaufman fallback (block_diag_ldlt.rs:97-114) is not restated: it only
//     triggers for indefinite blocks, and every RCS block here is PD (damped).
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <omp.h>
#include <set>
#include <string>
#include <vector>

#include "../include/sophus_gpu.h"
#include "oracle_lie.hpp"

namespace oracle {

thread_local std::string g_error;

sgpu_status fail(sgpu_status s, const std::string& msg) {
  g_error = msg;
  return s;
}

// Host threads for the per-term and per-point loops (oracle_set_threads; 1 by default).
// The parallel loops keep every output's accumulation order: per-term and per-point values
// are formed independently, each row of H, S and the right-hand side is written by one
// thread only (row % threads), which visits that row's contributions in the sequential
// order, and scalar sums are added in term order.  Results are bitwise the same for any
// thread count (the reference's rayon paths, cost_system.rs:119-160 and
// schur_ldlt.rs:282-360, instead re-associate the sums).
int g_threads = 1;

template <class Fn>
void run_threads(Fn&& fn) {
  if (g_threads <= 1) {
    fn(0, 1);
    return;
  }
#pragma omp parallel num_threads(g_threads)
  fn(omp_get_thread_num(), omp_get_num_threads());
}

int var_dof(int type) {
  return type == SGPU_VAR_ISOMETRY3 || type == SGPU_VAR_EUCM ? 6 : type == SGPU_VAR_VEC3 ? 3 : 4;
}
int var_params(int type) {
  return type == SGPU_VAR_ISOMETRY3 ? 7 : type == SGPU_VAR_VEC3 ? 3 : type == SGPU_VAR_EUCM ? 6 : 4;
}
bool is_reprojection(int kind) {
  return kind == SGPU_COST_PINHOLE_REPROJECTION || kind == SGPU_COST_EUCM_REPROJECTION;
}

// ---- robust kernel: crates/sophus_opt/src/robust_kernel.rs:13-35 -----------------------
double huber_apply(double delta, double x) {
  if (std::fabs(x) <= delta) return x * x;
  return 2.0 * delta * std::fabs(x) - delta * delta;
}
double robust_weight(int kind, double delta, double nrm) {
  if (kind != SGPU_ROBUST_HUBER) return 1.0;
  if (nrm == 0.0) return 0.0;
  return std::sqrt(huber_apply(delta, nrm)) / nrm;
}

// ---- variables -----------------------------------------------------------------------
struct Family {
  std::string name;
  int type, kind, dof, nparams;
  int64_t n;
  std::vector<double> members;
  std::vector<char> is_const;
  std::vector<int64_t> scalar_start;  // -1 for Conditioned / const
  std::vector<int64_t> block_id;      // global block id, -1 if none
};

struct Vars {
  std::vector<Family> fam;  // sorted by name (BTreeMap order)
  int64_t num_free_scalars = 0, num_marg_scalars = 0;
  int64_t num_blocks = 0, num_free_blocks = 0;
  std::vector<int64_t> block_start, block_dim;  // per global block id
  int find(const std::string& n) const {
    for (size_t i = 0; i < fam.size(); ++i)
      if (fam[i].name == n) return int(i);
    return -1;
  }
};

// VarFamilies::new (var_families.rs:60-113): Free families in name order, then
// Conditioned (no indices), then Marginalized; const members get -1.
void layout(Vars& v) {
  int64_t off = 0, blk = 0;
  v.block_start.clear();
  v.block_dim.clear();
  auto assign = [&](Family& f) {
    f.scalar_start.assign(f.n, -1);
    f.block_id.assign(f.n, -1);
    for (int64_t i = 0; i < f.n; ++i) {
      if (f.is_const[i]) continue;
      f.scalar_start[i] = off;
      f.block_id[i] = blk++;
      v.block_start.push_back(off);
      v.block_dim.push_back(f.dof);
      off += f.dof;
    }
  };
  for (auto& f : v.fam)
    if (f.kind == SGPU_VAR_FREE) assign(f);
  for (auto& f : v.fam)
    if (f.kind == SGPU_VAR_CONDITIONED) {
      f.scalar_start.assign(f.n, -1);
      f.block_id.assign(f.n, -1);
    }
  v.num_free_scalars = off;
  v.num_free_blocks = blk;
  for (auto& f : v.fam)
    if (f.kind == SGPU_VAR_MARGINALIZED) assign(f);
  v.num_marg_scalars = off - v.num_free_scalars;
  v.num_blocks = blk;
}

// IsVariable::update: SE(3) left-multiplicative exp(delta) * T (lie_group_manifold.rs:53-65),
// vectors and pinhole parameters additive (manifold.rs:66-75, perspective_camera.rs:81-89).
Vars update(const Vars& v, const std::vector<double>& delta) {
  Vars u = v;
  for (auto& f : u.fam) {
    for (int64_t i = 0; i < f.n; ++i) {
      if (f.scalar_start[i] < 0) continue;
      const double* d = &delta[f.scalar_start[i]];
      double* m = &f.members[i * f.nparams];
      if (f.type == SGPU_VAR_ISOMETRY3) {
        double e[7], out[7];
        se3_exp(d, e);
        se3_mul(e, m, out);
        std::memcpy(m, out, sizeof(out));
      } else {
        for (int k = 0; k < f.dof; ++k) m[k] += d[k];
      }
    }
  }
  return u;
}

// ---- cost terms -------------------------------------------------------------------------
struct CostFn {
  int kind, robust;
  double delta;
  int arity;
  int fam[3];
  std::vector<int64_t> idx;   // sorted terms, arity per term
  std::vector<double> meas;   // 2 or 7 per term
  std::vector<int64_t> range_start;  // reduction ranges (sorted)
  int mlen() const { return is_reprojection(kind) ? 2 : kind == SGPU_COST_ISOMETRY3_PRIOR ? 43 : 7; }
  int64_t num_terms() const { return int64_t(idx.size()) / arity; }
};

// Evaluated term: residual dim R, args with Jacobians (R x dof) row-major.
struct EvalTerm {
  int R = 0;
  double r[6];          // weighted residual r~
  double cost = 0;
  double J[3][6 * 6];   // per arg, R x dof
  bool has[3] = {false, false, false};
  const double* W = nullptr;  // precision (R x R row-major) or none
};

// PinholeCameraReprojectionCostTerm::eval (pinhole_reprojection.rs:60-95).
void eval_pinhole(const double* intr, const double* pose, const double* xw, const double* uv,
                  bool d_intr, bool d_pose, bool d_pt, double r[2], double Jc[8],
                  double Jp[12], double Jx[6]) {
  double cw[7], xc[3];
  se3_inverse(pose, cw);
  se3_transform(cw, xw, xc);
  const double px = xc[0] / xc[2], py = xc[1] / xc[2];
  r[0] = uv[0] - (px * intr[0] + intr[2]);
  r[1] = uv[1] - (py * intr[1] + intr[3]);
  // d_pi = dx_distort_x * dx_proj_x * R_cw (pinhole_reprojection.rs:76-80,
  // perspective.rs:54-72, affine.rs:77-85)
  const double iz = 1.0 / xc[2];
  const double zz = xc[2] * xc[2];
  const double dp[6] = {intr[0] * iz, 0.0, intr[0] * (-xc[0] / zz),
                        0.0, intr[1] * iz, intr[1] * (-xc[1] / zz)};
  double R[9];
  so3_matrix(cw, R);
  double A[6];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 3; ++j)
      A[3 * i + j] = dp[3 * i] * R[j] + dp[3 * i + 1] * R[3 + j] + dp[3 * i + 2] * R[6 + j];
  if (d_intr) {  // -dx_distort_params (affine.rs:87-108)
    const double v[8] = {px, 0, 1, 0, 0, py, 0, 1};
    for (int k = 0; k < 8; ++k) Jc[k] = -v[k];
  }
  if (d_pose) {  // A * [hat(-x_w) | I]  (affine_group_template.rs:505-510, rotation3.rs:575-577)
    const double nx[3] = {-xw[0], -xw[1], -xw[2]};
    double H[9];
    hat(nx, H);
    for (int i = 0; i < 2; ++i) {
      for (int j = 0; j < 3; ++j)
        Jp[6 * i + j] = A[3 * i] * H[j] + A[3 * i + 1] * H[3 + j] + A[3 * i + 2] * H[6 + j];
      for (int j = 0; j < 3; ++j) Jp[6 * i + 3 + j] = A[3 * i + j];
    }
  }
  if (d_pt)
    for (int k = 0; k < 6; ++k) Jx[k] = -A[k];
}

// EucmCameraReprojectionCostTerm::eval (example_problems/ba_problem.rs:68-110), with the
// analytic derivatives of crates/sophus_sensor/src/distortions/enhanced_unified.rs:125-215:
//   d_pi_r_cw = dx_distort_x(proj) * dx_proj_x(x_c) * R_cw          (2 x 3)
//   J_intr = -dx_distort_params(proj)                                (2 x 6)
//   J_pose = d_pi_r_cw * dx_exp_x_times_point_at_0(x_w) = d_pi_r_cw [hat(-x_w) | I]
//   J_point = -d_pi_r_cw
void eval_eucm(const double* p, const double* pose, const double* xw, const double* uv,
               bool d_intr, bool d_pose, bool d_pt, double r[2], double Jc[12], double Jp[12],
               double Jx[6]) {
  double cw[7], xc[3];
  se3_inverse(pose, cw);
  se3_transform(cw, xw, xc);
  const double x = xc[0] / xc[2], y = xc[1] / xc[2];
  const double fx = p[0], fy = p[1], alpha = p[4], beta = p[5];
  const double r2 = x * x + y * y;
  const double rho2 = beta * r2 + 1.0;
  const double rho = std::sqrt(rho2);
  const double norm = alpha * rho + (1.0 - alpha);
  const double mx = x / norm, my = y / norm;
  r[0] = uv[0] - (fx * mx + p[2]);
  r[1] = uv[1] - (fy * my + p[3]);
  const double nn = norm * norm;
  // dx_distort_x (enhanced_unified.rs:125-162)
  const double drho_dx = (2.0 * beta * x) / (2.0 * rho), drho_dy = (2.0 * beta * y) / (2.0 * rho);
  const double dnorm_dx = alpha * drho_dx, dnorm_dy = alpha * drho_dy;
  const double Dd[4] = {fx * ((norm - x * dnorm_dx) / nn), fx * (-x * dnorm_dy / nn),
                        fy * (-y * dnorm_dx / nn), fy * ((norm - y * dnorm_dy) / nn)};
  // dx_proj_x (perspective.rs:54-72)
  const double iz = 1.0 / xc[2], zz = xc[2] * xc[2];
  const double Dp[6] = {iz, 0.0, -xc[0] / zz, 0.0, iz, -xc[1] / zz};
  double DD[6];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 3; ++j) DD[3 * i + j] = Dd[2 * i] * Dp[j] + Dd[2 * i + 1] * Dp[3 + j];
  double R[9];
  so3_matrix(cw, R);
  double A[6];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 3; ++j)
      A[3 * i + j] = DD[3 * i] * R[j] + DD[3 * i + 1] * R[3 + j] + DD[3 * i + 2] * R[6 + j];
  if (d_intr) {  // -dx_distort_params (enhanced_unified.rs:164-215)
    const double drho_dbeta = r2 / (2.0 * rho);
    const double dnorm_dalpha = rho - 1.0, dnorm_dbeta = alpha * drho_dbeta;
    const double v[12] = {mx,  0.0, 1.0, 0.0, fx * (-x * dnorm_dalpha / nn), fx * (-x * dnorm_dbeta / nn),
                          0.0, my,  0.0, 1.0, fy * (-y * dnorm_dalpha / nn), fy * (-y * dnorm_dbeta / nn)};
    for (int k = 0; k < 12; ++k) Jc[k] = -v[k];
  }
  if (d_pose) {
    const double nx[3] = {-xw[0], -xw[1], -xw[2]};
    double H[9];
    hat(nx, H);
    for (int i = 0; i < 2; ++i) {
      for (int j = 0; j < 3; ++j)
        Jp[6 * i + j] = A[3 * i] * H[j] + A[3 * i + 1] * H[3 + j] + A[3 * i + 2] * H[6 + j];
      for (int j = 0; j < 3; ++j) Jp[6 * i + 3 + j] = A[3 * i + j];
    }
  }
  if (d_pt)
    for (int k = 0; k < 6; ++k) Jx[k] = -A[k];
}

// PoseGraph3CostTerm::eval (pose_graph3.rs:47-79): J_m = -dx_log_a_exp_x_b_at_0(a, b),
// J_n = +dx_log_a_exp_x_b_at_0(a, b), a = T_m^-1, b = T_n T_mn^-1;
// dx_log_a_exp_x_b_at_0 = J_l^-1(log(a b)) * Adj(a) (real_lie_group.rs:88-97).
void eval_pose_graph(const double* m, const double* n, const double* mn, double r[6],
                     double Jm[36], double Jn[36]) {
  pose_graph3_residual(m, n, mn, r);
  double a[7], mni[7], b[7], ab[7], t[6], Jl[36], Ad[36], D[36];
  se3_inverse(m, a);
  se3_inverse(mn, mni);
  se3_mul(n, mni, b);
  se3_mul(a, b, ab);
  se3_log(ab, t);
  se3_inv_left_jacobian(t, Jl);
  se3_adj(a, Ad);
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      double acc = 0;
      for (int k = 0; k < 6; ++k) acc += Jl[6 * i + k] * Ad[6 * k + j];
      D[6 * i + j] = acc;
    }
  for (int k = 0; k < 36; ++k) {
    Jm[k] = -D[k];
    Jn[k] = D[k];
  }
}

// Isometry3PriorCostTerm::eval (isometry3_prior.rs:46-88): r = log(T E^-1) and its Jacobian by
// forward-mode duals of log(exp(x) T E^-1) at x = 0, the reference's own derivative.
void eval_isometry3_prior(const double* pose, const double* mean, double r[6], double J[36]) {
  using D = Dual<6>;
  D x[6], E[7], T[7], Ei[7], Tx[7], TE[7], g[6];
  for (int k = 0; k < 6; ++k) x[k] = D::var(0.0, k);
  for (int k = 0; k < 7; ++k) {
    T[k] = D(pose[k]);
    Ei[k] = D(mean[k]);
  }
  se3_inverse(Ei, E);
  D ex[7];
  se3_exp(x, ex);
  se3_mul(ex, T, Tx);
  se3_mul(Tx, E, TE);
  se3_log(TE, g);
  for (int i = 0; i < 6; ++i) {
    r[i] = g[i].re;
    for (int k = 0; k < 6; ++k) J[6 * i + k] = g[i].d[k];
  }
}

// An equality-constraint set (EqConstraintFn, eq_constraint_fn.rs:61-140): terms of one kind;
// all of them share one constraint block (EqSystem::new gives each set one partition of one
// block, eq_system.rs:33-37, and populate_lower_triangular_kkt_mat adds every term into it,
// :207-262).
struct EqFn {
  int kind = SGPU_EQ_TRANSLATION_NORM;
  int fam = -1;
  std::vector<int64_t> idx;       // sorted terms (arity 1)
  std::vector<double> radius;     // one parameter per term
};

struct Problem {
  Vars vars;
  std::vector<CostFn> costs;
  std::vector<EqFn> eqs;
};

// TranslationNormConstraint (ba_scale_constraint.rs:44-99): c = |t(T)| - r, and its Jacobian
// by forward-mode duals of c(exp(x) T) at x = 0 (the reference's own derivative).
void eval_translation_norm(const double* pose, double r, double* c, double J[6]) {
  using D = Dual<6>;
  D x[6], E[7], T[7], Tx[7];
  for (int k = 0; k < 6; ++k) x[k] = D::var(0.0, k);
  for (int k = 0; k < 7; ++k) T[k] = D(pose[k]);
  se3_exp(x, E);
  se3_mul(E, T, Tx);
  const D n = sqrt(Tx[4] * Tx[4] + Tx[5] * Tx[5] + Tx[6] * Tx[6]) - D(r);
  *c = n.re;
  for (int k = 0; k < 6; ++k) J[k] = n.d[k];
}

// The KKT rows of the equality constraints at V (EqSystem, eq_system.rs:162-262): G (one
// dense row per set over the free scalars), -c per set, and G^T lambda added to the free
// rows of the neg-gradient.
struct Kkt {
  int64_t nc = 0;
  std::vector<double> G;      // nc x nfs row-major
  std::vector<double> rhs_c;  // -c
};

void eval_kkt(const Problem& P, const Vars& V, const std::vector<double>& lambda, Kkt& K,
              std::vector<double>* neg_grad) {
  const int64_t nfs = V.num_free_scalars;
  K.nc = int64_t(P.eqs.size());
  K.G.assign(size_t(K.nc * nfs), 0.0);
  K.rhs_c.assign(size_t(K.nc), 0.0);
  for (int64_t k = 0; k < K.nc; ++k) {
    const EqFn& e = P.eqs[k];
    const Family& f = V.fam[e.fam];
    for (size_t t = 0; t < e.idx.size(); ++t) {
      double c, J[6];
      eval_translation_norm(&f.members[7 * e.idx[t]], e.radius[t], &c, J);
      K.rhs_c[k] -= c;
      const int64_t s0 = f.kind == SGPU_VAR_CONDITIONED ? -1 : f.scalar_start[e.idx[t]];
      if (s0 < 0) continue;
      for (int i = 0; i < 6; ++i) {
        K.G[size_t(k * nfs + s0 + i)] += J[i];
        if (neg_grad) (*neg_grad)[s0 + i] += J[i] * lambda[k];
      }
    }
  }
}

// Evaluate one term.  `derivs` false = all args Conditioned (calc_total_cost).
void eval_term(const Problem& P, const Vars& V, const CostFn& c, int64_t t, bool derivs,
               EvalTerm& e) {
  const int64_t* ix = &c.idx[t * c.arity];
  const double* ms = &c.meas[t * c.mlen()];
  bool act[3] = {false, false, false};
  for (int a = 0; a < c.arity; ++a)
    act[a] = derivs && V.fam[c.fam[a]].kind != SGPU_VAR_CONDITIONED;
  double r[6];
  if (c.kind == SGPU_COST_PINHOLE_REPROJECTION) {
    e.R = 2;
    const double* intr = &V.fam[c.fam[0]].members[ix[0] * 4];
    const double* pose = &V.fam[c.fam[1]].members[ix[1] * 7];
    const double* xw = &V.fam[c.fam[2]].members[ix[2] * 3];
    eval_pinhole(intr, pose, xw, ms, act[0], act[1], act[2], r, e.J[0], e.J[1], e.J[2]);
  } else if (c.kind == SGPU_COST_EUCM_REPROJECTION) {
    e.R = 2;
    const double* intr = &V.fam[c.fam[0]].members[ix[0] * 6];
    const double* pose = &V.fam[c.fam[1]].members[ix[1] * 7];
    const double* xw = &V.fam[c.fam[2]].members[ix[2] * 3];
    eval_eucm(intr, pose, xw, ms, act[0], act[1], act[2], r, e.J[0], e.J[1], e.J[2]);
  } else if (c.kind == SGPU_COST_ISOMETRY3_PRIOR) {
    e.R = 6;
    eval_isometry3_prior(&V.fam[c.fam[0]].members[ix[0] * 7], ms, r, e.J[0]);
  } else {
    e.R = 6;
    const double* m = &V.fam[c.fam[0]].members[ix[0] * 7];
    const double* n = &V.fam[c.fam[1]].members[ix[1] * 7];
    eval_pose_graph(m, n, ms, r, e.J[0], e.J[1]);
  }
  for (int a = 0; a < 3; ++a) e.has[a] = act[a];
  // make (evaluated_term.rs:286-327): r~ = w(|r|) r; with a precision W the gradient is
  // J^T W r~, the Hessian J^T W J and the cost r~^T W r~
  e.W = c.kind == SGPU_COST_ISOMETRY3_PRIOR ? ms + 7 : nullptr;
  double nrm2 = 0;
  for (int i = 0; i < e.R; ++i) nrm2 += r[i] * r[i];
  const double w = c.robust == SGPU_ROBUST_HUBER ? robust_weight(c.robust, c.delta, std::sqrt(nrm2))
                                                 : 1.0;
  for (int i = 0; i < e.R; ++i) e.r[i] = c.robust == SGPU_ROBUST_HUBER ? w * r[i] : r[i];
  double cost = 0;
  if (e.W) {
    for (int i = 0; i < e.R; ++i) {
      double lr = 0;
      for (int k = 0; k < e.R; ++k) lr += e.W[e.R * i + k] * e.r[k];
      cost += e.r[i] * lr;
    }
  } else {
    for (int i = 0; i < e.R; ++i) cost += e.r[i] * e.r[i];
  }
  e.cost = cost;
}

// calc_merit / calc_total_cost (optimizer_helpers.rs:309-321, cost_fn.rs:403-433).
double merit(const Problem& P, const Vars& V) {
  constexpr int64_t kChunk = 1 << 16;
  std::vector<double> cost(kChunk);
  double total = 0;
  for (const auto& c : P.costs) {
    double s = 0;
    const int64_t nt = c.num_terms();
    for (int64_t t0 = 0; t0 < nt; t0 += kChunk) {
      const int64_t n = std::min(kChunk, nt - t0);
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
      for (int64_t k = 0; k < n; ++k) {
        EvalTerm e;
        eval_term(P, V, c, t0 + k, false, e);
        cost[k] = e.cost;
      }
      for (int64_t k = 0; k < n; ++k) s += cost[k];
    }
    total += s;
  }
  return total;
}

// ---- block-sparse lower matrix (pattern + values) ------------------------------------
struct BlockLower {
  int64_t nb = 0;
  std::vector<int64_t> dim, start;
  // CSR by block row: cols < = row, sorted
  std::vector<int64_t> row_ptr, col;
  std::vector<int64_t> val_off;  // offset of each block's values (row-major dim_r x dim_c)
  std::vector<double> val;
  int64_t find(int64_t r, int64_t c) const {
    auto b = col.begin() + row_ptr[r], e = col.begin() + row_ptr[r + 1];
    auto it = std::lower_bound(b, e, c);
    if (it == e || *it != c) return -1;
    return int64_t(it - col.begin());
  }
  double* block(int64_t k) { return &val[val_off[k]]; }
  const double* block(int64_t k) const { return &val[val_off[k]]; }
};

void build_pattern(BlockLower& M, const std::vector<std::set<int64_t>>& rows) {
  M.row_ptr.assign(M.nb + 1, 0);
  M.col.clear();
  M.val_off.clear();
  int64_t off = 0;
  for (int64_t r = 0; r < M.nb; ++r) {
    for (int64_t c : rows[r]) {
      M.col.push_back(c);
      M.val_off.push_back(off);
      off += M.dim[r] * M.dim[c];
    }
    M.row_ptr[r + 1] = int64_t(M.col.size());
  }
  M.val.assign(off, 0.0);
}

void add_block(BlockLower& M, int64_t r, int64_t c, const double* B /*row-major dr x dc*/) {
  int64_t k = M.find(r, c);
  double* d = M.block(k);
  const int64_t n = M.dim[r] * M.dim[c];
  for (int64_t i = 0; i < n; ++i) d[i] += B[i];
}

// Full-matrix entry (i, j) of block (r, c): lower storage, transpose when r < c.
double get_entry(const BlockLower& M, int64_t r, int64_t c, int i, int j) {
  if (r >= c) {
    int64_t k = M.find(r, c);
    if (k < 0) return 0.0;
    return M.block(k)[i * M.dim[c] + j];
  }
  int64_t k = M.find(c, r);
  if (k < 0) return 0.0;
  return M.block(k)[j * M.dim[r] + i];
}

// ---- linearization -------------------------------------------------------------------
struct Linearization {
  BlockLower H;                  // all active blocks (free, then marg)
  std::vector<double> neg_grad;  // scalar vector
};

// Symbolic pattern of H (LinearSystem::build_initial_pattern, linear_system.rs:90-105).
void symbolic(const Problem& P, Linearization& L) {
  const Vars& V = P.vars;
  L.H.nb = V.num_blocks;
  L.H.dim = V.block_dim;
  L.H.start = V.block_start;
  std::vector<std::set<int64_t>> rows(V.num_blocks);
  for (const auto& c : P.costs) {
    for (int64_t t = 0; t < c.num_terms(); ++t) {
      const int64_t* ix = &c.idx[t * c.arity];
      for (int a = 0; a < c.arity; ++a) {
        const Family& fa = V.fam[c.fam[a]];
        if (fa.kind == SGPU_VAR_CONDITIONED) continue;
        int64_t ba = fa.block_id[ix[a]];
        if (ba < 0) continue;
        rows[ba].insert(ba);
        for (int b = 0; b < c.arity; ++b) {
          if (b == a) continue;
          const Family& fb = V.fam[c.fam[b]];
          if (fb.kind == SGPU_VAR_CONDITIONED) continue;
          int64_t bb = fb.block_id[ix[b]];
          if (bb < 0) continue;
          if (V.block_start[bb] > V.block_start[ba]) continue;
          rows[ba].insert(bb);
        }
      }
    }
  }
  build_pattern(L.H, rows);
}

// cost_system.eval + populate_upper_triangular_normal_equation (sequential path):
// terms are evaluated, reduced over their reduction range, and accumulated with
// nu * I added per reduced term to each active diagonal block.
// Sum of the H blocks and gradients of the terms [t0, t1) of one reduction range.
struct RangeSums {
  double Hs[3][3][36];
  double gs[3][6];
};

void reduce_range(const Problem& P, const Vars& V, const CostFn& c, int64_t t0, int64_t t1,
                  const int dof[3], int R, RangeSums& out) {
  std::memset(&out, 0, sizeof(out));
  EvalTerm e;
  for (int64_t t = t0; t < t1; ++t) {
    eval_term(P, V, c, t, true, e);
    // with a precision: lambda r~ = W r~ and W J_b in place of r~ and J_b
    double lr[6], WJ[3][36];
    for (int k = 0; k < R; ++k) {
      lr[k] = e.r[k];
      if (e.W) {
        lr[k] = 0;
        for (int q = 0; q < R; ++q) lr[k] += e.W[R * k + q] * e.r[q];
      }
    }
    for (int b = 0; b < c.arity; ++b) {
      if (!e.has[b]) continue;
      for (int k = 0; k < R; ++k)
        for (int j = 0; j < dof[b]; ++j) {
          double acc = e.J[b][k * dof[b] + j];
          if (e.W) {
            acc = 0;
            for (int q = 0; q < R; ++q) acc += e.W[R * k + q] * e.J[b][q * dof[b] + j];
          }
          WJ[b][k * dof[b] + j] = acc;
        }
    }
    for (int a = 0; a < c.arity; ++a) {
      if (!e.has[a]) continue;
      for (int i = 0; i < dof[a]; ++i) {
        double acc = 0;
        for (int k = 0; k < R; ++k) acc += e.J[a][k * dof[a] + i] * lr[k];
        out.gs[a][i] += acc;
      }
      for (int b = 0; b < c.arity; ++b) {
        if (!e.has[b]) continue;
        for (int i = 0; i < dof[a]; ++i)
          for (int j = 0; j < dof[b]; ++j) {
            double acc = 0;
            for (int k = 0; k < R; ++k) acc += e.J[a][k * dof[a] + i] * WJ[b][k * dof[b] + j];
            out.Hs[a][b][i * dof[b] + j] += acc;
          }
      }
    }
  }
}

void linearize(const Problem& P, const Vars& V, double nu, Linearization& L) {
  std::fill(L.H.val.begin(), L.H.val.end(), 0.0);
  L.neg_grad.assign(V.num_free_scalars + V.num_marg_scalars, 0.0);
  constexpr int64_t kChunk = 1 << 10;
  std::vector<RangeSums> sums(kChunk);
  for (const auto& c : P.costs) {
    const int64_t nt = c.num_terms();
    const int R = is_reprojection(c.kind) ? 2 : 6;
    int dof[3] = {0, 0, 0};
    for (int a = 0; a < c.arity; ++a) dof[a] = V.fam[c.fam[a]].dof;
    const int64_t nr = int64_t(c.range_start.size());
    auto range_end = [&](int64_t rr) { return rr + 1 < nr ? c.range_start[rr + 1] : nt; };
    for (int64_t r0 = 0; r0 < nr; r0 += kChunk) {
      const int64_t n = std::min(kChunk, nr - r0);
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
      for (int64_t k = 0; k < n; ++k)
        reduce_range(P, V, c, c.range_start[r0 + k], range_end(r0 + k), dof, R, sums[k]);
      // accumulate in range order; H row ba and its gradient rows belong to one thread
      run_threads([&](int tid, int T) {
        for (int64_t k = 0; k < n; ++k) {
          const RangeSums& s = sums[k];
          const int64_t* ix = &c.idx[c.range_start[r0 + k] * c.arity];
          for (int a = 0; a < c.arity; ++a) {
            const Family& fa = V.fam[c.fam[a]];
            if (fa.kind == SGPU_VAR_CONDITIONED) continue;
            const int64_t sa = fa.scalar_start[ix[a]];
            if (sa < 0) continue;
            const int64_t ba = fa.block_id[ix[a]];
            if (ba % T != tid) continue;
            for (int i = 0; i < dof[a]; ++i) L.neg_grad[sa + i] -= s.gs[a][i];
            add_block(L.H, ba, ba, s.Hs[a][a]);
            if (nu != 0.0) {
              double ni[36] = {};
              for (int i = 0; i < dof[a]; ++i) ni[i * dof[a] + i] = nu;
              add_block(L.H, ba, ba, ni);
            }
            for (int b = 0; b < c.arity; ++b) {
              if (b == a) continue;
              const Family& fb = V.fam[c.fam[b]];
              if (fb.kind == SGPU_VAR_CONDITIONED) continue;
              const int64_t sb = fb.scalar_start[ix[b]];
              if (sb < 0 || sb > sa) continue;
              add_block(L.H, ba, fb.block_id[ix[b]], s.Hs[a][b]);
            }
          }
        }
      });
    }
  }
}

// ---- dense LDLt kernel (dense_ldlt.rs:164-256) ----------------------------------------
// A: n x n row-major (lower part read).  Out: unit-lower L (row-major), d.
bool ldlt_decompose(int n, const double* A, double* L, double* d, double tol_rel,
                    int* rank_out = nullptr, int* definiteness = nullptr,
                    double* pivot_condition = nullptr) {
  std::vector<double> a(A, A + n * n);
  double max_abs_pivot = 1.0, min_abs_pivot = DBL_MAX;
  int rank = 0;
  bool has_zero = false, has_neg = false;
  for (int j = 0; j < n; ++j) {
    for (int k = 0; k < j; ++k) {
      double s = a[j * n + k];
      for (int p = 0; p < k; ++p) s -= a[j * n + p] * d[p] * a[k * n + p];
      const double zero_tol = std::max(max_abs_pivot, 1.0) * tol_rel;
      a[j * n + k] = std::fabs(d[k]) <= zero_tol ? 0.0 : s / d[k];
    }
    double djj = a[j * n + j];
    for (int p = 0; p < j; ++p) {
      const double ljp = a[j * n + p];
      djj -= d[p] * ljp * ljp;
    }
    if (!std::isfinite(djj)) return false;
    const double zero_tol = std::max(max_abs_pivot, 1.0) * tol_rel;
    if (std::fabs(djj) <= zero_tol) {
      d[j] = 0.0;
      has_zero = true;
    } else {
      if (djj < 0.0) has_neg = true;
      d[j] = djj;
      max_abs_pivot = std::max(max_abs_pivot, std::fabs(djj));
      min_abs_pivot = std::min(min_abs_pivot, std::fabs(djj));
      ++rank;
    }
    a[j * n + j] = 1.0;
  }
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < n; ++c) L[r * n + c] = c > r ? 0.0 : a[r * n + c];
  if (rank_out) *rank_out = rank;
  if (definiteness) *definiteness = has_neg ? 2 : has_zero ? 1 : 0;
  if (pivot_condition) *pivot_condition = rank == 0 ? 0.0 : min_abs_pivot / max_abs_pivot;
  return true;
}

// kernel.rs solves.  y <- L^-1 y ; y <- D^-1 y (d == 0 -> 0) ; y <- L^-T y.
void ldlt_solve_vec(int n, const double* L, const double* d, double* y) {
  for (int i = 0; i < n; ++i) {
    double s = y[i];
    for (int k = 0; k < i; ++k) s -= L[i * n + k] * y[k];
    y[i] = s;
  }
  for (int i = 0; i < n; ++i) y[i] = d[i] == 0.0 ? 0.0 : y[i] / d[i];
  for (int i = n - 1; i >= 0; --i) {
    double s = y[i];
    for (int k = i + 1; k < n; ++k) s -= L[k * n + i] * y[k];
    y[i] = s;
  }
}

// ldlt_right_matsolve_inplace (dense_ldlt.rs:300-308): X <- X L^-T D^-1 L^-1, X m x n.
void ldlt_right_solve(int m, int n, const double* L, const double* d, double* X) {
  for (int r = 0; r < m; ++r) {  // lower_transpose_rightsolve_inplace (kernel.rs:130-153)
    for (int j = 0; j < n; ++j) {
      double s = X[r * n + j];
      for (int t = 0; t < j; ++t) s -= X[r * n + t] * L[j * n + t];
      X[r * n + j] = s;
    }
  }
  for (int j = 0; j < n; ++j) {  // diag_right_matsolve_inplace (kernel.rs:155-165)
    const double inv = d[j] == 0.0 ? 0.0 : 1.0 / d[j];
    for (int r = 0; r < m; ++r) X[r * n + j] *= inv;
  }
  for (int r = 0; r < m; ++r) {  // lower_right_solve_inplace (kernel.rs:110-127)
    for (int j = n - 1; j >= 0; --j) {
      double s = X[r * n + j];
      for (int t = j + 1; t < n; ++t) s -= X[r * n + t] * L[t * n + j];
      X[r * n + j] = s;
    }
  }
}

// ---- block-sparse LDLt (block_sparse_ldlt.rs) ------------------------------------------
// Minimum-degree block ordering (the oracle's stand-in for faer's AMD): repeatedly
// eliminate the node of least current degree (ties: lowest index), joining its
// neighbours into a clique.
std::vector<int64_t> min_degree_order(int64_t nb, const std::vector<std::set<int64_t>>& adj0) {
  std::vector<std::set<int64_t>> adj = adj0;
  std::vector<char> done(nb, 0);
  std::set<std::pair<int64_t, int64_t>> heap;
  for (int64_t i = 0; i < nb; ++i) heap.insert({int64_t(adj[i].size()), i});
  std::vector<int64_t> perm;
  perm.reserve(nb);
  while (!heap.empty()) {
    auto [deg, v] = *heap.begin();
    heap.erase(heap.begin());
    perm.push_back(v);
    done[v] = 1;
    std::vector<int64_t> nbrs(adj[v].begin(), adj[v].end());
    for (int64_t u : nbrs) {
      heap.erase({int64_t(adj[u].size()), u});
      adj[u].erase(v);
    }
    for (size_t i = 0; i < nbrs.size(); ++i)
      for (size_t j = i + 1; j < nbrs.size(); ++j) {
        adj[nbrs[i]].insert(nbrs[j]);
        adj[nbrs[j]].insert(nbrs[i]);
      }
    for (int64_t u : nbrs) heap.insert({int64_t(adj[u].size()), u});
    adj[v].clear();
  }
  return perm;  // perm[new] = old
}

// Ordering override: perm[new] = old over the RCS blocks; empty = minimum degree.
std::vector<int64_t> g_rcs_order;

// Pivot statistics of every block LDLt factorization since the last reset: pivots set to
// zero by the PSD rule, negative pivots, and the smallest |d| / max(1, max |d|) over the
// kept pivots of a diagonal block (how close the factorization came to the zero rule).
struct FactorStat {
  int64_t zero, neg;
  double min_ratio;
};
std::vector<FactorStat> g_factor_stats;

struct BlockLdltFactor {
  int64_t nb = 0;
  std::vector<int64_t> perm, dim, start;  // perm[new] = old; dim/start in new order
  std::vector<std::vector<int64_t>> lrows;                 // rows i > j of L(:, j)
  std::vector<std::vector<std::vector<double>>> lval;      // per column, per row block (di x dj)
  std::vector<std::vector<double>> Ljj, d;                 // per column
};

// factorize_impl (block_sparse_ldlt.rs:479-535) with the workspace of :554-747.
bool block_ldlt_factorize(const BlockLower& A, BlockLdltFactor& F, double tol_rel) {
  const int64_t nb = A.nb;
  F.nb = nb;
  std::vector<std::set<int64_t>> adj(nb);
  for (int64_t r = 0; r < nb; ++r)
    for (int64_t k = A.row_ptr[r]; k < A.row_ptr[r + 1]; ++k)
      if (A.col[k] != r) {
        adj[r].insert(A.col[k]);
        adj[A.col[k]].insert(r);
      }
  if (int64_t(g_rcs_order.size()) == nb)
    F.perm = g_rcs_order;
  else
    F.perm = min_degree_order(nb, adj);
  std::vector<int64_t> inv(nb);
  for (int64_t i = 0; i < nb; ++i) inv[F.perm[i]] = i;
  F.dim.resize(nb);
  F.start.resize(nb + 1);
  F.start[0] = 0;
  for (int64_t i = 0; i < nb; ++i) {
    F.dim[i] = A.dim[F.perm[i]];
    F.start[i + 1] = F.start[i] + F.dim[i];
  }
  // permuted lower matrix by column: Acol[j] = list of (i >= j, block i x j row-major)
  std::vector<std::vector<std::pair<int64_t, std::vector<double>>>> Acol(nb);
  for (int64_t r = 0; r < nb; ++r)
    for (int64_t k = A.row_ptr[r]; k < A.row_ptr[r + 1]; ++k) {
      const int64_t c = A.col[k];
      int64_t pi = inv[r], pj = inv[c];
      const double* B = A.block(k);
      const int64_t dr = A.dim[r], dc = A.dim[c];
      std::vector<double> blk(dr * dc);
      if (pi >= pj) {
        std::copy(B, B + dr * dc, blk.begin());
        Acol[pj].push_back({pi, blk});
      } else {  // transpose into (pj, pi)
        for (int64_t i = 0; i < dr; ++i)
          for (int64_t j = 0; j < dc; ++j) blk[j * dr + i] = B[i * dc + j];
        Acol[pi].push_back({pj, blk});
      }
    }
  // elimination tree on the upper pattern (elimination_tree.rs:20-60)
  std::vector<std::vector<int64_t>> upper(nb);  // upper[j] = rows i < j with A(j, i) != 0
  for (int64_t j = 0; j < nb; ++j)
    for (auto& e : Acol[j])
      if (e.first != j) upper[e.first].push_back(j);
  std::vector<int64_t> parent(nb, -1), ancestor(nb, -1);
  for (int64_t j = 0; j < nb; ++j)
    for (int64_t i0 : upper[j]) {
      int64_t i = i0;
      while (i != j) {
        int64_t next = ancestor[i];
        ancestor[i] = j;
        if (next < 0) {
          parent[i] = j;
          break;
        }
        i = next;
      }
    }
  std::vector<int64_t> stamp(nb, -1);
  auto reach = [&](int64_t j) {  // elimination_tree.rs:67-105
    // Nodes are pushed onto a stack that grows downwards from the end of the buffer,
    // so the returned slice lists them in reverse push order.
    stamp[j] = j;
    std::vector<int64_t> pushed;
    for (int64_t i0 : upper[j]) {
      int64_t i = i0;
      while (i >= 0 && stamp[i] != j) {
        pushed.push_back(i);
        stamp[i] = j;
        i = parent[i];
      }
    }
    std::reverse(pushed.begin(), pushed.end());
    return pushed;
  };
  F.lrows.assign(nb, {});
  F.lval.assign(nb, {});
  F.Ljj.assign(nb, {});
  F.d.assign(nb, {});
  std::vector<std::map<int64_t, std::vector<double>>> Lmap(nb);  // per column
  std::vector<int64_t> slot_of(nb, -1);  // row -> update target of the current column
  FactorStat stat{0, 0, 1.0};
  for (int64_t j = 0; j < nb; ++j) {
    const int64_t dj = F.dim[j];
    std::map<int64_t, std::vector<double>> C;  // row -> block (di x dj)
    for (auto& e : Acol[j]) {
      auto& b = C[e.first];
      if (b.empty()) b.assign(F.dim[e.first] * dj, 0.0);
      for (size_t t = 0; t < e.second.size(); ++t) b[t] += e.second[t];
    }
    if (!C.count(j)) C[j].assign(dj * dj, 0.0);
    // apply_to_col_k_in_reach (:614-720): every column k in the reach with L(j, k) != 0
    // updates C(j) and each C(i), i > j, with L(i, k) != 0, in reach order.  Each target
    // block's updates are listed in that order and the blocks are updated in parallel.
    struct Upd {
      int64_t q;                       // index into kcol / Qj
      const std::vector<double>* Lik;  // L(i, k) (null for the diagonal target)
      int64_t slot;                    // target block
    };
    std::vector<Upd> upd;
    std::vector<int64_t> kcol;            // listed columns k, in reach order
    std::vector<std::vector<double>> Qj;  // per listed k: Qjk = L(j, k) L_kk (dj x dk)
    std::vector<std::vector<double>*> targets{&C[j]};
    std::vector<int64_t> target_row{j};
    slot_of[j] = 0;
    for (int64_t k : reach(j)) {
      auto it = Lmap[k].find(j);
      if (it == Lmap[k].end()) continue;
      const int64_t dk = F.dim[k];
      const std::vector<double>& Ljk = it->second;  // dj x dk
      const std::vector<double>& Lkk = F.Ljj[k];
      std::vector<double> Qjk(dj * dk, 0.0);
      for (int64_t a = 0; a < dj; ++a)
        for (int64_t b = 0; b < dk; ++b) {
          double acc = 0;
          for (int64_t t = 0; t < dk; ++t) acc += Ljk[a * dk + t] * Lkk[t * dk + b];
          Qjk[a * dk + b] = acc;
        }
      const int64_t q = int64_t(kcol.size());
      kcol.push_back(k);
      Qj.push_back(std::move(Qjk));
      upd.push_back({q, nullptr, 0});
      for (auto& [i, Lik] : Lmap[k]) {
        if (i <= j) continue;
        if (slot_of[i] < 0) {
          auto& Cb = C[i];
          if (Cb.empty()) Cb.assign(F.dim[i] * dj, 0.0);
          slot_of[i] = int64_t(targets.size());
          targets.push_back(&Cb);
          target_row.push_back(i);
        }
        upd.push_back({q, &Lik, slot_of[i]});
      }
    }
    // group the updates by target block (counting sort keeps reach order in each group)
    const int64_t nt = int64_t(targets.size());
    std::vector<int64_t> gptr(nt + 1, 0), gidx(upd.size());
    for (const Upd& u : upd) ++gptr[u.slot + 1];
    for (int64_t t = 0; t < nt; ++t) gptr[t + 1] += gptr[t];
    {
      std::vector<int64_t> fill(gptr.begin(), gptr.end() - 1);
      for (size_t u = 0; u < upd.size(); ++u) gidx[fill[upd[u].slot]++] = int64_t(u);
    }
    for (int64_t r : target_row) slot_of[r] = -1;
#pragma omp parallel for schedule(dynamic, 4) num_threads(g_threads) if (g_threads > 1)
    for (int64_t t = 0; t < nt; ++t) {
      std::vector<double>& Cb = *targets[t];
      const int64_t di = int64_t(Cb.size()) / dj;
      std::vector<double> Qik;
      for (int64_t g = gptr[t]; g < gptr[t + 1]; ++g) {
        const Upd& u = upd[gidx[g]];
        const int64_t k = kcol[u.q], dk = F.dim[k];
        const std::vector<double>& dkv = F.d[k];
        const std::vector<double>& Qjk = Qj[u.q];
        const std::vector<double>* Q = &Qjk;
        if (u.Lik) {
          const std::vector<double>& Lik = *u.Lik;
          const std::vector<double>& Lkk = F.Ljj[k];
          Qik.assign(di * dk, 0.0);
          for (int64_t a = 0; a < di; ++a)
            for (int64_t b = 0; b < dk; ++b) {
              double acc = 0;
              for (int64_t s2 = 0; s2 < dk; ++s2) acc += Lik[a * dk + s2] * Lkk[s2 * dk + b];
              Qik[a * dk + b] = acc;
            }
          Q = &Qik;
        }
        // C -= Q diag(d) Qjk^T (sub_mat_diag_mat_t_inplace, kernel.rs:169-195)
        for (int64_t s2 = 0; s2 < dk; ++s2) {
          const double dt = dkv[s2];
          for (int64_t c2 = 0; c2 < dj; ++c2) {
            const double bj = dt * Qjk[c2 * dk + s2];
            for (int64_t r2 = 0; r2 < di; ++r2) Cb[r2 * dj + c2] -= (*Q)[r2 * dk + s2] * bj;
          }
        }
      }
    }
    // append_to_ldlt (:722-747)
    F.Ljj[j].assign(dj * dj, 0.0);
    F.d[j].assign(dj, 0.0);
    if (!ldlt_decompose(int(dj), C[j].data(), F.Ljj[j].data(), F.d[j].data(), tol_rel))
      return false;
    {
      double mx = 1.0;
      for (double v : F.d[j]) mx = std::max(mx, std::fabs(v));
      for (double v : F.d[j]) {
        if (v == 0.0) ++stat.zero;
        else stat.min_ratio = std::min(stat.min_ratio, std::fabs(v) / mx);
        if (v < 0.0) ++stat.neg;
      }
    }
    std::vector<std::pair<int64_t, std::vector<double>*>> below;
    for (auto& [i, Cb] : C)
      if (i > j) below.push_back({i, &Cb});
#pragma omp parallel for schedule(dynamic, 8) num_threads(g_threads) if (g_threads > 1)
    for (int64_t e = 0; e < int64_t(below.size()); ++e)
      ldlt_right_solve(int(F.dim[below[e].first]), int(dj), F.Ljj[j].data(), F.d[j].data(),
                       below[e].second->data());
    for (auto& [i, Cb] : below) Lmap[j][i] = std::move(*Cb);
  }
  for (int64_t j = 0; j < nb; ++j)
    for (auto& [i, b] : Lmap[j]) {
      F.lrows[j].push_back(i);
      F.lval[j].push_back(b);
    }
  g_factor_stats.push_back(stat);
  return true;
}

// solve_inplace (block_sparse_ldlt.rs:96-190)
void block_ldlt_solve(const BlockLdltFactor& F, const BlockLower& A, std::vector<double>& x) {
  const int64_t nb = F.nb;
  std::vector<double> y(F.start[nb]);
  for (int64_t i = 0; i < nb; ++i)
    for (int64_t t = 0; t < F.dim[i]; ++t) y[F.start[i] + t] = x[A.start[F.perm[i]] + t];
  for (int64_t j = 0; j < nb; ++j)
    for (size_t e = 0; e < F.lrows[j].size(); ++e) {
      const int64_t i = F.lrows[j][e];
      const auto& Lij = F.lval[j][e];
      for (int64_t r = 0; r < F.dim[i]; ++r) {
        double acc = 0;
        for (int64_t c = 0; c < F.dim[j]; ++c) acc += Lij[r * F.dim[j] + c] * y[F.start[j] + c];
        y[F.start[i] + r] -= acc;
      }
    }
  for (int64_t j = 0; j < nb; ++j) {  // per-block LDLt solve
    const int n = int(F.dim[j]);
    double* v = &y[F.start[j]];
    const double* L = F.Ljj[j].data();
    const double* d = F.d[j].data();
    ldlt_solve_vec(n, L, d, v);
  }
  for (int64_t j = nb - 1; j >= 0; --j)
    for (size_t e = 0; e < F.lrows[j].size(); ++e) {
      const int64_t i = F.lrows[j][e];
      const auto& Lij = F.lval[j][e];
      for (int64_t c = 0; c < F.dim[j]; ++c) {
        double acc = 0;
        for (int64_t r = 0; r < F.dim[i]; ++r) acc += Lij[r * F.dim[j] + c] * y[F.start[i] + r];
        y[F.start[j] + c] -= acc;
      }
    }
  for (int64_t i = 0; i < nb; ++i)
    for (int64_t t = 0; t < F.dim[i]; ++t) x[A.start[F.perm[i]] + t] = y[F.start[i] + t];
}

// ---- 3x3 inverse (nalgebra try_inverse closed form; det == 0 -> singular) -----------------
bool inverse3(const double* m, double* out) {
  const double m11 = m[0], m12 = m[1], m13 = m[2];
  const double m21 = m[3], m22 = m[4], m23 = m[5];
  const double m31 = m[6], m32 = m[7], m33 = m[8];
  const double minor_m12_m23 = m22 * m33 - m32 * m23;
  const double minor_m11_m23 = m21 * m33 - m31 * m23;
  const double minor_m11_m22 = m21 * m32 - m31 * m22;
  const double det = m11 * minor_m12_m23 - m12 * minor_m11_m23 + m13 * minor_m11_m22;
  if (det == 0.0) return false;
  out[0] = minor_m12_m23 / det;
  out[1] = (m13 * m32 - m33 * m12) / det;
  out[2] = (m12 * m23 - m22 * m13) / det;
  out[3] = -minor_m11_m23 / det;
  out[4] = (m11 * m33 - m31 * m13) / det;
  out[5] = (m13 * m21 - m23 * m11) / det;
  out[6] = minor_m11_m22 / det;
  out[7] = (m12 * m31 - m32 * m11) / det;
  out[8] = (m11 * m22 - m21 * m12) / det;
  return true;
}

// Dense inverse via LU with partial pivoting, for marg blocks of other sizes.
bool inverse_dense(int n, const double* m, double* out) {
  if (n == 3) return inverse3(m, out);
  std::vector<double> a(m, m + n * n);
  std::vector<double> inv(n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(a[r * n + c]) > std::fabs(a[p * n + c])) p = r;
    if (a[p * n + c] == 0.0) return false;
    for (int k = 0; k < n; ++k) {
      std::swap(a[c * n + k], a[p * n + k]);
      std::swap(inv[c * n + k], inv[p * n + k]);
    }
    const double piv = a[c * n + c];
    for (int k = 0; k < n; ++k) {
      a[c * n + k] /= piv;
      inv[c * n + k] /= piv;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = a[r * n + c];
      for (int k = 0; k < n; ++k) {
        a[r * n + k] -= f * a[c * n + k];
        inv[r * n + k] -= f * inv[c * n + k];
      }
    }
  }
  std::copy(inv.begin(), inv.end(), out);
  return true;
}

// ---- scalar sparse LDLt (sparse_ldlt.rs) ------------------------------------------------
// SparseLdlt::factorize_impl (:571-650): the block order (faer's AMD on the block graph
// there, :97-115; here the order block_ldlt_factorize uses) expanded to scalars
// (scalar_perm_from_block_perm, :121-160), then the left-looking scalar LDLt of the
// LdltWorkspace (:335-445): column j is loaded from the permuted lower matrix, every
// earlier column k with L(j, k) != 0 subtracts (L(j, k) d_k) L(i, k) for its rows i >= j,
// and the pivot is tested on its own, |d_jj| <= max(|d_jj|, 1) tol -> 0 with an empty
// column, else L(i, j) = c_i / d_jj (append_to_ldlt, :399-445).  The reference visits the
// columns k in the elimination-tree reach order; here they go in increasing k, which
// changes only the order of the subtractions.
struct ScalarLdltFactor {
  int64_t n = 0;
  std::vector<int64_t> perm;                                   // scalar perm[new] = old
  std::vector<std::vector<std::pair<int64_t, double>>> col;    // L(i, j), i > j ascending
  std::vector<double> d;
};

bool sparse_ldlt_factorize(const BlockLower& A, ScalarLdltFactor& F, double tol_rel) {
  const int64_t nb = A.nb;
  std::vector<int64_t> bperm;
  {
    std::vector<std::set<int64_t>> adj(nb);
    for (int64_t r = 0; r < nb; ++r)
      for (int64_t k = A.row_ptr[r]; k < A.row_ptr[r + 1]; ++k)
        if (A.col[k] != r) {
          adj[r].insert(A.col[k]);
          adj[A.col[k]].insert(r);
        }
    bperm = int64_t(g_rcs_order.size()) == nb ? g_rcs_order : min_degree_order(nb, adj);
  }
  const int64_t n = nb == 0 ? 0 : A.start[nb - 1] + A.dim[nb - 1];
  F.n = n;
  F.perm.clear();
  for (int64_t b : bperm)
    for (int64_t t = 0; t < A.dim[b]; ++t) F.perm.push_back(A.start[b] + t);
  std::vector<int64_t> inv(n);
  for (int64_t i = 0; i < n; ++i) inv[F.perm[i]] = i;
  // permuted lower matrix by column (permute_sparse_lower, :163-205): acol[j] = (i >= j, a)
  std::vector<std::vector<std::pair<int64_t, double>>> acol(n);
  for (int64_t r = 0; r < nb; ++r)
    for (int64_t k = A.row_ptr[r]; k < A.row_ptr[r + 1]; ++k) {
      const int64_t c = A.col[k], dr = A.dim[r], dc = A.dim[c];
      const double* B = A.block(k);
      for (int64_t i = 0; i < dr; ++i)
        for (int64_t j = 0; j < dc; ++j) {
          if (r == c && j > i) continue;  // the lower triangle of a diagonal block
          int64_t pi = inv[A.start[r] + i], pj = inv[A.start[c] + j];
          if (pi < pj) std::swap(pi, pj);
          acol[pj].push_back({pi, B[i * dc + j]});
        }
    }
  F.col.assign(n, {});
  F.d.assign(n, 0.0);
  std::vector<std::vector<std::pair<int64_t, double>>> lrow(n);  // (k, L(i, k)) per row i
  std::vector<double> c(n, 0.0);
  std::vector<char> touched(n, 0);
  std::vector<int64_t> rows;
  for (int64_t j = 0; j < n; ++j) {
    rows.clear();
    for (const auto& [i, v] : acol[j])  // load_column (:344-357)
      if (!touched[i]) {
        c[i] = v;
        touched[i] = 1;
        rows.push_back(i);
      }
    for (const auto& [k, ljk] : lrow[j]) {  // apply_to_col_k_in_reach (:359-397)
      const double ljk_dk = ljk * F.d[k];
      for (const auto& [i, lik] : F.col[k]) {
        if (i < j) continue;
        const double delta = ljk_dk * lik;
        if (!touched[i]) {
          c[i] = -delta;
          touched[i] = 1;
          rows.push_back(i);
        } else {
          c[i] -= delta;
        }
      }
    }
    const double djj_raw = touched[j] ? c[j] : 0.0;
    if (!std::isfinite(djj_raw)) return false;  // NonFinitePivot
    const double tau = std::max(std::fabs(djj_raw), 1.0) * tol_rel;
    const double djj = std::fabs(djj_raw) <= tau ? 0.0 : djj_raw;
    F.d[j] = djj;
    if (djj != 0.0) {
      std::sort(rows.begin(), rows.end());
      for (int64_t i : rows)
        if (i > j) {
          const double lij = c[i] / djj;
          if (lij != 0.0) {
            F.col[j].push_back({i, lij});
            lrow[i].push_back({j, lij});
          }
        }
    }
    for (int64_t i : rows) {
      touched[i] = 0;
      c[i] = 0.0;
    }
  }
  FactorStat stat{0, 0, DBL_MAX};
  for (double v : F.d) {
    if (v == 0.0) ++stat.zero;
    if (v < 0.0) ++stat.neg;
  }
  g_factor_stats.push_back(stat);
  return true;
}

// SparseLdltFactor::solve_inplace (:252-330): L y = P b, z = y / d (a zero pivot divides by
// zero there), L^T x = z, x <- P^T x.
void sparse_ldlt_solve(const ScalarLdltFactor& F, std::vector<double>& x) {
  const int64_t n = F.n;
  std::vector<double> b(n);
  for (int64_t i = 0; i < n; ++i) b[i] = x[F.perm[i]];
  for (int64_t j = 0; j < n; ++j)
    for (const auto& [i, l] : F.col[j]) b[i] -= l * b[j];
  for (int64_t i = 0; i < n; ++i) b[i] /= F.d[i];
  for (int64_t j = n - 1; j >= 0; --j)
    for (const auto& [i, l] : F.col[j]) b[j] -= l * b[i];
  for (int64_t i = 0; i < n; ++i) x[F.perm[i]] = b[i];
}

// ---- Schur complement solve (schur_ldlt.rs:82-182, :536-812) -------------------------------
// H: all active blocks; the first nfb blocks are free.  g: full rhs (neg_grad).
// Returns false on SingularMargBlock or a factorization failure.
// scalar_rule: the RCS is factored by SparseLdlt (SchurSparseLdlt) instead of
// BlockSparseLdlt (SchurBlockSparseLdlt).
bool schur_solve(const BlockLower& H, int64_t nfb, const std::vector<double>& g,
                 std::vector<double>& dx, double tol_rel = 1e-12, const Kkt* kkt = nullptr,
                 bool scalar_rule = false) {
  const int64_t nb = H.nb;
  const int64_t nfs = nfb == 0 ? 0 : H.start[nfb - 1] + H.dim[nfb - 1];
  // S pattern: H_ff pattern plus every (f_i, f_j) pair sharing a marg block.
  std::vector<std::set<int64_t>> srows(nfb);
  run_threads([&](int tid, int T) {
    for (int64_t r = tid; r < nfb; r += T) {
      for (int64_t k = H.row_ptr[r]; k < H.row_ptr[r + 1]; ++k) srows[r].insert(H.col[k]);
      srows[r].insert(r);
    }
    for (int64_t b = nfb; b < nb; ++b) {
      // free columns of marg row b (ascending); row fs[i] of S belongs to one thread
      const int64_t k0 = H.row_ptr[b], k1 = H.row_ptr[b + 1];
      for (int64_t ki = k0; ki < k1 && H.col[ki] < nfb; ++ki) {
        if (H.col[ki] % T != tid) continue;
        for (int64_t kj = k0; kj <= ki; ++kj) srows[H.col[ki]].insert(H.col[kj]);
      }
    }
  });
  BlockLower S;
  S.nb = nfb;
  S.dim.assign(H.dim.begin(), H.dim.begin() + nfb);
  S.start.assign(H.start.begin(), H.start.begin() + nfb);
  build_pattern(S, srows);
  // S_ff := H_ff
  for (int64_t r = 0; r < nfb; ++r)
    for (int64_t k = H.row_ptr[r]; k < H.row_ptr[r + 1]; ++k) add_block(S, r, H.col[k], H.block(k));
  std::vector<double> rhs(g.begin(), g.begin() + nfs);
  std::vector<std::vector<double>> hinv(nb - nfb);
  // forward pass (schur_forward_pass_sequential, :241-278), over chunks of marg blocks:
  // each block's H_mm^-1, H_mm^-1 g_m and H_mm^-1 H_mf are formed in parallel, then
  // applied to S and the rhs in block order (apply_contrib_to_sparse, :138-182).
  constexpr int64_t kChunk = 1 << 10;
  std::vector<int64_t> off(kChunk + 1);  // per block: start of [hig | w_0 | w_1 ...]
  std::vector<double> work;
  std::vector<char> singular(kChunk);
  for (int64_t b0 = nfb; b0 < nb; b0 += kChunk) {
    const int64_t n = std::min(kChunk, nb - b0);
    off[0] = 0;
    for (int64_t q = 0; q < n; ++q) {
      const int64_t b = b0 + q, dm = H.dim[b];
      int64_t len = dm;
      for (int64_t k = H.row_ptr[b]; k < H.row_ptr[b + 1] && H.col[k] < nfb; ++k)
        len += dm * H.dim[H.col[k]];
      off[q + 1] = off[q] + len;
    }
    if (int64_t(work.size()) < off[n]) work.resize(off[n]);
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (int64_t q = 0; q < n; ++q) {
      const int64_t b = b0 + q, dm = H.dim[b];
      int64_t kd = H.find(b, b);
      double hmm[36] = {};
      if (kd >= 0) std::copy(H.block(kd), H.block(kd) + dm * dm, hmm);
      std::vector<double>& hi = hinv[b - nfb];
      hi.assign(dm * dm, 0.0);
      singular[q] = !inverse_dense(int(dm), hmm, hi.data());
      if (singular[q]) continue;
      double* hig = &work[off[q]];
      for (int64_t i = 0; i < dm; ++i) {
        double acc = 0;
        for (int64_t k = 0; k < dm; ++k) acc += hi[i * dm + k] * g[H.start[b] + k];
        hig[i] = acc;
      }
      double* wb = hig + dm;
      for (int64_t k = H.row_ptr[b]; k < H.row_ptr[b + 1] && H.col[k] < nfb; ++k) {
        const int64_t df = H.dim[H.col[k]];
        for (int64_t i = 0; i < dm; ++i)
          for (int64_t j = 0; j < df; ++j) {
            double acc = 0;
            for (int64_t t = 0; t < dm; ++t) acc += hi[i * dm + t] * H.block(k)[t * df + j];
            wb[i * df + j] = acc;
          }
        wb += dm * df;
      }
    }
    for (int64_t q = 0; q < n; ++q)
      if (singular[q]) return false;
    run_threads([&](int tid, int T) {
      for (int64_t q = 0; q < n; ++q) {
        const int64_t b = b0 + q, dm = H.dim[b];
        const double* hig = &work[off[q]];
        const int64_t k0 = H.row_ptr[b];
        for (int64_t ki = k0; ki < H.row_ptr[b + 1] && H.col[ki] < nfb; ++ki) {
          const int64_t fi = H.col[ki], di = H.dim[fi];
          const double* hmf_i = H.block(ki);  // dm x di
          if (fi % T == tid) {
            for (int64_t c = 0; c < di; ++c) {
              double acc = 0;
              for (int64_t t = 0; t < dm; ++t) acc += hmf_i[t * di + c] * hig[t];
              rhs[H.start[fi] + c] -= acc;
            }
            const double* wj = hig + dm;
            for (int64_t kj = k0; kj <= ki; ++kj) {
              const int64_t fj = H.col[kj], dj = H.dim[fj];
              double blk[36];
              for (int64_t r = 0; r < di; ++r)
                for (int64_t c = 0; c < dj; ++c) {
                  double acc = 0;
                  for (int64_t t = 0; t < dm; ++t) acc += hmf_i[t * di + r] * wj[t * dj + c];
                  blk[r * dj + c] = -acc;
                }
              add_block(S, fi, fj, blk);
              wj += dm * dj;
            }
          }
        }
      }
    });
  }
  BlockLdltFactor F;
  ScalarLdltFactor Fs;
  if (scalar_rule ? !sparse_ldlt_factorize(S, Fs, tol_rel) : !block_ldlt_factorize(S, F, tol_rel))
    return false;
  auto solve_s = [&](std::vector<double>& v) {
    if (scalar_rule)
      sparse_ldlt_solve(Fs, v);
    else
      block_ldlt_solve(F, S, v);
  };
  std::vector<double> dxf = rhs;
  std::vector<double> dlam;
  if (kkt && kkt->nc > 0) {
    // range-space KKT (schur_ldlt.rs:450-480, 674-730, 762-772): Y = S^-1 G^T, M = G Y,
    // u = S^-1 rhs, dlam = M^-1 (G u - rhs_c), dx_f = S^-1 (rhs - G^T dlam)
    const int64_t nc = kkt->nc;
    std::vector<std::vector<double>> Y(nc);
    for (int64_t k = 0; k < nc; ++k) {
      Y[k].assign(kkt->G.begin() + k * nfs, kkt->G.begin() + (k + 1) * nfs);
      solve_s(Y[k]);
    }
    std::vector<double> M(size_t(nc * nc)), Mi(size_t(nc * nc));
    for (int64_t a = 0; a < nc; ++a)
      for (int64_t b = 0; b < nc; ++b) {
        double acc = 0;
        for (int64_t i = 0; i < nfs; ++i) acc += kkt->G[size_t(a * nfs + i)] * Y[b][i];
        M[size_t(a * nc + b)] = acc;
      }
    if (!inverse_dense(int(nc), M.data(), Mi.data())) return false;  // SingularKktConstraint
    std::vector<double> u = rhs;
    solve_s(u);
    std::vector<double> w(nc);
    for (int64_t a = 0; a < nc; ++a) {
      double acc = 0;
      for (int64_t i = 0; i < nfs; ++i) acc += kkt->G[size_t(a * nfs + i)] * u[i];
      w[a] = acc - kkt->rhs_c[a];
    }
    dlam.assign(nc, 0.0);
    for (int64_t a = 0; a < nc; ++a)
      for (int64_t b = 0; b < nc; ++b) dlam[a] += Mi[size_t(a * nc + b)] * w[b];
    for (int64_t i = 0; i < nfs; ++i)
      for (int64_t a = 0; a < nc; ++a) dxf[i] -= kkt->G[size_t(a * nfs + i)] * dlam[a];
  }
  solve_s(dxf);
  dx.assign(g.size() + dlam.size(), 0.0);
  std::copy(dxf.begin(), dxf.end(), dx.begin());
  std::copy(dlam.begin(), dlam.end(), dx.begin() + g.size());
  // back-substitution (:785-806)
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
  for (int64_t b = nfb; b < nb; ++b) {
    const int64_t dm = H.dim[b];
    double mv[6];
    std::copy(g.begin() + H.start[b], g.begin() + H.start[b] + dm, mv);
    for (int64_t k = H.row_ptr[b]; k < H.row_ptr[b + 1]; ++k) {
      const int64_t f = H.col[k];
      if (f >= nfb) continue;
      const int64_t df = H.dim[f];
      for (int64_t i = 0; i < dm; ++i) {
        double acc = 0;
        for (int64_t j = 0; j < df; ++j) acc += H.block(k)[i * df + j] * dxf[H.start[f] + j];
        mv[i] -= acc;
      }
    }
    const auto& hi = hinv[b - nfb];
    for (int64_t i = 0; i < dm; ++i) {
      double acc = 0;
      for (int64_t k = 0; k < dm; ++k) acc += hi[i * dm + k] * mv[k];
      dx[H.start[b] + i] = acc;
    }
  }
  return true;
}

// ---- problem construction / validation ----------------------------------------------------
sgpu_status build_problem(const sgpu_var_family* fams, int64_t nf, const sgpu_cost_fn* cfs,
                          int64_t nc, Problem& P, const sgpu_eq_fn* eqs = nullptr,
                          int64_t neq = 0) {
  if (!fams || nf <= 0) return fail(SGPU_ERR_INVALID_ARGUMENT, "no variable families");
  for (int64_t i = 0; i < nf; ++i) {
    const auto& s = fams[i];
    if (!s.name || s.type < 0 || s.type > 3 || s.kind < 0 || s.kind > 2 || s.num_members < 0 ||
        (s.num_members > 0 && !s.members))
      return fail(SGPU_ERR_INVALID_ARGUMENT, "bad variable family");
    Family f;
    f.name = s.name;
    f.type = s.type;
    f.kind = s.kind;
    f.dof = var_dof(s.type);
    f.nparams = var_params(s.type);
    f.n = s.num_members;
    f.members.assign(s.members, s.members + f.n * f.nparams);
    f.is_const.assign(f.n, 0);
    for (int64_t k = 0; k < s.num_const_ids; ++k) {
      if (s.const_ids[k] < 0 || s.const_ids[k] >= f.n)
        return fail(SGPU_ERR_INDEX_OUT_OF_RANGE, "const id out of range");
      f.is_const[s.const_ids[k]] = 1;
    }
    for (auto& g : P.vars.fam)
      if (g.name == f.name) return fail(SGPU_ERR_INVALID_ARGUMENT, "duplicate family name");
    P.vars.fam.push_back(std::move(f));
  }
  std::sort(P.vars.fam.begin(), P.vars.fam.end(),
            [](const Family& a, const Family& b) { return a.name < b.name; });
  layout(P.vars);
  // validate_solver_config (optimizer.rs:649-717)
  if (P.vars.num_free_scalars == 0 || P.vars.num_marg_scalars == 0)
    return fail(SGPU_ERR_SCHUR_MISSING_FREE_OR_MARG,
                "Schur solver requires at least one active free and one active marginalized "
                "variable");
  for (int64_t i = 0; i < nc; ++i) {
    const auto& s = cfs[i];
    const int arity = is_reprojection(s.kind) ? 3 : s.kind == SGPU_COST_ISOMETRY3_PRIOR ? 1 : 2;
    int marg = 0;
    for (int a = 0; a < arity; ++a) {
      if (!s.family_names[a]) return fail(SGPU_ERR_INVALID_ARGUMENT, "missing family name");
      int fi = P.vars.find(s.family_names[a]);
      if (fi >= 0 && P.vars.fam[fi].kind == SGPU_VAR_MARGINALIZED) ++marg;
    }
    if (marg > 1)
      return fail(SGPU_ERR_SCHUR_COST_CONNECTS_TWO_MARG_FAMILIES,
                  "cost term connects two marginalized families");
  }
  for (int64_t i = 0; i < nc; ++i) {
    const auto& s = cfs[i];
    if (!is_reprojection(s.kind) && s.kind != SGPU_COST_POSE_GRAPH3 && s.kind != SGPU_COST_ISOMETRY3_PRIOR)
      return fail(SGPU_ERR_INVALID_ARGUMENT, "unknown cost kind");
    CostFn c;
    c.kind = s.kind;
    c.robust = s.robust_kernel;
    c.delta = s.huber_delta;
    c.arity = is_reprojection(s.kind) ? 3 : s.kind == SGPU_COST_ISOMETRY3_PRIOR ? 1 : 2;
    const int want[4][3] = {{SGPU_VAR_PINHOLE, SGPU_VAR_ISOMETRY3, SGPU_VAR_VEC3},
                            {SGPU_VAR_ISOMETRY3, SGPU_VAR_ISOMETRY3, -1},
                            {SGPU_VAR_EUCM, SGPU_VAR_ISOMETRY3, SGPU_VAR_VEC3},
                            {SGPU_VAR_ISOMETRY3, -1, -1}};
    for (int a = 0; a < c.arity; ++a) {
      c.fam[a] = P.vars.find(s.family_names[a]);
      if (c.fam[a] < 0) return fail(SGPU_ERR_UNKNOWN_FAMILY, "unknown family");
      if (P.vars.fam[c.fam[a]].type != want[s.kind][a])
        return fail(SGPU_ERR_FAMILY_TYPE, "family type mismatch");
    }
    if (s.num_terms <= 0 || !s.entity_indices || !s.measurements)
      return fail(SGPU_ERR_INVALID_ARGUMENT, "empty cost function");
    const int64_t nt = s.num_terms;
    for (int64_t t = 0; t < nt; ++t)
      for (int a = 0; a < c.arity; ++a) {
        const int64_t v = s.entity_indices[t * c.arity + a];
        if (v < 0 || v >= P.vars.fam[c.fam[a]].n)
          return fail(SGPU_ERR_INDEX_OUT_OF_RANGE, "entity index out of range");
      }
    // CostFn::sort (cost_fn.rs:354-393): stable sort by (marg, free, conditioned) args
    int perm_keys[3];
    int nk = 0;
    for (int kind : {SGPU_VAR_MARGINALIZED, SGPU_VAR_FREE, SGPU_VAR_CONDITIONED})
      for (int a = 0; a < c.arity; ++a)
        if (P.vars.fam[c.fam[a]].kind == kind) perm_keys[nk++] = a;
    std::vector<int64_t> order(nt);
    for (int64_t t = 0; t < nt; ++t) order[t] = t;
    const int64_t* E = s.entity_indices;
    const int ar = c.arity;
    std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) {
      for (int k = 0; k < nk; ++k) {
        const int64_t a = E[x * ar + perm_keys[k]], b = E[y * ar + perm_keys[k]];
        if (a != b) return a < b;
      }
      return false;
    });
    const int ml = c.mlen();
    c.idx.resize(nt * ar);
    c.meas.resize(nt * ml);
    for (int64_t t = 0; t < nt; ++t) {
      for (int a = 0; a < ar; ++a) c.idx[t * ar + a] = E[order[t] * ar + a];
      for (int k = 0; k < ml; ++k) c.meas[t * ml + k] = s.measurements[order[t] * ml + k];
    }
    for (int64_t t = 0; t < nt; ++t) {
      bool same = t > 0;
      if (t > 0)
        for (int a = 0; a < ar; ++a)
          if (P.vars.fam[c.fam[a]].kind != SGPU_VAR_CONDITIONED &&
              c.idx[t * ar + a] != c.idx[c.range_start.back() * ar + a])
            same = false;
      if (!same) c.range_start.push_back(t);
    }
    P.costs.push_back(std::move(c));
  }
  if (P.costs.empty()) return fail(SGPU_ERR_INVALID_ARGUMENT, "no cost functions");
  // equality-constraint sets (EqSystem::new, eq_system.rs:24-48; validate_solver_config,
  // optimizer.rs:649-683)
  for (int64_t i = 0; i < neq; ++i) {
    const sgpu_eq_fn& q = eqs[i];
    if (q.kind != SGPU_EQ_TRANSLATION_NORM || !q.family_names[0] || q.num_terms < 0 ||
        (q.num_terms > 0 && (!q.entity_indices || !q.params)))
      return fail(SGPU_ERR_INVALID_ARGUMENT, "bad equality-constraint set");
    if (q.num_terms == 0) continue;  // residual_dim 0 sets add nothing
    EqFn e;
    e.fam = P.vars.find(q.family_names[0]);
    if (e.fam < 0) return fail(SGPU_ERR_UNKNOWN_FAMILY, "unknown constraint family");
    const Family& f = P.vars.fam[e.fam];
    if (f.kind == SGPU_VAR_MARGINALIZED)
      return fail(SGPU_ERR_SCHUR_CONSTRAINT_ON_MARGINALIZED_FAMILY,
                  "equality constraint on a marginalized family");
    if (f.type != SGPU_VAR_ISOMETRY3) return fail(SGPU_ERR_FAMILY_TYPE, "constraint family type");
    std::vector<int64_t> order(q.num_terms);
    for (int64_t t = 0; t < q.num_terms; ++t) {
      if (q.entity_indices[t] < 0 || q.entity_indices[t] >= f.n)
        return fail(SGPU_ERR_INDEX_OUT_OF_RANGE, "constraint entity index out of range");
      order[t] = t;
    }
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
      return q.entity_indices[a] < q.entity_indices[b];
    });
    for (int64_t t : order) {
      e.idx.push_back(q.entity_indices[t]);
      e.radius.push_back(q.params[t]);
    }
    P.eqs.push_back(std::move(e));
  }
  return SGPU_OK;
}

struct StepRecord {
  double merit, lm_damping, new_merit, gain_ratio, max_abs_delta;
  int32_t accepted, termination;
};

struct Optimizer {
  Problem P;
  sgpu_opt_params params;
  double merit_v = 0, nu = 0;
  int64_t rejects = 0, steps = 0;
  bool terminated = false;
  Linearization L;
  std::vector<double> lambda;  // Lagrange multipliers, one per constraint set (zero at start)
};

// step_nlls_inner (optimizer_step.rs:366-443) + Optimizer::step (optimizer.rs:345-377)
int step(Optimizer& O, StepRecord& rec) {
  Problem& P = O.P;
  const Vars& V = P.vars;
  rec.new_merit = NAN;
  rec.gain_ratio = NAN;
  rec.accepted = 0;
  rec.max_abs_delta = NAN;
  int reason = SGPU_TERM_NONE;
  linearize(P, V, O.nu, O.L);
  // equality constraints: G^T lambda in the free rows, -c below (eq_system.rs:207-262)
  Kkt K;
  std::vector<double> full_grad;
  if (!P.eqs.empty()) {
    O.lambda.resize(P.eqs.size(), 0.0);
    eval_kkt(P, V, O.lambda, K, &O.L.neg_grad);
  }
  full_grad = O.L.neg_grad;
  full_grad.insert(full_grad.end(), K.rhs_c.begin(), K.rhs_c.end());
  double maxg = 0;
  for (double v : O.L.neg_grad) maxg = std::max(maxg, std::fabs(v));
  bool done = false;
  // check_gradient_tolerance_unless_eq (optimizer.rs:527-535)
  if (P.eqs.empty() && maxg < O.params.gradient_tolerance) {
    reason = SGPU_TERM_GRADIENT_TOLERANCE;
    done = true;
  }
  std::vector<double> dx;
  if (!done && !schur_solve(O.L.H, V.num_free_blocks, O.L.neg_grad, dx, 1e-12, P.eqs.empty() ? nullptr : &K,
                            O.params.solver == SGPU_SOLVER_SCHUR_SPARSE_LDLT)) {
    O.nu *= 2.0;  // handle_solve_failure -> reject_step
    ++O.rejects;
    if (O.rejects >= O.params.max_consecutive_rejects) reason = SGPU_TERM_MAX_CONSECUTIVE_REJECTS;
    done = true;
  }
  if (!done) {
    Vars U = update(V, dx);
    // update_lambdas (eq_system.rs:149-159): lambda - delta's multiplier rows
    std::vector<double> lam_new = O.lambda;
    for (size_t k = 0; k < lam_new.size(); ++k) lam_new[k] -= dx[O.L.neg_grad.size() + k];
    const double nm = merit(P, U);
    rec.new_merit = nm;
    double max_step = 0;
    for (double v : dx) max_step = std::max(max_step, std::fabs(v));
    rec.max_abs_delta = max_step;
    if (max_step < O.params.parameter_tolerance) {
      if (nm < O.merit_v) {
        P.vars = U;
        O.lambda = lam_new;
        O.merit_v = nm;
        rec.accepted = 1;
      }
      reason = SGPU_TERM_PARAMETER_TOLERANCE;
    } else {
      // compute_gain_ratio (optimizer_helpers.rs:159-173), over the whole step and the whole
      // right-hand side (multiplier rows included)
      double dd = 0, dg = 0;
      for (size_t i = 0; i < dx.size(); ++i) {
        dd += dx[i] * dx[i];
        dg += dx[i] * full_grad[i];
      }
      const double pred = O.nu * dd + dg;
      const double actual = O.merit_v - nm;
      const double rho = std::fabs(pred) > DBL_MIN ? actual / pred : (actual > 0 ? 1.0 : 0.0);
      rec.gain_ratio = rho;
      if (rho > 0.0 && nm < O.merit_v) {
        const double f = 1.0 - std::pow(2.0 * rho - 1.0, 3);
        O.nu *= std::max(f, 1.0 / 3.0);
        O.rejects = 0;
        const double rel = (O.merit_v - nm) / (std::fabs(O.merit_v) + DBL_MIN);
        if (rel < O.params.function_tolerance) reason = SGPU_TERM_FUNCTION_TOLERANCE;
        P.vars = U;
        O.lambda = lam_new;
        O.merit_v = nm;
        rec.accepted = 1;
      } else {
        O.nu *= 2.0;
        ++O.rejects;
        if (O.rejects >= O.params.max_consecutive_rejects)
          reason = SGPU_TERM_MAX_CONSECUTIVE_REJECTS;
      }
    }
  }
  ++O.steps;
  if (O.steps >= O.params.num_iterations && reason == SGPU_TERM_NONE)
    reason = SGPU_TERM_MAX_ITERATIONS;
  rec.merit = O.merit_v;
  rec.lm_damping = O.nu;
  rec.termination = reason;
  return reason;
}

}  // namespace oracle

using namespace oracle;

namespace {
double g_step_seconds = 0.0;  // wall time spent inside step() by the last optimize call
}

extern "C" {

const char* oracle_last_error(void) { return g_error.c_str(); }

// Sets the RCS block ordering used by every later factorization (perm[new] = old over the
// free blocks, in the layout order of the variables); n = 0 restores minimum degree.
// Returns 1 if perm is not a permutation of 0..n-1.
int32_t oracle_set_rcs_order(const int64_t* perm, int64_t n) {
  std::vector<char> seen(size_t(std::max<int64_t>(n, 0)), 0);
  for (int64_t i = 0; i < n; ++i) {
    if (perm[i] < 0 || perm[i] >= n || seen[perm[i]]) return 1;
    seen[perm[i]] = 1;
  }
  g_rcs_order.assign(perm, perm + n);
  return 0;
}

// Host threads for the oracle's per-term and per-point loops (n < 1 -> 1); results do not
// depend on it.  Returns the previous value.
int32_t oracle_set_threads(int32_t n) {
  const int32_t prev = g_threads;
  g_threads = std::max<int32_t>(1, n);
  return prev;
}

// Pivot statistics of the block LDLt factorizations since the last reset (one entry per
// factorization, oldest first); returns how many were written (<= capacity).
int64_t oracle_factor_stats(int64_t* zero, int64_t* neg, double* min_ratio, int64_t capacity) {
  const int64_t n = std::min<int64_t>(capacity, int64_t(g_factor_stats.size()));
  for (int64_t i = 0; i < n; ++i) {
    if (zero) zero[i] = g_factor_stats[i].zero;
    if (neg) neg[i] = g_factor_stats[i].neg;
    if (min_ratio) min_ratio[i] = g_factor_stats[i].min_ratio;
  }
  return n;
}

void oracle_reset_factor_stats(void) { g_factor_stats.clear(); }

double oracle_last_step_seconds(void) { return g_step_seconds; }

typedef struct oracle_step_record {
  double merit, lm_damping, new_merit, gain_ratio, max_abs_delta;
  int32_t accepted, termination;
} oracle_step_record;

// optimize_nlls (nlls.rs:73-104): create, step until termination, build_solution.
// trace (may be NULL) receives up to trace_capacity step records.
sgpu_status oracle_optimize_nlls_eq(const sgpu_var_family* fams, int64_t nf,
                                    const sgpu_cost_fn* cfs, int64_t nc, const sgpu_eq_fn* eqs,
                                    int64_t neq, const sgpu_opt_params* params,
                                    double* const* out_members, sgpu_solution* sol,
                                    oracle_step_record* trace, int64_t trace_capacity,
                                    int64_t* trace_len, double* lambda_out);

sgpu_status oracle_optimize_nlls(const sgpu_var_family* fams, int64_t nf, const sgpu_cost_fn* cfs,
                                 int64_t nc, const sgpu_opt_params* params,
                                 double* const* out_members, sgpu_solution* sol,
                                 oracle_step_record* trace, int64_t trace_capacity,
                                 int64_t* trace_len) {
  return oracle_optimize_nlls_eq(fams, nf, cfs, nc, nullptr, 0, params, out_members, sol, trace,
                                 trace_capacity, trace_len, nullptr);
}

// optimize_nlls with equality-constraint sets (the reference's eq_fns argument); lambda_out
// (may be NULL) receives the final multipliers, one per non-empty set.
sgpu_status oracle_optimize_nlls_eq(const sgpu_var_family* fams, int64_t nf,
                                    const sgpu_cost_fn* cfs, int64_t nc, const sgpu_eq_fn* eqs,
                                    int64_t neq, const sgpu_opt_params* params,
                                    double* const* out_members, sgpu_solution* sol,
                                    oracle_step_record* trace, int64_t trace_capacity,
                                    int64_t* trace_len, double* lambda_out) {
  g_error.clear();
  if (!params || !sol) return fail(SGPU_ERR_INVALID_ARGUMENT, "params and solution required");
  Optimizer O;
  O.params = *params;
  sgpu_status st = build_problem(fams, nf, cfs, nc, O.P, eqs, neq);
  if (st != SGPU_OK) return st;
  symbolic(O.P, O.L);
  O.nu = params->initial_lm_damping;
  O.merit_v = merit(O.P, O.P.vars);
  sol->initial_cost = O.merit_v;
  int64_t n = 0;
  int reason = SGPU_TERM_NONE;
  g_step_seconds = 0.0;
  while (reason == SGPU_TERM_NONE) {
    StepRecord rec;
    const auto t0 = std::chrono::steady_clock::now();
    reason = step(O, rec);
    g_step_seconds +=
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (trace && n < trace_capacity) {
      trace[n] = {rec.merit, rec.lm_damping, rec.new_merit, rec.gain_ratio, rec.max_abs_delta,
                  rec.accepted, rec.termination};
    }
    ++n;
  }
  if (trace_len) *trace_len = n;
  sol->final_cost = O.merit_v;
  sol->num_steps = O.steps;
  sol->termination = reason;
  // build_solution: one more linearization with nu = 0 unless skip_final_hessian
  linearize(O.P, O.P.vars, params->skip_final_hessian ? O.nu : 0.0, O.L);
  double maxg = 0;
  for (double v : O.L.neg_grad) maxg = std::max(maxg, std::fabs(v));
  sol->final_max_abs_neg_grad = maxg;
  if (out_members) {
    for (int64_t i = 0; i < nf; ++i) {
      if (!out_members[i]) continue;
      const int fi = O.P.vars.find(fams[i].name);
      const auto& f = O.P.vars.fam[fi];
      std::memcpy(out_members[i], f.members.data(), f.members.size() * sizeof(double));
    }
  }
  if (lambda_out)
    for (size_t k = 0; k < O.P.eqs.size(); ++k) lambda_out[k] = k < O.lambda.size() ? O.lambda[k] : 0.0;
  return SGPU_OK;
}

// One equality-constrained linear step at the given values (damping nu, multipliers lambda):
// G (nc x nfs) and -c of the constraint rows, the neg-gradient with G^T lambda added, and the
// range-space Schur solve's [dx_f; dx_m; dlambda] (status 1 if the solve failed).
int32_t oracle_kkt_step(const sgpu_var_family* fams, int64_t nf, const sgpu_cost_fn* cfs,
                        int64_t nc, const sgpu_eq_fn* eqs, int64_t neq, double nu,
                        const double* lambda, double* G, double* rhs_c, double* neg_grad,
                        double* dx) {
  g_error.clear();
  Problem P;
  if (build_problem(fams, nf, cfs, nc, P, eqs, neq) != SGPU_OK) return 2;
  Linearization L;
  symbolic(P, L);
  linearize(P, P.vars, nu, L);
  std::vector<double> lam(lambda, lambda + P.eqs.size());
  Kkt K;
  eval_kkt(P, P.vars, lam, K, &L.neg_grad);
  std::copy(K.G.begin(), K.G.end(), G);
  std::copy(K.rhs_c.begin(), K.rhs_c.end(), rhs_c);
  std::copy(L.neg_grad.begin(), L.neg_grad.end(), neg_grad);
  std::vector<double> d;
  if (!schur_solve(L.H, P.vars.num_free_blocks, L.neg_grad, d, 1e-12, &K)) return 1;
  std::copy(d.begin(), d.end(), dx);
  return 0;
}

// Linearization at the given values with damping nu, as a dense matrix over the
// active scalars (free then marg) and the neg-gradient.  For small problems.
sgpu_status oracle_linearize_dense(const sgpu_var_family* fams, int64_t nf,
                                   const sgpu_cost_fn* cfs, int64_t nc, double nu, double* H,
                                   double* neg_grad, int64_t* dim) {
  g_error.clear();
  Problem P;
  sgpu_status st = build_problem(fams, nf, cfs, nc, P);
  if (st != SGPU_OK) return st;
  Linearization L;
  symbolic(P, L);
  linearize(P, P.vars, nu, L);
  const int64_t n = P.vars.num_free_scalars + P.vars.num_marg_scalars;
  *dim = n;
  if (!H) return SGPU_OK;
  for (int64_t r = 0; r < L.H.nb; ++r)
    for (int64_t c = 0; c < L.H.nb; ++c)
      for (int i = 0; i < L.H.dim[r]; ++i)
        for (int j = 0; j < L.H.dim[c]; ++j)
          H[(L.H.start[r] + i) * n + L.H.start[c] + j] = get_entry(L.H, r, c, i, j);
  std::copy(L.neg_grad.begin(), L.neg_grad.end(), neg_grad);
  return SGPU_OK;
}

// One LM step's linear solve at the given values: Schur + block LDLt of the damped
// system (the dx of step_nlls_inner).  Returns 1 on solver failure.
sgpu_status oracle_schur_step(const sgpu_var_family* fams, int64_t nf, const sgpu_cost_fn* cfs,
                              int64_t nc, double nu, double* dx, int32_t* failed) {
  g_error.clear();
  Problem P;
  sgpu_status st = build_problem(fams, nf, cfs, nc, P);
  if (st != SGPU_OK) return st;
  Linearization L;
  symbolic(P, L);
  linearize(P, P.vars, nu, L);
  std::vector<double> d;
  *failed = schur_solve(L.H, P.vars.num_free_blocks, L.neg_grad, d) ? 0 : 1;
  if (!*failed) std::copy(d.begin(), d.end(), dx);
  return SGPU_OK;
}

double oracle_merit(const sgpu_var_family* fams, int64_t nf, const sgpu_cost_fn* cfs, int64_t nc) {
  Problem P;
  if (build_problem(fams, nf, cfs, nc, P) != SGPU_OK) return NAN;
  return merit(P, P.vars);
}

// Schur solve of a dense symmetric matrix partitioned into blocks (dims[nb]); the first
// nfb blocks are free, the rest marginalized (each marg block must only couple to free
// blocks).  SchurFactor::factorize + solve.  Returns 1 in *failed on failure.
void oracle_schur_solve_dense(int64_t nb, const int64_t* dims, int64_t nfb, const double* Hd,
                              const double* g, double* dx, int32_t* failed) {
  BlockLower H;
  H.nb = nb;
  H.dim.assign(dims, dims + nb);
  H.start.resize(nb);
  int64_t n = 0;
  for (int64_t i = 0; i < nb; ++i) {
    H.start[i] = n;
    n += dims[i];
  }
  std::vector<std::set<int64_t>> rows(nb);
  for (int64_t r = 0; r < nb; ++r)
    for (int64_t c = 0; c <= r; ++c) {
      bool nz = false;
      for (int64_t i = 0; i < dims[r]; ++i)
        for (int64_t j = 0; j < dims[c]; ++j)
          if (Hd[(H.start[r] + i) * n + H.start[c] + j] != 0.0) nz = true;
      if (nz || r == c) rows[r].insert(c);
    }
  build_pattern(H, rows);
  for (int64_t r = 0; r < nb; ++r)
    for (int64_t k = H.row_ptr[r]; k < H.row_ptr[r + 1]; ++k) {
      const int64_t c = H.col[k];
      double* B = H.block(k);
      for (int64_t i = 0; i < dims[r]; ++i)
        for (int64_t j = 0; j < dims[c]; ++j) B[i * dims[c] + j] = Hd[(H.start[r] + i) * n + H.start[c] + j];
    }
  std::vector<double> gv(g, g + n), d;
  *failed = schur_solve(H, nfb, gv, d) ? 0 : 1;
  if (!*failed) std::copy(d.begin(), d.end(), dx);
}

// Block-sparse LDLt solve (factorize_impl + solve_inplace) of a symmetric matrix given
// dense (n x n) with the block partition dims[nb]; structurally zero blocks are skipped.
// scalar_rule: SparseLdlt's factorize + solve_inplace on the same input instead.
static int32_t direct_solve_dense(int64_t nb, const int64_t* dims, const double* Ad,
                                  const double* b, double* x, bool scalar_rule) {
  BlockLower A;
  A.nb = nb;
  A.dim.assign(dims, dims + nb);
  A.start.resize(nb);
  int64_t n = 0;
  for (int64_t i = 0; i < nb; ++i) {
    A.start[i] = n;
    n += dims[i];
  }
  std::vector<std::set<int64_t>> rows(nb);
  for (int64_t r = 0; r < nb; ++r)
    for (int64_t c = 0; c <= r; ++c) {
      bool nz = r == c;
      for (int64_t i = 0; i < dims[r]; ++i)
        for (int64_t j = 0; j < dims[c]; ++j)
          if (Ad[(A.start[r] + i) * n + A.start[c] + j] != 0.0) nz = true;
      if (nz) rows[r].insert(c);
    }
  build_pattern(A, rows);
  for (int64_t r = 0; r < nb; ++r)
    for (int64_t k = A.row_ptr[r]; k < A.row_ptr[r + 1]; ++k) {
      const int64_t c = A.col[k];
      for (int64_t i = 0; i < dims[r]; ++i)
        for (int64_t j = 0; j < dims[c]; ++j)
          A.block(k)[i * dims[c] + j] = Ad[(A.start[r] + i) * n + A.start[c] + j];
    }
  std::vector<double> v(b, b + n);
  if (scalar_rule) {
    ScalarLdltFactor F;
    if (!sparse_ldlt_factorize(A, F, 1e-12)) return 1;
    sparse_ldlt_solve(F, v);
  } else {
    BlockLdltFactor F;
    if (!block_ldlt_factorize(A, F, 1e-12)) return 1;
    block_ldlt_solve(F, A, v);
  }
  std::copy(v.begin(), v.end(), x);
  return 0;
}

int32_t oracle_block_ldlt_solve_dense(int64_t nb, const int64_t* dims, const double* Ad,
                                      const double* b, double* x) {
  return direct_solve_dense(nb, dims, Ad, b, x, false);
}

int32_t oracle_sparse_ldlt_solve_dense(int64_t nb, const int64_t* dims, const double* Ad,
                                       const double* b, double* x) {
  return direct_solve_dense(nb, dims, Ad, b, x, true);
}

// ldlt_decompose_inplace on a dense n x n matrix (row-major).  definiteness: 0 PD,
// 1 PSD, 2 indefinite.  Returns 1 on a non-finite pivot.
int32_t oracle_dense_ldlt(int32_t n, const double* A, double tol_rel, double* L, double* d,
                          int32_t* rank, int32_t* definiteness, double* pivot_condition) {
  int r = 0, df = 0;
  double pc = 0;
  if (!ldlt_decompose(n, A, L, d, tol_rel, &r, &df, &pc)) return 1;
  *rank = r;
  *definiteness = df;
  *pivot_condition = pc;
  return 0;
}

void oracle_dense_ldlt_solve(int32_t n, const double* L, const double* d, double* y) {
  ldlt_solve_vec(n, L, d, y);
}

// Per-term evaluation of pinhole reprojection terms (analytic Jacobians, FP64).
void oracle_eval_pinhole_terms(int64_t n, const double* intr, const double* pose,
                               const double* pt, const double* uv, int32_t robust, double delta,
                               double* res, double* jpose, double* jpt, double* cost,
                               double* weight) {
  for (int64_t t = 0; t < n; ++t) {
    double r[2], Jc[8], Jp[12], Jx[6];
    eval_pinhole(intr + 4 * t, pose + 7 * t, pt + 3 * t, uv + 2 * t, false, true, true, r, Jc, Jp,
                 Jx);
    const double nrm = std::sqrt(r[0] * r[0] + r[1] * r[1]);
    const double w = robust == SGPU_ROBUST_HUBER ? robust_weight(robust, delta, nrm) : 1.0;
    const double r0 = robust == SGPU_ROBUST_HUBER ? w * r[0] : r[0];
    const double r1 = robust == SGPU_ROBUST_HUBER ? w * r[1] : r[1];
    if (res) {
      res[2 * t] = r[0];
      res[2 * t + 1] = r[1];
    }
    if (jpose) std::memcpy(jpose + 12 * t, Jp, sizeof(Jp));
    if (jpt) std::memcpy(jpt + 6 * t, Jx, sizeof(Jx));
    if (cost) cost[t] = r0 * r0 + r1 * r1;
    if (weight) weight[t] = w;
  }
}

// Forward-mode dual-number Jacobians of the generic residual (pinhole_reprojection.rs:42-50)
// with respect to the left-multiplicative pose tangent at 0 and the point.
void oracle_pinhole_dual_jacobians(int64_t n, const double* intr, const double* pose,
                                   const double* pt, const double* uv, double* jpose,
                                   double* jpt) {
  using D = Dual<9>;
  for (int64_t t = 0; t < n; ++t) {
    D tangent[6], xw[3], in[4], T[7], Td[7], E[7], r[2];
    for (int k = 0; k < 6; ++k) tangent[k] = D::var(0.0, k);
    for (int k = 0; k < 3; ++k) xw[k] = D::var(pt[3 * t + k], 6 + k);
    for (int k = 0; k < 4; ++k) in[k] = D(intr[4 * t + k]);
    for (int k = 0; k < 7; ++k) T[k] = D(pose[7 * t + k]);
    se3_exp(tangent, E);
    se3_mul(E, T, Td);
    pinhole_residual(in, Td, xw, uv + 2 * t, r);
    for (int i = 0; i < 2; ++i) {
      for (int k = 0; k < 6; ++k) jpose[12 * t + 6 * i + k] = r[i].d[k];
      for (int k = 0; k < 3; ++k) jpt[6 * t + 3 * i + k] = r[i].d[6 + k];
    }
  }
}

// Per-term evaluation of EUCM reprojection terms (analytic Jacobians, FP64).
void oracle_eval_eucm_terms(int64_t n, const double* intr, const double* pose, const double* pt,
                            const double* uv, int32_t robust, double delta, double* res,
                            double* jintr, double* jpose, double* jpt, double* cost,
                            double* weight) {
  for (int64_t t = 0; t < n; ++t) {
    double r[2], Jc[12], Jp[12], Jx[6];
    eval_eucm(intr + 6 * t, pose + 7 * t, pt + 3 * t, uv + 2 * t, true, true, true, r, Jc, Jp, Jx);
    const double nrm = std::sqrt(r[0] * r[0] + r[1] * r[1]);
    const double w = robust == SGPU_ROBUST_HUBER ? robust_weight(robust, delta, nrm) : 1.0;
    const double r0 = robust == SGPU_ROBUST_HUBER ? w * r[0] : r[0];
    const double r1 = robust == SGPU_ROBUST_HUBER ? w * r[1] : r[1];
    if (res) {
      res[2 * t] = r[0];
      res[2 * t + 1] = r[1];
    }
    if (jintr) std::memcpy(jintr + 12 * t, Jc, sizeof(Jc));
    if (jpose) std::memcpy(jpose + 12 * t, Jp, sizeof(Jp));
    if (jpt) std::memcpy(jpt + 6 * t, Jx, sizeof(Jx));
    if (cost) cost[t] = r0 * r0 + r1 * r1;
    if (weight) weight[t] = w;
  }
}

// Forward-mode dual-number Jacobians of the generic EUCM residual (ba_problem.rs:58-66)
// with respect to the intrinsics, the left-multiplicative pose tangent at 0 and the point.
void oracle_eucm_dual_jacobians(int64_t n, const double* intr, const double* pose,
                                const double* pt, const double* uv, double* jintr, double* jpose,
                                double* jpt) {
  using D = Dual<15>;
  for (int64_t t = 0; t < n; ++t) {
    D in[6], tangent[6], xw[3], T[7], Td[7], E[7], r[2];
    for (int k = 0; k < 6; ++k) in[k] = D::var(intr[6 * t + k], k);
    for (int k = 0; k < 6; ++k) tangent[k] = D::var(0.0, 6 + k);
    for (int k = 0; k < 3; ++k) xw[k] = D::var(pt[3 * t + k], 12 + k);
    for (int k = 0; k < 7; ++k) T[k] = D(pose[7 * t + k]);
    se3_exp(tangent, E);
    se3_mul(E, T, Td);
    eucm_residual(in, Td, xw, uv + 2 * t, r);
    for (int i = 0; i < 2; ++i) {
      for (int k = 0; k < 6; ++k) jintr[12 * t + 6 * i + k] = r[i].d[k];
      for (int k = 0; k < 6; ++k) jpose[12 * t + 6 * i + k] = r[i].d[6 + k];
      for (int k = 0; k < 3; ++k) jpt[6 * t + 3 * i + k] = r[i].d[12 + k];
    }
  }
}

// EUCM distort / undistort of n points on the z = 1 plane (enhanced_unified.rs:60-124).
void oracle_eucm_distort(int64_t n, const double* intr, const double* xy, double* uv) {
  for (int64_t t = 0; t < n; ++t) eucm_distort(intr + 6 * t, xy + 2 * t, uv + 2 * t);
}
void oracle_eucm_undistort(int64_t n, const double* intr, const double* uv, double* xy) {
  for (int64_t t = 0; t < n; ++t) eucm_undistort(intr + 6 * t, uv + 2 * t, xy + 2 * t);
}

void oracle_eval_pose_graph3_terms(int64_t n, const double* m, const double* nn,
                                   const double* mn, double* res, double* jm, double* jn) {
  for (int64_t t = 0; t < n; ++t)
    eval_pose_graph(m + 7 * t, nn + 7 * t, mn + 7 * t, res + 6 * t, jm + 36 * t, jn + 36 * t);
}

// Dual-number Jacobians of the pose-graph residual w.r.t. left perturbations of m and n.
// Isometry3PriorCostTerm terms: the residual (6 per term), the Jacobian (36, row-major) by
// duals and, with the precision W (36 per term) and Huber delta (robust 1), the cost.
void oracle_eval_isometry3_prior_terms(int64_t n, const double* poses, const double* means,
                                       const double* precisions, int32_t robust, double delta,
                                       double* res, double* jac, double* cost) {
  for (int64_t t = 0; t < n; ++t) {
    double r[6];
    eval_isometry3_prior(poses + 7 * t, means + 7 * t, r, jac + 36 * t);
    double nrm2 = 0;
    for (int k = 0; k < 6; ++k) nrm2 += r[k] * r[k];
    const double w = robust == SGPU_ROBUST_HUBER ? robust_weight(robust, delta, std::sqrt(nrm2)) : 1.0;
    double rw[6], c = 0;
    for (int k = 0; k < 6; ++k) {
      res[6 * t + k] = r[k];
      rw[k] = robust == SGPU_ROBUST_HUBER ? w * r[k] : r[k];
    }
    const double* W = precisions + 36 * t;
    for (int i = 0; i < 6; ++i) {
      double lr = 0;
      for (int k = 0; k < 6; ++k) lr += W[6 * i + k] * rw[k];
      c += rw[i] * lr;
    }
    cost[t] = c;
  }
}

void oracle_pose_graph3_dual_jacobians(int64_t n, const double* m, const double* nn,
                                       const double* mn, double* jm, double* jn) {
  using D = Dual<12>;
  for (int64_t t = 0; t < n; ++t) {
    D tm[6], tn[6], M[7], N[7], MN[7], Em[7], En[7], Mp[7], Np[7], r[6];
    for (int k = 0; k < 6; ++k) {
      tm[k] = D::var(0.0, k);
      tn[k] = D::var(0.0, 6 + k);
    }
    for (int k = 0; k < 7; ++k) {
      M[k] = D(m[7 * t + k]);
      N[k] = D(nn[7 * t + k]);
      MN[k] = D(mn[7 * t + k]);
    }
    se3_exp(tm, Em);
    se3_exp(tn, En);
    se3_mul(Em, M, Mp);
    se3_mul(En, N, Np);
    pose_graph3_residual(Mp, Np, MN, r);
    for (int i = 0; i < 6; ++i)
      for (int k = 0; k < 6; ++k) {
        jm[36 * t + 6 * i + k] = r[i].d[k];
        jn[36 * t + 6 * i + k] = r[i].d[6 + k];
      }
  }
}

void oracle_so3_exp(const double* w, double* q) { so3_exp(w, q); }
void oracle_so3_log(const double* q, double* w) { so3_log(q, w); }
void oracle_so3_matrix(const double* q, double* R) { so3_matrix(q, R); }
void oracle_se3_exp(const double* t, double* T) { se3_exp(t, T); }
void oracle_se3_log(const double* T, double* t) { se3_log(T, t); }
void oracle_se3_mul(const double* a, const double* b, double* out) { se3_mul(a, b, out); }
void oracle_se3_inverse(const double* a, double* out) { se3_inverse(a, out); }
void oracle_so3_mat_v(const double* w, double* V) { so3_mat_v(w, V); }
void oracle_so3_mat_v_inverse(const double* w, double* V) { so3_mat_v_inverse(w, V); }
void oracle_se3_inv_left_jacobian(const double* t, double* J) { se3_inv_left_jacobian(t, J); }
double oracle_huber_weight(double delta, double nrm) {
  return robust_weight(SGPU_ROBUST_HUBER, delta, nrm);
}
int32_t oracle_inverse3(const double* m, double* out) { return inverse3(m, out) ? 0 : 1; }

}  // extern "C"
