// C ABI: Optimizer::new / step / build_solution and optimize_nlls on the GPU.
//
// The LM control (optimizer_step.rs:366-443, optimizer.rs:345-377,463-520,
// optimizer_helpers.rs:159-173) runs on the host in FP64 from a handful of scalars the
// device reduces per step; everything that touches observations, points or the reduced
// camera system runs in the kernels of kernels.hip on the handle's HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "api_internal.hpp"
#include "kernels.hpp"
#include "learned.hpp"
#include "optimizer_internal.hpp"
#include "pcg.hpp"
#include "problem.hpp"
#include "symbolic.hpp"

namespace sgpu {

namespace {
thread_local std::string g_last_error;
}

void clear_error() { g_last_error.clear(); }
sgpu_status set_error(sgpu_status status, const std::string& message) {
  g_last_error = message;
  return status;
}

namespace {

enum KernelId {
  K_LINEARIZE_POINTS = 0,
  K_CAMERA_ROWS,
  K_FACTOR,
  K_TRIANGULAR,
  K_BACKSUB,
  K_UPDATE,
  K_COST,
  K_MF_STEP,  // k_mf_step launches, per-launch events (timing mode 2 only)
  K_COUNT
};
const char* kKernelNames[K_COUNT] = {"point_prep",  "camera_rows", "rcs_factor", "rcs_solve",
                                     "backsub",     "pose_update", "cost",       "mf_step"};

double bits_to_double(unsigned long long b) {
  double d;
  std::memcpy(&d, &b, sizeof(d));
  return d;
}


}  // namespace
}  // namespace sgpu

using namespace sgpu;

struct sgpu_optimizer {
  HostProblem P;
  Symbolic S;
  sgpu_opt_params params{};
  int device = 0;
  hipStream_t stream = nullptr;
  DeviceArena arena;
  DevProblem dp{};
  DevFactor df{};
  DevLin dl{};
  std::vector<DevPoseGraph> dpg;
  double* poses[2] = {nullptr, nullptr};
  double* points[2] = {nullptr, nullptr};
  int cur = 0;
  // intrinsics members (both copies equal unless the intrinsics are free)
  double* intr[2] = {nullptr, nullptr};
  double* cost_partial = nullptr;
  DevScalars* d_sc = nullptr;
  DevScalars* h_sc = nullptr;  // pinned
  DevPcg pcg{};                // iterative RCS solver state (PCG solvers only)
  PcgState* h_pcg = nullptr;   // pinned poll buffer
  DevLearned learned{};        // learned block-Jacobi correction (PCG_LEARNED only)
  DevEq deq{};                 // equality-constraint sets (none: nset = 0)
  double merit = 0, nu = 0;
  int64_t rejects = 0, steps = 0;
  double grad0 = -1.0, grad_last = -1.0;  // max |g| of the first and the last linearization read
  bool terminated = false;
  int32_t last_termination = SGPU_TERM_NONE;
  // timing: 0 off, 1 per phase, 2 also per k_mf_step launch
  int timing = 0;
  std::vector<hipEvent_t> step_ev;
  hipEvent_t ev[2 * K_COUNT] = {};
  double kernel_ms[K_COUNT] = {};
  int64_t kernel_count[K_COUNT] = {};
  hipEvent_t phase_ev[5] = {};
  // Inverse-block queries: the final linearization (build_solution's nu) factored, S's null
  // vectors (from the zeroed pivots, orthonormal, camera order) and the query buffers.
  bool cov_ready = false;
  std::vector<std::vector<double>> cov_null;
  double *cov_b = nullptr, *cov_bperm = nullptr, *cov_z = nullptr, *cov_x = nullptr;
  double *cov_coup = nullptr, *cov_hinv = nullptr;
  int32_t* cov_cam = nullptr;
  int32_t cov_max_obs = 0;
  // smallest kept pivot of the factored S relative to the diagonal entry of S in its column
  // (scale-invariant: a column nearly in the span of the ones eliminated before it has a
  // pivot far below its diagonal entry)
  double cov_pivot_ratio = 0.0;
  double* cov_sdiag = nullptr;

  ~sgpu_optimizer() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : phase_ev)
      if (e) (void)hipEventDestroy(e);
    if (h_sc) (void)hipHostFree(h_sc);
    if (h_pcg) (void)hipHostFree(h_pcg);
    for (auto e : step_ev)
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void tic(int k) {
    if (timing) SGPU_HIP(hipEventRecord(ev[2 * k], stream));
  }
  void toc(int k) {
    if (timing) SGPU_HIP(hipEventRecord(ev[2 * k + 1], stream));
  }
  // accumulate after a stream sync; kernels recorded this step only
  void collect(const bool* used) {
    if (!timing) return;
    if (timing == 2 && used[K_FACTOR] && df.step_events) {
      for (size_t li = 0; li < S.launches.size(); ++li) {
        if (S.launches[li].tile_count == 0) continue;
        float ms = 0;
        SGPU_HIP(hipEventElapsedTime(&ms, step_ev[2 * li], step_ev[2 * li + 1]));
        kernel_ms[K_MF_STEP] += ms;
        kernel_count[K_MF_STEP] += 1;
      }
    }
    for (int k = 0; k < K_COUNT; ++k) {
      if (!used[k] || k == K_MF_STEP) continue;
      float ms = 0;
      SGPU_HIP(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
      kernel_ms[k] += ms;
      kernel_count[k] += 1;
    }
  }

  bool iterative() const {
    return params.solver == SGPU_SOLVER_SCHUR_PCG_JACOBI || params.solver == SGPU_SOLVER_SCHUR_PCG_LEARNED;
  }

  const Family& pose_fam() const { return P.fams[P.f_pose]; }
  const Family& point_fam() const { return P.fams[P.f_point]; }

  void linearize(const double* ps, const double* pts, const double* in, double nu_lin,
                 bool* used) {
    cov_ready = false;  // the fronts and point records are rewritten
    launch_zero_scalars(stream, d_sc);
    launch_cam_cache(stream, S.npose, ps, dl.cw);
    if (!iterative()) launch_zero_fronts(stream, df);
    else if (!dpg.empty()) launch_fill(stream, dl.bsr, int64_t(36) * dp.nslot, 0.0);
    tic(K_LINEARIZE_POINTS);
    launch_point_prep(stream, params.precision, dp, ps, pts, in, dl.cw, nu_lin, dl, d_sc);
    toc(K_LINEARIZE_POINTS);
    tic(K_CAMERA_ROWS);
    launch_camera_rows(stream, params.precision, dp, pts, in, dl.cw, nu_lin, dl, df, d_sc);
    launch_intrinsics_rows(stream, dp, pts, in, dl.cw, dl, df);
    toc(K_CAMERA_ROWS);
    const double* centered = params.precision == SGPU_PRECISION_MIXED ? dl.cw : nullptr;
    for (auto& g : dpg) launch_pose_graph_linearize(stream, g, dp, ps, nu_lin, dl, df, centered);
    // the multipliers of the variables' buffer (the variables linearized are the current ones)
    launch_eq_eval(stream, dp, deq, ps, deq.lambda[cur], dl, true);
    launch_max_grad(stream, dl.grad_cam, S.nb, d_sc);
    used[K_LINEARIZE_POINTS] = used[K_CAMERA_ROWS] = true;
  }

  void cost_of(const double* ps, const double* pts, const double* in, bool* used) {
    tic(K_COST);
    launch_cost(stream, dp, ps, pts, in, cost_partial, 2 * kCostBlocks, d_sc);
    toc(K_COST);
    int nblk = kCostBlocks;
    for (auto& g : dpg) {
      launch_pose_graph_cost(stream, g, ps, cost_partial + 2 * nblk, 2 * kCostBlocks, d_sc);
      nblk += kCostBlocks;
    }
    launch_finalize_cost(stream, cost_partial, nblk, d_sc);
    if (used) used[K_COST] = true;
  }

  double compute_merit(const double* ps, const double* pts, const double* in) {
    bool used[K_COUNT] = {};
    launch_zero_scalars(stream, d_sc);
    cost_of(ps, pts, in, used);
    SGPU_HIP(hipMemcpyAsync(h_sc, d_sc, sizeof(DevScalars), hipMemcpyDeviceToHost, stream));
    SGPU_HIP(hipStreamSynchronize(stream));
    collect(used);
    return h_sc->cost;
  }

  void init(const sgpu_var_family* families, int64_t nf, const sgpu_cost_fn* cfs, int64_t nc,
            const sgpu_eq_fn* eqs, int64_t neq, const sgpu_opt_params* prm) {
    params = *prm;
    if (params.solver != SGPU_SOLVER_SCHUR_BLOCK_SPARSE_LDLT && params.solver != SGPU_SOLVER_SCHUR_SPARSE_LDLT &&
        params.solver != SGPU_SOLVER_SCHUR_PCG_JACOBI && params.solver != SGPU_SOLVER_SCHUR_PCG_LEARNED)
      throw Error(SGPU_ERR_UNSUPPORTED, "solver not available");
    if (iterative() && !(params.pcg_relative_tolerance > 0.0 && params.pcg_max_iterations > 0))
      throw Error(SGPU_ERR_INVALID_ARGUMENT, "PCG needs a positive tolerance and iteration cap");
    if (!(params.pcg_forcing >= 0.0 && params.pcg_forcing < 1.0))
      throw Error(SGPU_ERR_INVALID_ARGUMENT, "pcg_forcing must be in [0, 1)");
    if (params.solver == SGPU_SOLVER_SCHUR_PCG_LEARNED && !params.precond_weights)
      throw Error(SGPU_ERR_INVALID_ARGUMENT, "the learned preconditioner needs precond_weights");
    const float* precond_weights = params.precond_weights;
    params.precond_weights = nullptr;  // copied below; the caller's array is not kept
    if (params.precision != SGPU_PRECISION_F64 && params.precision != SGPU_PRECISION_MIXED)
      throw Error(SGPU_ERR_INVALID_ARGUMENT, "unknown precision");
    // validation (and, with host_symbolic, the term sort) on the host, before any device call
    P = build_host_problem_sorted(families, nf, cfs, nc, eqs, neq, params.host_symbolic != 0);
    if (!P.eqs.empty() && (iterative() || params.precision != SGPU_PRECISION_F64))
      throw Error(SGPU_ERR_UNSUPPORTED,
                  "equality constraints are implemented for the direct RCS solver in FP64 (the "
                  "range-space KKT solve needs S^-1 G^T from the factorization)");
    if (P.intr_block >= 0 && params.precision != SGPU_PRECISION_F64)
      throw Error(SGPU_ERR_UNSUPPORTED,
                  "free intrinsics are implemented in the FP64 precision (the mixed path's centred "
                  "camera tangents do not cover the intrinsics block)");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
      throw Error(SGPU_ERR_NO_DEVICE, "no HIP device");
    device = params.device;
    if (device < 0 || device >= ndev) throw Error(SGPU_ERR_NO_DEVICE, "device ordinal out of range");
    SGPU_HIP(hipSetDevice(device));
    SGPU_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    // symbolic analysis: the term sort, the S pattern and the pair lists on the device (the
    // pairs stay there), or all of it on the host
    int32_t* d_pairs = nullptr;
    if (params.host_symbolic) {
      S = analyze(P);
    } else {
      sort_terms_on_device(P, stream);
      S = analyze_device(P, stream, arena, &d_pairs);
    }
    SGPU_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_sc), sizeof(DevScalars), hipHostMallocDefault));
    for (auto& e : ev) SGPU_HIP(hipEventCreate(&e));
    for (auto& e : phase_ev) SGPU_HIP(hipEventCreate(&e));
    // variables
    const Family& pf = pose_fam();
    const Family& xf = point_fam();
    const Family& kf = P.fams[P.f_intr];
    for (int i = 0; i < 2; ++i) {
      poses[i] = arena.alloc<double>(pf.members.size());
      points[i] = arena.alloc<double>(xf.members.size());
    }
    SGPU_HIP(hipMemcpy(poses[0], pf.members.data(), pf.members.size() * 8, hipMemcpyHostToDevice));
    SGPU_HIP(hipMemcpy(points[0], xf.members.data(), xf.members.size() * 8, hipMemcpyHostToDevice));
    for (int i = 0; i < 2; ++i) intr[i] = arena.upload(kf.members, stream);
    // problem arrays
    std::vector<int32_t> pt_block(xf.block.begin(), xf.block.end());
    dp.nb = S.nb;
    dp.npt = S.npt;
    dp.nobs = S.nobs;
    dp.npose = S.npose;
    dp.nintr = int32_t(kf.n);
    dp.obs_pose = arena.upload(P.obs_pose, stream);
    dp.obs_point = arena.upload(P.obs_point, stream);
    dp.obs_intr = arena.upload(P.obs_intr, stream);
    dp.obs_uv = arena.upload(P.obs_uv, stream);
    dp.pt_ptr = arena.upload(S.pt_ptr, stream);
    dp.pt_nu = arena.upload(S.pt_nu, stream);
    dp.pt_block = arena.upload(pt_block, stream);
    dp.cam_nu = arena.upload(S.cam_nu, stream);
    dp.pose_block = arena.upload(S.pose_block, stream);
    dp.cam_member = arena.upload(P.cam_member, stream);
    dp.cam_ptr = arena.upload(S.cam_ptr, stream);
    dp.cam_obs = arena.upload(S.cam_obs, stream);
    {
      const size_t ncm = S.cam_obs.size();
      std::vector<int32_t> cm_point(ncm), cm_intr(ncm);
      std::vector<double> cm_uv(2 * ncm);
      for (size_t q = 0; q < ncm; ++q) {
        const int32_t o = S.cam_obs[q];
        cm_point[q] = P.obs_point[o];
        cm_intr[q] = P.obs_intr[o];
        cm_uv[2 * q] = P.obs_uv[2 * size_t(o)];
        cm_uv[2 * q + 1] = P.obs_uv[2 * size_t(o) + 1];
      }
      dp.cm_point = arena.upload(cm_point, stream);
      dp.cm_intr = arena.upload(cm_intr, stream);
      dp.cm_uv = arena.upload(cm_uv, stream);
    }
    dp.s_ptr = arena.upload(S.s_ptr, stream);
    dp.s_col = arena.upload(S.s_col, stream);
    {
      std::vector<int32_t> srow(S.s_col.size());
      for (int32_t c = 0; c < S.nb; ++c)
        for (int32_t k = S.s_ptr[c]; k < S.s_ptr[c + 1]; ++k) srow[k] = c;
      dp.s_row = arena.upload(srow, stream);
    }
    dp.pair_ptr = arena.upload(S.pair_ptr, stream);
    dp.pairs = d_pairs ? d_pairs : arena.upload(S.pairs, stream);
    dp.own_chunks = arena.upload(S.own_chunks, stream);
    dp.cam_chunk_ptr = arena.upload(S.cam_chunk_ptr, stream);
    dp.n_own_chunks = int32_t(S.own_chunks.size() / 3);
    dp.nslot = int64_t(S.s_col.size());
    dp.s_off = arena.upload(S.s_off, stream);
    dp.s_ld = arena.upload(S.s_ld, stream);
    dp.robust = P.robust;
    dp.delta = P.delta;
    dp.cam_model = P.cam_model;
    dp.intr_block = P.intr_block;
    dp.intr_member = P.intr_member;
    dp.intr_slot = P.intr_block >= 0 ? arena.upload(S.intr_slot, stream) : nullptr;
    dp.intr_diag_slot = S.intr_diag_slot;
    dp.n_eq = int32_t(P.eqs.size());
    if (!P.eqs.empty()) {
      std::vector<int32_t> member, set_ptr{0};
      std::vector<double> radius;
      for (const auto& e : P.eqs) {
        member.insert(member.end(), e.member.begin(), e.member.end());
        radius.insert(radius.end(), e.radius.begin(), e.radius.end());
        set_ptr.push_back(int32_t(member.size()));
      }
      deq.nset = int32_t(P.eqs.size());
      deq.nterm = int32_t(member.size());
      deq.term_member = arena.upload(member, stream);
      deq.set_ptr = arena.upload(set_ptr, stream);
      deq.radius = arena.upload(radius, stream);
      deq.G = arena.alloc<double>(size_t(7) * deq.nterm);
      deq.c = arena.alloc<double>(size_t(deq.nset));
      deq.Y = arena.alloc<double>(size_t(6) * S.nb * deq.nset);
      deq.b = arena.alloc<double>(size_t(6) * S.nb);
      deq.bperm = arena.alloc<double>(size_t(6) * S.nb);
      deq.z = arena.alloc<double>(size_t(6) * S.nb);
      // EqSystem::new starts the multipliers at zero (eq_system.rs:37)
      const std::vector<double> zero(size_t(deq.nset), 0.0);
      for (int i = 0; i < 2; ++i) deq.lambda[i] = arena.upload(zero, stream);
    }
    // factor structure
    df.nb = S.nb;
    df.scalar_pivots = params.solver == SGPU_SOLVER_SCHUR_SPARSE_LDLT ? 1 : 0;
    df.nsn = S.nsn;
    df.perm = arena.upload(S.perm, stream);
    df.sn_c0 = arena.upload(S.sn_c0, stream);
    df.sn_nc = arena.upload(S.sn_nc, stream);
    df.sn_nf = arena.upload(S.sn_nf, stream);
    df.sn_off = arena.upload(S.sn_off, stream);
    df.sn_loff = arena.upload(S.sn_loff, stream);
    df.sn_rows_ptr = arena.upload(S.sn_rows_ptr, stream);
    df.sn_rows = arena.upload(S.sn_rows, stream);
    df.sn_parent = arena.upload(S.sn_parent, stream);
    df.sn_upos_ptr = arena.upload(S.sn_upos_ptr, stream);
    df.sn_upos = arena.upload(S.sn_upos, stream);
    df.lvl_sn = arena.upload(S.lvl_sn, stream);
    df.launch_tiles = arena.upload(S.launch_tiles, stream);
    df.leaves = arena.upload(S.leaves, stream);
    df.sn_hand = arena.upload(S.sn_hand, stream);
    df.sn_toff = arena.upload(S.sn_toff, stream);
    df.sn_poff = arena.upload(S.sn_poff, stream);
    df.panel_items = arena.upload(S.panel_items, stream);
    df.sn_woff = arena.upload(S.sn_woff, stream);
    df.solve_cols = arena.upload(S.solve_cols, stream);
    df.sn_uoff = arena.upload(S.sn_uoff, stream);
    df.sn_child_ptr = arena.upload(S.sn_child_ptr, stream);
    df.sn_child = arena.upload(S.sn_child, stream);
    df.usolve = arena.alloc<double>(size_t(S.u_elems));
    if (!iterative()) {
      df.front = arena.alloc<double>(size_t(S.front_elems));
      df.lbuf = arena.alloc<double>(size_t(S.l_elems));
      df.ltbuf = arena.alloc<double>(size_t(S.t_elems));
      df.tinv = arena.alloc<double>(size_t(S.p_elems));
      df.wbuf = arena.alloc<double>(size_t(S.w_elems));
    }
    df.piv_L = arena.alloc<double>(size_t(36) * S.nb);
    df.piv_d = arena.alloc<double>(size_t(6) * S.nb);
    df.piv_M = arena.alloc<double>(size_t(36) * S.nb);
    df.front_elems = S.front_elems;
    df.host = &S;
    // workspace
    dl.cw = arena.alloc<double>(size_t(12) * S.npose);
    dl.prec = arena.alloc<double>(size_t(kPtRec) * S.npt);
    dl.pt_s = arena.alloc<double>(size_t(3) * S.npt);
    dl.rhs_nat = arena.alloc<double>(size_t(6) * S.nb);
    dl.grad_cam = arena.alloc<double>(size_t(6) * S.nb);
    dl.x_perm = arena.alloc<double>(size_t(6) * S.nb);
    dl.x_aux = arena.alloc<double>(size_t(6) * S.nb);
    dl.dx_cam = arena.alloc<double>(size_t(6) * S.nb);
    dl.vobs = arena.alloc<double>(size_t(12) * std::max<int32_t>(S.cam_ptr[S.nb], 1));
    dl.own_part = arena.alloc<double>(size_t(48) * std::max<size_t>(S.own_chunks.size() / 3, 1));
    dl.bsr = nullptr;
    if (iterative()) {
      std::vector<int32_t> mv_ptr, mv_ent, diag_slot, mv_row;
      build_pcg_pattern(S.nb, S.s_ptr, S.s_col, mv_ptr, mv_ent, diag_slot, mv_row);
      dl.bsr = arena.alloc<double>(size_t(36) * S.s_col.size());
      pcg.nb = S.nb;
      pcg.mv_ptr = arena.upload(mv_ptr, stream);
      pcg.mv_ent = arena.upload(mv_ent, stream);
      pcg.n_full = mv_ptr[S.nb];
      pcg.frow = arena.upload(mv_row, stream);
      pcg.fbsr = arena.alloc<double>(size_t(36) * pcg.n_full);
      pcg.fcol = arena.alloc<int32_t>(size_t(pcg.n_full));
      pcg.diag_slot = arena.upload(diag_slot, stream);
      pcg.bsr = dl.bsr;
      pcg.minv = arena.alloc<double>(size_t(36) * S.nb);
      pcg.x = dl.dx_cam;
      pcg.r = arena.alloc<double>(size_t(6) * S.nb);
      pcg.z = arena.alloc<double>(size_t(6) * S.nb);
      pcg.p = arena.alloc<double>(size_t(6) * S.nb);
      pcg.q = arena.alloc<double>(size_t(6) * S.nb);
      pcg.part = arena.alloc<double>(size_t(pcg_partial_len(S.nb)));
      pcg.state = arena.alloc<PcgState>(1);
      pcg.n_vec_blocks = pcg_vec_blocks(S.nb);
      pcg.n_mv_blocks = pcg_mv_blocks(S.nb);
      SGPU_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_pcg), sizeof(PcgState), hipHostMallocDefault));
      if (params.solver == SGPU_SOLVER_SCHUR_PCG_LEARNED) {
        learned.wbf16 = arena.upload(weights_to_bf16(precond_weights, kPcWeights), stream);
        learned.feat = arena.alloc<float>(size_t(kPcFeat) * S.nb);
        learned.h1 = arena.alloc<float>(size_t(kPcHid) * S.nb);
        learned.h2 = arena.alloc<float>(size_t(kPcHid) * S.nb);
        learned.out = arena.alloc<float>(size_t(kPcOut) * S.nb);
      }
    }
    dl.pt_vw = nullptr;
    dl.intr_part = nullptr;
    dl.intr_cpart = nullptr;
    if (P.intr_block >= 0) {
      dl.pt_vw = arena.alloc<double>(size_t(36) * std::max<int32_t>(S.npt, 1));
      dl.intr_part = arena.alloc<double>(size_t(intrinsics_partials(dp)));
      dl.intr_cpart = arena.alloc<double>(size_t(36) * std::max<size_t>(S.own_chunks.size() / 3, 1));
    }
    dl.partial_cap = 4 * kCostBlocks + 4;
    dl.partial = arena.alloc<double>(dl.partial_cap);
    cost_partial = arena.alloc<double>(size_t(2) * kCostBlocks * (1 + P.pg.size()));
    d_sc = arena.alloc<DevScalars>(1);
    for (size_t i = 0; i < P.pg.size(); ++i) {
      const auto& g = P.pg[i];
      DevPoseGraph q{};
      q.n = int32_t(g.m.size());
      q.m = arena.upload(g.m, stream);
      q.nn = arena.upload(g.n, stream);
      q.meas = arena.upload(g.meas, stream);
      q.first = arena.upload(g.first, stream);
      std::vector<int32_t> dest;
      for (auto& d : S.pg_dest[i]) dest.insert(dest.end(), {d.mm, d.nn, d.mn, d.mn_flip});
      q.dest = arena.upload(dest, stream);
      q.robust = g.robust;
      q.delta = g.delta;
      q.prior = g.prior ? 1 : 0;
      q.prec = g.prior ? arena.upload(g.prec, stream) : nullptr;
      dpg.push_back(q);
    }
    nu = params.initial_lm_damping;
    merit = compute_merit(poses[0], points[0], intr[0]);
  }

  // The CG tolerance of this step: pcg_relative_tolerance, or with pcg_forcing = eta > 0
  // the forcing term max(tol, eta min(1, sqrt(|g| / |g0|))) from the gradients the host has
  // read (the previous step's linearization and the first one).
  double pcg_tolerance() const {
    const double tol = params.pcg_relative_tolerance;
    if (!(params.pcg_forcing > 0.0)) return tol;
    const double ratio = grad0 > 0.0 && grad_last >= 0.0 ? std::sqrt(grad_last / grad0) : 1.0;
    return std::max(tol, params.pcg_forcing * std::min(1.0, ratio));
  }

  void step(sgpu_step_info* info) {
    if (terminated) throw Error(SGPU_ERR_TERMINATED, "optimizer already terminated");
    const auto t0 = std::chrono::steady_clock::now();
    bool used[K_COUNT] = {};
    const int nxt = 1 - cur;
    SGPU_HIP(hipEventRecord(phase_ev[0], stream));
    linearize(poses[cur], points[cur], intr[cur], nu, used);
    SGPU_HIP(hipEventRecord(phase_ev[1], stream));
    if (iterative()) {
      // block-Jacobi PCG on the resident S (two launches per CG iteration); its iteration count
      // and breakdown flag travel back with the step scalars
      tic(K_FACTOR);
      launch_pcg_setup(stream, pcg);
      if (params.solver == SGPU_SOLVER_SCHUR_PCG_LEARNED)
        launch_learned_preconditioner(stream, pcg, learned);
      launch_pcg(stream, pcg, dl.rhs_nat, pcg_tolerance(),
                 int32_t(std::min<int64_t>(params.pcg_max_iterations, INT32_MAX)),
                 &d_sc->pcg_iterations, &d_sc->bad_pivot, h_pcg);
      toc(K_FACTOR);
      SGPU_HIP(hipEventRecord(phase_ev[2], stream));
      tic(K_TRIANGULAR);
    } else {
      tic(K_FACTOR);
      launch_rcs_factor(stream, df, d_sc, dl.rhs_nat, dl.x_perm);
      toc(K_FACTOR);
      SGPU_HIP(hipEventRecord(phase_ev[2], stream));
      tic(K_TRIANGULAR);
      launch_rcs_solve(stream, df, dl.x_perm, dl.dx_cam);
      // equality constraints: the range-space KKT correction of dx and the multipliers' step
      launch_eq_solve(stream, dp, deq, df, deq.lambda[cur], deq.lambda[nxt], dl, d_sc);
    }
    // the mixed path solves for centred camera tangents (see k_schur_slots)
    if (params.precision == SGPU_PRECISION_MIXED) launch_uncenter(stream, dp, dl.cw, dl.dx_cam);
    toc(K_TRIANGULAR);
    tic(K_BACKSUB);
    launch_backsub_points(stream, params.precision, dp, points[cur], intr[cur], dl.cw, dl,
                          points[nxt], d_sc);
    toc(K_BACKSUB);
    SGPU_HIP(hipMemcpyAsync(poses[nxt], poses[cur], pose_fam().members.size() * 8,
                            hipMemcpyDeviceToDevice, stream));
    if (P.intr_block >= 0)
      SGPU_HIP(hipMemcpyAsync(intr[nxt], intr[cur], P.fams[P.f_intr].members.size() * 8,
                              hipMemcpyDeviceToDevice, stream));
    tic(K_UPDATE);
    launch_update_poses(stream, dp, poses[cur], intr[cur], dl, poses[nxt], intr[nxt], d_sc);
    toc(K_UPDATE);
    launch_finalize_step(stream, dp, dl, d_sc);
    SGPU_HIP(hipEventRecord(phase_ev[3], stream));
    cost_of(poses[nxt], points[nxt], intr[nxt], used);
    SGPU_HIP(hipEventRecord(phase_ev[4], stream));
    SGPU_HIP(hipMemcpyAsync(h_sc, d_sc, sizeof(DevScalars), hipMemcpyDeviceToHost, stream));
    SGPU_HIP(hipStreamSynchronize(stream));
    used[K_FACTOR] = used[K_TRIANGULAR] = used[K_BACKSUB] = used[K_UPDATE] = true;
    collect(used);
    const DevScalars sc = *h_sc;
    if (grad0 < 0.0) grad0 = bits_to_double(sc.max_grad_bits);
    grad_last = bits_to_double(sc.max_grad_bits);
    // LM control, step_nlls_inner order
    int reason = SGPU_TERM_NONE;
    int accepted = 0;
    double new_merit = NAN, rho = NAN;
    const double max_grad = bits_to_double(sc.max_grad_bits);
    const double max_step = bits_to_double(sc.max_delta_bits);
    // check_gradient_tolerance_unless_eq (optimizer.rs:527-535)
    if (P.eqs.empty() && max_grad < params.gradient_tolerance) {
      reason = SGPU_TERM_GRADIENT_TOLERANCE;
    } else if (sc.singular_marg || sc.bad_pivot) {
      nu *= 2.0;  // handle_solve_failure -> reject_step
      ++rejects;
      if (rejects >= params.max_consecutive_rejects) reason = SGPU_TERM_MAX_CONSECUTIVE_REJECTS;
    } else {
      new_merit = sc.cost;
      if (max_step < params.parameter_tolerance) {
        if (new_merit < merit) {
          cur = nxt;
          merit = new_merit;
          accepted = 1;
        }
        reason = SGPU_TERM_PARAMETER_TOLERANCE;
      } else {
        const double pred = nu * sc.dd + sc.dg;
        const double actual = merit - new_merit;
        rho = std::fabs(pred) > DBL_MIN ? actual / pred : (actual > 0 ? 1.0 : 0.0);
        if (rho > 0.0 && new_merit < merit) {
          const double f = 1.0 - std::pow(2.0 * rho - 1.0, 3);
          nu *= std::max(f, 1.0 / 3.0);
          rejects = 0;
          const double rel = (merit - new_merit) / (std::fabs(merit) + DBL_MIN);
          if (rel < params.function_tolerance) reason = SGPU_TERM_FUNCTION_TOLERANCE;
          cur = nxt;
          merit = new_merit;
          accepted = 1;
        } else {
          nu *= 2.0;
          ++rejects;
          if (rejects >= params.max_consecutive_rejects) reason = SGPU_TERM_MAX_CONSECUTIVE_REJECTS;
        }
      }
    }
    ++steps;
    if (steps >= params.num_iterations && reason == SGPU_TERM_NONE) reason = SGPU_TERM_MAX_ITERATIONS;
    if (reason != SGPU_TERM_NONE) {
      terminated = true;
      last_termination = reason;
    }
    if (info) {
      float a = 0, b = 0, c = 0, d = 0;
      SGPU_HIP(hipEventElapsedTime(&a, phase_ev[0], phase_ev[1]));
      SGPU_HIP(hipEventElapsedTime(&b, phase_ev[1], phase_ev[2]));
      SGPU_HIP(hipEventElapsedTime(&c, phase_ev[2], phase_ev[3]));
      SGPU_HIP(hipEventElapsedTime(&d, phase_ev[3], phase_ev[4]));
      info->merit = merit;
      info->smooth_cost = merit;
      info->inner_step = steps - 1;
      info->lm_damping = nu;
      info->termination = reason;
      info->accepted = accepted;
      info->new_merit = new_merit;
      info->gain_ratio = rho;
      info->max_abs_neg_grad = max_grad;
      info->max_abs_delta = max_step;
      info->pcg_iterations = sc.pcg_iterations;
      info->ms_linearize = a;
      info->ms_factor = b;
      info->ms_solve = c;
      info->ms_cost = d;
      info->ms_total =
          std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
  }

  void copy_members(int fam_sorted, double* out, int64_t capacity) {
    const Family& f = P.fams[fam_sorted];
    const int64_t need = int64_t(f.members.size());
    if (capacity < need) throw Error(SGPU_ERR_INVALID_ARGUMENT, "output capacity too small");
    if (fam_sorted == P.f_pose || fam_sorted == P.f_point ||
        (fam_sorted == P.f_intr && P.intr_block >= 0)) {
      const double* src = fam_sorted == P.f_pose    ? poses[cur]
                          : fam_sorted == P.f_point ? points[cur]
                                                    : intr[cur];
      SGPU_HIP(hipMemcpyAsync(out, src, need * 8, hipMemcpyDeviceToHost, stream));
      SGPU_HIP(hipStreamSynchronize(stream));
    } else {
      std::memcpy(out, f.members.data(), need * 8);
    }
  }

  // ---- inverse blocks (SchurFactor::inverse_block, schur_ldlt.rs:814-958) ----------------
  // The matrix is the damped Hessian of build_solution's final linearization at the current
  // variables (nu = 0, or the current nu with skip_final_hessian), as DampedHessian holds it
  // (damped_hessian.rs:98).  S is factored with the direct solver's kernels, S = L B L^T with
  // B = blockdiag(L_jj diag(d) L_jj^T).  Like the reference's min-norm LDLt
  // (block_sparse_min_norm_ldlt.rs:63-88), only positive pivots span the range: with S' = S
  // minus the non-positive pivots' terms, S'^+ b = P S'^- P b where S'^- = L^-T M' L^-1,
  // M'_j = L_jj^-T diag(d+)^-1 L_jj^-1 (non-positive pivots give 0), and P projects off the
  // null space of S', spanned by n = L^-T (L_jj^-T e_c) for each non-positive pivot c of
  // block j (S' n = 0).  This is the Moore-Penrose inverse the reference forms from E and G.
  void ensure_cov() {
    if (cov_ready) return;
    if (iterative()) throw Error(SGPU_ERR_UNSUPPORTED, "inverse blocks need the direct RCS solver");
    if (!P.eqs.empty())
      throw Error(SGPU_ERR_UNSUPPORTED, "inverse blocks of an equality-constrained problem");
    const int64_t n = 6 * int64_t(S.nb);
    if (!cov_b) {
      cov_b = arena.alloc<double>(size_t(n));
      cov_bperm = arena.alloc<double>(size_t(n));
      cov_z = arena.alloc<double>(size_t(n));
      cov_x = arena.alloc<double>(size_t(n));
      for (int32_t b = 0; b < S.npt; ++b) cov_max_obs = std::max(cov_max_obs, S.pt_ptr[b + 1] - S.pt_ptr[b]);
      cov_coup = arena.alloc<double>(size_t(18) * std::max(cov_max_obs, 1));
      cov_cam = arena.alloc<int32_t>(size_t(std::max(cov_max_obs, 1)));
      cov_hinv = arena.alloc<double>(9);
      cov_sdiag = arena.alloc<double>(size_t(n));
    }
    bool used[K_COUNT] = {};
    linearize(poses[cur], points[cur], intr[cur], params.skip_final_hessian ? nu : 0.0, used);
    launch_gather_s_diag(stream, dp, df.front, cov_sdiag);
    launch_rcs_factor(stream, df, d_sc, dl.rhs_nat, dl.x_perm);
    SGPU_HIP(hipMemcpyAsync(h_sc, d_sc, sizeof(DevScalars), hipMemcpyDeviceToHost, stream));
    std::vector<double> d(n), Lp(size_t(6) * n), sdiag(n);
    SGPU_HIP(hipMemcpyAsync(d.data(), df.piv_d, n * 8, hipMemcpyDeviceToHost, stream));
    SGPU_HIP(hipMemcpyAsync(sdiag.data(), cov_sdiag, n * 8, hipMemcpyDeviceToHost, stream));
    SGPU_HIP(hipMemcpyAsync(Lp.data(), df.piv_L, size_t(6) * n * 8, hipMemcpyDeviceToHost, stream));
    SGPU_HIP(hipStreamSynchronize(stream));
    if (h_sc->singular_marg) throw Error(SGPU_ERR_NON_FINITE, "singular marginal block in the final Hessian");
    if (h_sc->bad_pivot) throw Error(SGPU_ERR_NON_FINITE, "non-finite pivot in the final Hessian");
    cov_null.clear();
    {
      // pivots are in elimination order: pivot block i is camera block perm[i]
      double lo = INFINITY;
      for (int64_t i = 0; i < n; ++i) {
        const double sd = sdiag[6 * size_t(S.perm[i / 6]) + i % 6];
        lo = std::min(lo, sd > 0.0 ? d[i] / sd : 0.0);
      }
      cov_pivot_ratio = std::isfinite(lo) ? lo : 0.0;
    }
    std::vector<double> e(n), v(n);
    for (int64_t gb = 0; gb < S.nb; ++gb) {
      bool any = false;
      for (int c = 0; c < 6; ++c) any = any || !(d[6 * gb + c] > 0.0);
      if (!any) continue;
      // Li = L_jj^-1 (unit lower, row-major) by forward substitution
      const double* Lj = &Lp[36 * gb];
      double Li[36];
      for (int c = 0; c < 6; ++c)
        for (int r = 0; r < 6; ++r) {
          double x = r == c ? 1.0 : 0.0;
          for (int t = c; t < r; ++t) x -= Lj[6 * r + t] * Li[6 * t + c];
          Li[6 * r + c] = r < c ? 0.0 : x;
        }
      double Mj[36];  // L_jj^-T diag(1/d+) L_jj^-1
      for (int a = 0; a < 6; ++a)
        for (int b = 0; b < 6; ++b) {
          double acc = 0.0;
          for (int k = 0; k < 6; ++k) {
            const double dk = d[6 * gb + k];
            if (dk > 0.0) acc += Li[6 * k + a] * (1.0 / dk) * Li[6 * k + b];
          }
          Mj[6 * a + b] = acc;
        }
      SGPU_HIP(hipMemcpyAsync(df.piv_M + 36 * gb, Mj, 36 * 8, hipMemcpyHostToDevice, stream));
      for (int c = 0; c < 6; ++c) {
        if (d[6 * gb + c] > 0.0) continue;
        std::fill(e.begin(), e.end(), 0.0);
        for (int r = 0; r < 6; ++r) e[6 * gb + r] = Li[6 * c + r];  // L_jj^-T e_c
        SGPU_HIP(hipMemcpyAsync(cov_z, e.data(), n * 8, hipMemcpyHostToDevice, stream));
        launch_rcs_solve(stream, df, cov_z, cov_x);
        SGPU_HIP(hipMemcpyAsync(v.data(), cov_x, n * 8, hipMemcpyDeviceToHost, stream));
        SGPU_HIP(hipStreamSynchronize(stream));
        for (const auto& q : cov_null) {  // Gram-Schmidt against the earlier ones
          double dot = 0.0;
          for (int64_t i = 0; i < n; ++i) dot += q[i] * v[i];
          for (int64_t i = 0; i < n; ++i) v[i] -= dot * q[i];
        }
        double nrm = 0.0;
        for (double x : v) nrm += x * x;
        nrm = std::sqrt(nrm);
        if (nrm > 0.0) {
          for (double& x : v) x /= nrm;
          cov_null.push_back(v);
        }
      }
    }
    SGPU_HIP(hipStreamSynchronize(stream));
    cov_ready = true;
  }

  void project(std::vector<double>& v) const {
    for (const auto& q : cov_null) {
      double dot = 0.0;
      for (size_t i = 0; i < v.size(); ++i) dot += q[i] * v[i];
      for (size_t i = 0; i < v.size(); ++i) v[i] -= dot * q[i];
    }
  }

  // x = S^+ b, both in camera (layout) order
  std::vector<double> solve_pinv(std::vector<double> b) {
    const int64_t n = 6 * int64_t(S.nb);
    project(b);
    SGPU_HIP(hipMemcpyAsync(cov_b, b.data(), n * 8, hipMemcpyHostToDevice, stream));
    launch_rcs_forward(stream, df, cov_b, cov_bperm, cov_z);
    launch_rcs_solve(stream, df, cov_z, cov_x);
    std::vector<double> x(n);
    SGPU_HIP(hipMemcpyAsync(x.data(), cov_x, n * 8, hipMemcpyDeviceToHost, stream));
    SGPU_HIP(hipStreamSynchronize(stream));
    project(x);
    return x;
  }

  // H_{f, b}: three columns (camera order) of the point's coupling with the free cameras,
  // and its (damped) H_ll^-1 (row-major)
  void coupling(int32_t b, std::vector<std::vector<double>>& cols, double hinv[9]) {
    const int32_t nq = S.pt_ptr[b + 1] - S.pt_ptr[b];
    launch_point_coupling(stream, dp, points[cur], intr[cur], dl.cw, dl.prec, b, cov_coup, cov_cam,
                          cov_hinv);
    std::vector<double> H(size_t(18) * std::max(nq, 1));
    std::vector<int32_t> cams(std::max(nq, 1));
    double Wb[18] = {};
    SGPU_HIP(hipMemcpyAsync(H.data(), cov_coup, size_t(18) * nq * 8, hipMemcpyDeviceToHost, stream));
    SGPU_HIP(hipMemcpyAsync(cams.data(), cov_cam, size_t(nq) * 4, hipMemcpyDeviceToHost, stream));
    SGPU_HIP(hipMemcpyAsync(hinv, cov_hinv, 9 * 8, hipMemcpyDeviceToHost, stream));
    // free intrinsics: their coupling H_cl = W_b from the intrinsics point pass
    if (P.intr_block >= 0)
      SGPU_HIP(hipMemcpyAsync(Wb, dl.pt_vw + 36 * size_t(b) + 18, 18 * 8, hipMemcpyDeviceToHost, stream));
    SGPU_HIP(hipStreamSynchronize(stream));
    cols.assign(3, std::vector<double>(6 * size_t(S.nb), 0.0));
    for (int32_t q = 0; q < nq; ++q) {
      if (cams[q] < 0) continue;
      for (int r = 0; r < 6; ++r)
        for (int k = 0; k < 3; ++k) cols[k][6 * size_t(cams[q]) + r] += H[18 * q + 3 * r + k];
    }
    if (P.intr_block >= 0)
      for (int r = 0; r < 6; ++r)
        for (int k = 0; k < 3; ++k) cols[k][6 * size_t(P.intr_block) + r] += Wb[3 * r + k];
  }

  // (row, col) = (block, is_free); out row-major
  void inverse_block(int32_t rb, bool rfree, int32_t cb, bool cfree, double* out) {
    ensure_cov();
    const int64_t n = 6 * int64_t(S.nb);
    if (rfree && cfree) {  // S^+[row, col]
      for (int q = 0; q < 6; ++q) {
        std::vector<double> e(n, 0.0);
        e[6 * int64_t(cb) + q] = 1.0;
        const std::vector<double> x = solve_pinv(e);
        for (int r = 0; r < 6; ++r) out[6 * r + q] = x[6 * int64_t(rb) + r];
      }
      return;
    }
    if (rfree != cfree) {  // -(S^+ H_fm H_mm^-1)[free, marg], or its transpose
      const int32_t f = rfree ? rb : cb, m = rfree ? cb : rb;
      std::vector<std::vector<double>> cols;
      double hi[9];
      coupling(m, cols, hi);
      double T[18];
      for (int q = 0; q < 3; ++q) {
        const std::vector<double> x = solve_pinv(cols[q]);
        for (int r = 0; r < 6; ++r) T[3 * r + q] = x[6 * int64_t(f) + r];
      }
      for (int r = 0; r < 6; ++r)
        for (int q = 0; q < 3; ++q) {
          double acc = 0.0;
          for (int t = 0; t < 3; ++t) acc += T[3 * r + t] * hi[3 * t + q];
          if (rfree)
            out[3 * r + q] = -acc;
          else
            out[6 * q + r] = -acc;
        }
      return;
    }
    // H_mm[i]^-1 delta_ij + H_mm[i]^-1 (H_mf_i S^+ H_fm_j) H_mm[j]^-1
    std::vector<std::vector<double>> ci, cj;
    double hi[9], hj[9];
    coupling(rb, ci, hi);
    coupling(cb, cj, hj);
    double core[9];
    for (int q = 0; q < 3; ++q) {
      const std::vector<double> x = solve_pinv(cj[q]);
      for (int p = 0; p < 3; ++p) {
        double acc = 0.0;
        for (int64_t i = 0; i < n; ++i) acc += ci[p][i] * x[i];
        core[3 * p + q] = acc;
      }
    }
    double t1[9];
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) {
        double acc = 0.0;
        for (int t = 0; t < 3; ++t) acc += hi[3 * a + t] * core[3 * t + c];
        t1[3 * a + c] = acc;
      }
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) {
        double acc = 0.0;
        for (int t = 0; t < 3; ++t) acc += t1[3 * a + t] * hj[3 * t + c];
        out[3 * a + c] = acc + (rb == cb ? hi[3 * a + c] : 0.0);
      }
  }

  double final_max_grad(double nu_lin) {
    bool used[K_COUNT] = {};
    linearize(poses[cur], points[cur], intr[cur], nu_lin, used);
    SGPU_HIP(hipMemcpyAsync(h_sc, d_sc, sizeof(DevScalars), hipMemcpyDeviceToHost, stream));
    SGPU_HIP(hipStreamSynchronize(stream));
    collect(used);
    return bits_to_double(h_sc->max_grad_bits);
  }
};

extern "C" {

int32_t sgpu_abi_version(void) { return SGPU_ABI_VERSION; }

const char* sgpu_last_error(void) { return g_last_error.c_str(); }

void sgpu_opt_params_default(sgpu_opt_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->num_iterations = 100;
  p->initial_lm_damping = 10.0;
  p->parallelize = 1;
  p->solver = SGPU_SOLVER_SCHUR_BLOCK_SPARSE_LDLT;
  p->skip_final_hessian = 0;
  p->precision = SGPU_PRECISION_F64;
  p->function_tolerance = 1e-10;
  p->gradient_tolerance = 1e-14;
  p->parameter_tolerance = 1e-12;
  p->max_consecutive_rejects = 20;
  p->pcg_relative_tolerance = 1e-8;
  p->pcg_max_iterations = 500;
  p->device = 0;
}

sgpu_status sgpu_optimizer_create(const sgpu_var_family* families, int64_t num_families,
                                  const sgpu_cost_fn* cost_fns, int64_t num_cost_fns,
                                  const sgpu_opt_params* params, sgpu_optimizer** out) {
  return sgpu_optimizer_create_eq(families, num_families, cost_fns, num_cost_fns, nullptr, 0,
                                  params, out);
}

sgpu_status sgpu_optimizer_create_eq(const sgpu_var_family* families, int64_t num_families,
                                     const sgpu_cost_fn* cost_fns, int64_t num_cost_fns,
                                     const sgpu_eq_fn* eq_fns, int64_t num_eq_fns,
                                     const sgpu_opt_params* params, sgpu_optimizer** out) {
  return guarded([&] {
    if (!out || !params) throw Error(SGPU_ERR_INVALID_ARGUMENT, "params and out are required");
    *out = nullptr;
    auto opt = std::make_unique<sgpu_optimizer>();
    opt->init(families, num_families, cost_fns, num_cost_fns, eq_fns, num_eq_fns, params);
    *out = opt.release();
  });
}

sgpu_status sgpu_optimizer_eq_state(sgpu_optimizer* opt, double* lambda, double* residual,
                                    int64_t capacity) {
  return guarded([&] {
    if (!opt) throw Error(SGPU_ERR_INVALID_ARGUMENT, "null optimizer");
    const int64_t n = int64_t(opt->P.eqs.size());
    if (capacity < n) throw Error(SGPU_ERR_INVALID_ARGUMENT, "output capacity too small");
    if (n == 0) return;
    SGPU_HIP(hipSetDevice(opt->device));
    launch_eq_eval(opt->stream, opt->dp, opt->deq, opt->poses[opt->cur], opt->deq.lambda[opt->cur],
                   opt->dl, false);
    if (lambda)
      SGPU_HIP(hipMemcpyAsync(lambda, opt->deq.lambda[opt->cur], n * 8, hipMemcpyDeviceToHost, opt->stream));
    if (residual) SGPU_HIP(hipMemcpyAsync(residual, opt->deq.c, n * 8, hipMemcpyDeviceToHost, opt->stream));
    SGPU_HIP(hipStreamSynchronize(opt->stream));
  });
}

sgpu_status sgpu_optimizer_step(sgpu_optimizer* opt, sgpu_step_info* info) {
  return guarded([&] {
    if (!opt) throw Error(SGPU_ERR_INVALID_ARGUMENT, "null optimizer");
    SGPU_HIP(hipSetDevice(opt->device));
    opt->step(info);
  });
}

sgpu_status sgpu_optimizer_get_members(sgpu_optimizer* opt, const char* name, double* out,
                                       int64_t capacity) {
  return guarded([&] {
    if (!opt || !name || !out) throw Error(SGPU_ERR_INVALID_ARGUMENT, "null argument");
    for (size_t i = 0; i < opt->P.fams.size(); ++i)
      if (opt->P.fams[i].name == name) {
        SGPU_HIP(hipSetDevice(opt->device));
        opt->copy_members(int(i), out, capacity);
        return;
      }
    throw Error(SGPU_ERR_UNKNOWN_FAMILY, std::string("unknown family '") + name + "'");
  });
}

double sgpu_optimizer_merit(const sgpu_optimizer* opt) { return opt ? opt->merit : NAN; }
double sgpu_optimizer_lm_damping(const sgpu_optimizer* opt) { return opt ? opt->nu : NAN; }

sgpu_status sgpu_optimizer_eval_cost_at(sgpu_optimizer* opt, const double* const* members,
                                        double* cost) {
  return guarded([&] {
    if (!opt || !members || !cost) throw Error(SGPU_ERR_INVALID_ARGUMENT, "null argument");
    SGPU_HIP(hipSetDevice(opt->device));
    const int nxt = 1 - opt->cur;
    for (size_t i = 0; i < opt->P.sorted_of_input.size(); ++i) {
      const int s = opt->P.sorted_of_input[i];
      const auto& f = opt->P.fams[s];
      if (!members[i]) throw Error(SGPU_ERR_INVALID_ARGUMENT, "missing family values");
      if (s == opt->P.f_pose)
        SGPU_HIP(hipMemcpyAsync(opt->poses[nxt], members[i], f.members.size() * 8, hipMemcpyHostToDevice, opt->stream));
      else if (s == opt->P.f_point)
        SGPU_HIP(hipMemcpyAsync(opt->points[nxt], members[i], f.members.size() * 8, hipMemcpyHostToDevice, opt->stream));
      else if (s == opt->P.f_intr && opt->P.intr_block >= 0)
        SGPU_HIP(hipMemcpyAsync(opt->intr[nxt], members[i], f.members.size() * 8, hipMemcpyHostToDevice, opt->stream));
      else if (s == opt->P.f_intr &&
               std::memcmp(members[i], f.members.data(), f.members.size() * 8) != 0)
        throw Error(SGPU_ERR_UNSUPPORTED, "conditioned intrinsics differ from the handle's");
    }
    *cost = opt->compute_merit(opt->poses[nxt], opt->points[nxt],
                               opt->P.intr_block >= 0 ? opt->intr[nxt] : opt->intr[opt->cur]);
  });
}

sgpu_status sgpu_optimizer_build_solution(sgpu_optimizer* opt, sgpu_solution* sol) {
  return guarded([&] {
    if (!opt || !sol) throw Error(SGPU_ERR_INVALID_ARGUMENT, "null argument");
    SGPU_HIP(hipSetDevice(opt->device));
    sol->final_cost = opt->merit;
    sol->num_steps = opt->steps;
    sol->termination = opt->last_termination;
    sol->final_max_abs_neg_grad = opt->final_max_grad(opt->params.skip_final_hessian ? opt->nu : 0.0);
  });
}

static void fill_stats(const HostProblem& P, const Symbolic& S, sgpu_rcs_stats* st) {
  st->num_free_blocks = S.nb;
  st->num_marg_blocks = P.num_marg_blocks;
  st->num_observations = S.nobs;
  int64_t npg = 0;
  for (auto& g : P.pg) npg += int64_t(g.m.size());
  st->num_pose_terms = npg;
  st->nnz_s_blocks = int64_t(S.s_col.size());
  st->nnz_l_blocks = S.nnz_l;
  st->num_schur_pairs = S.num_schur_pairs;
  st->num_levels = int64_t(S.lvl_ptr.size()) - 1;
  st->factor_flops = S.factor_flops;
  st->num_supernodes = S.nsn;
  st->max_front_blocks = S.max_front;
  st->front_bytes = S.front_elems * 8;
  st->mf_step_flops = 0.0;
  st->mf_step_bytes = 0.0;
  st->mf_step_launches = 0;
  for (const auto& L : S.launches) {
    st->mf_step_flops += L.flops;
    st->mf_step_bytes += L.bytes;
    st->mf_step_launches += L.tile_count > 0 ? 1 : 0;
  }
}

sgpu_status sgpu_optimizer_rcs_stats(const sgpu_optimizer* opt, sgpu_rcs_stats* st) {
  return guarded([&] {
    if (!opt || !st) throw Error(SGPU_ERR_INVALID_ARGUMENT, "null argument");
    fill_stats(opt->P, opt->S, st);
  });
}

sgpu_status sgpu_analyze(const sgpu_var_family* families, int64_t num_families,
                         const sgpu_cost_fn* cost_fns, int64_t num_cost_fns,
                         sgpu_rcs_stats* st) {
  return guarded([&] {
    if (!st) throw Error(SGPU_ERR_INVALID_ARGUMENT, "null argument");
    const HostProblem P = build_host_problem(families, num_families, cost_fns, num_cost_fns);
    const Symbolic S = analyze(P);
    fill_stats(P, S, st);
  });
}

sgpu_status sgpu_optimizer_inverse_block(sgpu_optimizer* opt, const char* row_family,
                                         int64_t row_member, const char* col_family,
                                         int64_t col_member, double* out, int64_t capacity) {
  return guarded([&] {
    if (!opt || !row_family || !col_family || !out) throw Error(SGPU_ERR_INVALID_ARGUMENT, "null argument");
    auto resolve = [&](const char* name, int64_t member, int32_t& block, bool& is_free) {
      for (size_t i = 0; i < opt->P.fams.size(); ++i) {
        const Family& f = opt->P.fams[i];
        if (f.name != name) continue;
        if (member < 0 || member >= f.n) throw Error(SGPU_ERR_INDEX_OUT_OF_RANGE, "member index out of range");
        if (f.kind == SGPU_VAR_CONDITIONED || f.block[member] < 0)
          throw Error(SGPU_ERR_INVALID_ARGUMENT,
                      std::string("member of '") + name + "' is conditioned or constant: it has no block");
        is_free = f.kind == SGPU_VAR_FREE;
        block = !is_free ? int32_t(member) : int(i) == opt->P.f_intr ? opt->P.intr_block : f.block[member];
        return;
      }
      throw Error(SGPU_ERR_UNKNOWN_FAMILY, std::string("unknown family '") + name + "'");
    };
    int32_t rb = 0, cb = 0;
    bool rf = false, cf = false;
    resolve(row_family, row_member, rb, rf);
    resolve(col_family, col_member, cb, cf);
    const int64_t need = int64_t(rf ? 6 : 3) * (cf ? 6 : 3);
    if (capacity < need) throw Error(SGPU_ERR_INVALID_ARGUMENT, "output capacity too small");
    SGPU_HIP(hipSetDevice(opt->device));
    opt->inverse_block(rb, rf, cb, cf, out);
  });
}

sgpu_status sgpu_optimizer_covariance_block(sgpu_optimizer* opt, const char* row_family,
                                            int64_t row_member, const char* col_family,
                                            int64_t col_member, double* out, int64_t capacity) {
  if (!opt) return set_error(SGPU_ERR_INVALID_ARGUMENT, "null optimizer");
  if (opt->params.skip_final_hessian)
    return set_error(SGPU_ERR_UNSUPPORTED,
                     "covariance blocks are of the undamped final Hessian (build_solution without "
                     "skip_final_hessian)");
  const sgpu_status st = sgpu_optimizer_inverse_block(opt, row_family, row_member, col_family,
                                                      col_member, out, capacity);
  if (st != SGPU_OK) return st;
  // rank-deficient (a non-positive pivot) or numerically singular (a pivot below 1e-9 of the
  // diagonal entry of S in its column, e.g. a gauge direction left at rounding level with
  // either sign)
  if (!opt->cov_null.empty() || opt->cov_pivot_ratio < 1e-9)
    return set_error(SGPU_ERR_UNSUPPORTED,
                     "the final Hessian is rank-deficient (" + std::to_string(opt->cov_null.size()) +
                         " non-positive pivots of its Schur complement, smallest pivot / diagonal " +
                         std::to_string(opt->cov_pivot_ratio) +
                         ", e.g. an unfixed gauge): its min-norm inverse is not the "
                         "Schur-complement inverse block; use sgpu_optimizer_inverse_block for "
                         "the latter");
  return SGPU_OK;
}

// The elimination order in the reference's free-block numbering (var_families.rs:60-113).
static void copy_order(const HostProblem& P, const Symbolic& S, int64_t* perm, int64_t capacity,
                       int64_t* num_blocks) {
  if (num_blocks) *num_blocks = S.nb;
  if (!perm) return;
  if (capacity < S.nb) throw Error(SGPU_ERR_INVALID_ARGUMENT, "output capacity too small");
  for (int32_t i = 0; i < S.nb; ++i) perm[i] = P.ref_block[S.perm[i]];
}

sgpu_status sgpu_analyze_rcs_order(const sgpu_var_family* families, int64_t num_families,
                                   const sgpu_cost_fn* cost_fns, int64_t num_cost_fns,
                                   int64_t* perm, int64_t capacity, int64_t* num_blocks) {
  return guarded([&] {
    const HostProblem P = build_host_problem(families, num_families, cost_fns, num_cost_fns);
    copy_order(P, analyze(P), perm, capacity, num_blocks);
  });
}

sgpu_status sgpu_optimizer_rcs_order(const sgpu_optimizer* opt, int64_t* perm, int64_t capacity,
                                     int64_t* num_blocks) {
  return guarded([&] {
    if (!opt) throw Error(SGPU_ERR_INVALID_ARGUMENT, "null optimizer");
    copy_order(opt->P, opt->S, perm, capacity, num_blocks);
  });
}

sgpu_status sgpu_optimizer_rcs_pivots(sgpu_optimizer* opt, double* d, int64_t capacity) {
  return guarded([&] {
    if (!opt || !d) throw Error(SGPU_ERR_INVALID_ARGUMENT, "null argument");
    if (opt->iterative()) throw Error(SGPU_ERR_UNSUPPORTED, "the PCG solvers have no pivots");
    const int64_t need = int64_t(6) * opt->S.nb;
    if (capacity < need) throw Error(SGPU_ERR_INVALID_ARGUMENT, "output capacity too small");
    SGPU_HIP(hipSetDevice(opt->device));
    SGPU_HIP(hipMemcpyAsync(d, opt->df.piv_d, need * 8, hipMemcpyDeviceToHost, opt->stream));
    SGPU_HIP(hipStreamSynchronize(opt->stream));
  });
}

int32_t sgpu_optimizer_kernel_times(const sgpu_optimizer* opt, const char** names, int64_t* counts,
                                    double* total_ms, int32_t capacity) {
  if (!opt) return 0;
  int32_t n = 0;
  for (int k = 0; k < K_COUNT && n < capacity; ++k, ++n) {
    if (names) names[n] = kKernelNames[k];
    if (counts) counts[n] = opt->kernel_count[k];
    if (total_ms) total_ms[n] = opt->kernel_ms[k];
  }
  return n;
}

void sgpu_optimizer_reset_kernel_times(sgpu_optimizer* opt) {
  if (!opt) return;
  for (int k = 0; k < K_COUNT; ++k) {
    opt->kernel_ms[k] = 0;
    opt->kernel_count[k] = 0;
  }
}

void sgpu_optimizer_set_kernel_timing(sgpu_optimizer* opt, int32_t enable) {
  if (!opt) return;
  opt->timing = enable <= 0 ? 0 : enable == 1 ? 1 : 2;
  if (opt->timing == 2 && !opt->iterative() && opt->step_ev.empty()) {
    opt->step_ev.assign(2 * opt->S.launches.size(), nullptr);
    for (auto& e : opt->step_ev)
      if (hipEventCreate(&e) != hipSuccess) {
        for (auto& d : opt->step_ev)
          if (d) (void)hipEventDestroy(d);
        opt->step_ev.clear();
        break;
      }
  }
  opt->df.step_events = opt->timing == 2 && !opt->step_ev.empty() ? opt->step_ev.data() : nullptr;
}

void sgpu_optimizer_destroy(sgpu_optimizer* opt) { delete opt; }

sgpu_status sgpu_optimize_nlls(const sgpu_var_family* families, int64_t num_families,
                               const sgpu_cost_fn* cost_fns, int64_t num_cost_fns,
                               const sgpu_opt_params* params, double* const* out_members,
                               sgpu_solution* solution) {
  return sgpu_optimize_nlls_eq(families, num_families, cost_fns, num_cost_fns, nullptr, 0, params,
                               out_members, solution);
}

sgpu_status sgpu_optimize_nlls_eq(const sgpu_var_family* families, int64_t num_families,
                                  const sgpu_cost_fn* cost_fns, int64_t num_cost_fns,
                                  const sgpu_eq_fn* eq_fns, int64_t num_eq_fns,
                                  const sgpu_opt_params* params, double* const* out_members,
                                  sgpu_solution* solution) {
  return guarded([&] {
    if (!params || !solution) throw Error(SGPU_ERR_INVALID_ARGUMENT, "params and solution required");
    auto opt = std::make_unique<sgpu_optimizer>();
    opt->init(families, num_families, cost_fns, num_cost_fns, eq_fns, num_eq_fns, params);
    solution->initial_cost = opt->merit;
    while (!opt->terminated) opt->step(nullptr);
    solution->final_cost = opt->merit;
    solution->num_steps = opt->steps;
    solution->termination = opt->last_termination;
    solution->final_max_abs_neg_grad =
        opt->final_max_grad(params->skip_final_hessian ? opt->nu : 0.0);
    if (out_members)
      for (int64_t i = 0; i < num_families; ++i)
        if (out_members[i]) {
          const int s = opt->P.sorted_of_input[i];
          opt->copy_members(s, out_members[i], int64_t(opt->P.fams[s].members.size()));
        }
  });
}

}  // extern "C"

namespace sgpu {
const DevPcg& optimizer_pcg(const sgpu_optimizer* opt) { return opt->pcg; }
hipStream_t optimizer_stream(const sgpu_optimizer* opt) { return opt->stream; }
int64_t optimizer_nslot(const sgpu_optimizer* opt) { return opt->dp.nslot; }
}  // namespace sgpu
