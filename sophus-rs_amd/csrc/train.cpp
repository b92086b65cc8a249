// Training of the learned block-Jacobi correction and the stage-level C entry points of
// the preconditioner (see learned.hpp).
#include <algorithm>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "api_internal.hpp"
#include "learned.hpp"
#include "optimizer_internal.hpp"

namespace sgpu {

// training launchers (learned.hip)
void launch_whiten(hipStream_t s, const DevPcg& g, double* Bm, int32_t* ok);
void launch_probe(hipStream_t s, int64_t n, const double* Bm, uint64_t seed, double* w, double* v);
void launch_rcs_matvec(hipStream_t s, const DevPcg& g, const double* v, double* u);
void launch_utilde(hipStream_t s, int64_t n, const double* Bm, const double* u, double* ut);
void launch_loss_grad(hipStream_t s, int64_t n, const int32_t* ok, const double* ut, const double* w,
                      const float* out, float* dO, double* loss, int accumulate);
void launch_mlp_backward(hipStream_t s, int64_t n, const float* feat, const float* w32, const float* h1,
                         const float* h2, const float* dO, float* dA2, float* dA1, float* grad);
void launch_adam(hipStream_t s, float* w, const float* g, float* m, float* v, float lr, int step,
                 float scale);

namespace {

struct Snapshot {
  size_t problem;
  double* bsr;
};

// One seeded random-walk BA problem as C-ABI families / cost function.
struct TrainProblem {
  std::vector<double> intr, poses_gt, poses, points_gt, points, uv;
  std::vector<int64_t> idx;
  sgpu_optimizer* opt = nullptr;
};

void make_problem(const sgpu_precond_train_params& p, int i, TrainProblem& tp) {
  sgpu_ba_spec spec{};
  spec.layout = 1;
  spec.num_cameras = p.num_cameras;
  spec.num_points = p.num_points;
  spec.obs_per_point_ring = 5;
  spec.min_obs = 2;
  spec.max_obs = 30;
  spec.geometric_p = p.geometric_p;
  spec.pixel_noise = 0.5;
  spec.rot_noise = 0.05;
  spec.trans_noise = 0.1;
  spec.point_noise = 0.1;
  spec.step_length = 0.3;
  spec.yaw_jitter = 5.0 * 3.14159265358979323846 / 180.0;
  spec.seed = p.seed + uint64_t(i);
  sgpu_ba_sizes sz{};
  sgpu_status st = sgpu_generate_ba(&spec, &sz, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                    nullptr);
  if (st != SGPU_OK) throw Error(st, "training problem generation failed");
  tp.intr.resize(4);
  tp.poses_gt.resize(7 * sz.num_cameras);
  tp.poses.resize(7 * sz.num_cameras);
  tp.points_gt.resize(3 * sz.num_points);
  tp.points.resize(3 * sz.num_points);
  tp.idx.resize(3 * sz.num_obs);
  tp.uv.resize(2 * sz.num_obs);
  st = sgpu_generate_ba(&spec, &sz, tp.intr.data(), tp.poses_gt.data(), tp.poses.data(),
                        tp.points_gt.data(), tp.points.data(), tp.idx.data(), tp.uv.data());
  if (st != SGPU_OK) throw Error(st, "training problem generation failed");
  const int64_t const_ids[1] = {0};
  sgpu_var_family fams[3] = {
      {"cams", SGPU_VAR_PINHOLE, SGPU_VAR_CONDITIONED, 1, tp.intr.data(), 0, nullptr},
      {"poses", SGPU_VAR_ISOMETRY3, SGPU_VAR_FREE, sz.num_cameras, tp.poses.data(), 1, const_ids},
      {"points", SGPU_VAR_VEC3, SGPU_VAR_MARGINALIZED, sz.num_points, tp.points.data(), 0, nullptr}};
  sgpu_cost_fn cf{};
  cf.kind = SGPU_COST_PINHOLE_REPROJECTION;
  cf.robust_kernel = SGPU_ROBUST_HUBER;
  cf.huber_delta = 1.0;
  cf.family_names[0] = "cams";
  cf.family_names[1] = "poses";
  cf.family_names[2] = "points";
  cf.num_terms = sz.num_obs;
  cf.entity_indices = tp.idx.data();
  cf.measurements = tp.uv.data();
  sgpu_opt_params prm;
  sgpu_opt_params_default(&prm);
  prm.num_iterations = std::max(1, p.lm_iterations);
  prm.initial_lm_damping = 1.0;
  prm.solver = SGPU_SOLVER_SCHUR_PCG_JACOBI;
  prm.pcg_relative_tolerance = 1e-6;
  prm.pcg_max_iterations = 400;
  prm.device = p.device;
  st = sgpu_optimizer_create(fams, 3, &cf, 1, &prm, &tp.opt);
  if (st != SGPU_OK) throw Error(st, std::string("training optimizer: ") + sgpu_last_error());
}

}  // namespace
}  // namespace sgpu

using namespace sgpu;

extern "C" {

sgpu_status sgpu_precond_train(const sgpu_precond_train_params* params, float* weights_inout,
                               double* loss_history, int64_t* num_snapshots) {
  std::vector<TrainProblem> probs;
  auto cleanup = [&]() {
    for (auto& tp : probs)
      if (tp.opt) sgpu_optimizer_destroy(tp.opt);
  };
  const sgpu_status status = guarded([&] {
    if (!params || !weights_inout) throw Error(SGPU_ERR_INVALID_ARGUMENT, "params and weights required");
    const sgpu_precond_train_params p = *params;
    if (p.num_problems < 1 || p.num_cameras < 3 || p.num_points < 10 || p.epochs < 0 || p.probes < 1 ||
        !(p.learning_rate > 0.0) || !(p.geometric_p > 0.0 && p.geometric_p <= 1.0))
      throw Error(SGPU_ERR_INVALID_ARGUMENT, "bad training parameters");
    SGPU_HIP(hipSetDevice(p.device));
    // 1. RCS snapshots from LM runs with the plain block-Jacobi PCG
    probs.resize(size_t(p.num_problems));
    DeviceArena arena;
    std::vector<Snapshot> snaps;
    int32_t max_nb = 0;
    for (int i = 0; i < p.num_problems; ++i) {
      make_problem(p, i, probs[i]);
      sgpu_optimizer* opt = probs[i].opt;
      const DevPcg& g = optimizer_pcg(opt);
      max_nb = std::max(max_nb, g.nb);
      const int64_t nslot = optimizer_nslot(opt);
      for (int k = 0; k < p.lm_iterations; ++k) {
        sgpu_step_info info;
        const sgpu_status st = sgpu_optimizer_step(opt, &info);
        if (st != SGPU_OK) throw Error(st, std::string("training step: ") + sgpu_last_error());
        double* copy = arena.alloc<double>(size_t(36) * nslot);
        SGPU_HIP(hipMemcpyAsync(copy, g.bsr, size_t(36) * nslot * 8, hipMemcpyDeviceToDevice,
                                optimizer_stream(opt)));
        SGPU_HIP(hipStreamSynchronize(optimizer_stream(opt)));
        snaps.push_back({size_t(i), copy});
        if (info.termination != SGPU_TERM_NONE) break;
      }
    }
    if (num_snapshots) *num_snapshots = int64_t(snaps.size());
    // 2. training buffers
    const size_t nb = size_t(max_nb);
    float* feat = arena.alloc<float>(kPcFeat * nb);
    float* h1 = arena.alloc<float>(kPcHid * nb);
    float* h2 = arena.alloc<float>(kPcHid * nb);
    float* out = arena.alloc<float>(kPcOut * nb);
    float* dO = arena.alloc<float>(kPcOut * nb);
    float* dA2 = arena.alloc<float>(kPcHid * nb);
    float* dA1 = arena.alloc<float>(kPcHid * nb);
    double* Bm = arena.alloc<double>(36 * nb);
    int32_t* ok = arena.alloc<int32_t>(nb);
    double* wv = arena.alloc<double>(6 * nb);
    double* vv = arena.alloc<double>(6 * nb);
    double* uu = arena.alloc<double>(6 * nb);
    double* ut = arena.alloc<double>(6 * nb);
    double* loss = arena.alloc<double>(nb);
    float* w32 = arena.alloc<float>(kPcWeights);
    float* grad = arena.alloc<float>(kPcWeights);
    float* am = arena.alloc<float>(kPcWeights);
    float* av = arena.alloc<float>(kPcWeights);
    hipStream_t s = optimizer_stream(probs[0].opt);
    SGPU_HIP(hipMemcpy(w32, weights_inout, size_t(kPcWeights) * 4, hipMemcpyHostToDevice));
    SGPU_HIP(hipMemset(am, 0, size_t(kPcWeights) * 4));
    SGPU_HIP(hipMemset(av, 0, size_t(kPcWeights) * 4));
    std::vector<double> hloss(nb);
    std::mt19937_64 rng(p.seed ^ 0x5DEECE66Dull);
    std::vector<size_t> order(snaps.size());
    std::iota(order.begin(), order.end(), size_t(0));
    int adam_step = 0;
    // one pass over the snapshots: loss (and an Adam step per snapshot when training)
    auto pass = [&](int epoch, bool train) -> double {
      double tot = 0.0;
      int64_t count = 0;
      for (size_t q = 0; q < order.size(); ++q) {
        const Snapshot& sn = snaps[order[q]];
        DevPcg g = optimizer_pcg(probs[sn.problem].opt);
        g.bsr = sn.bsr;
        const int64_t n = g.nb;
        launch_pcg_setup(s, g);
        launch_precond_features(s, g, feat);
        launch_whiten(s, g, Bm, ok);
        launch_mlp_forward(s, n, feat, w32, nullptr, h1, h2, out);
        for (int k = 0; k < p.probes; ++k) {
          const uint64_t seed = p.seed * 0x9E3779B97F4A7C15ull + uint64_t(epoch) * 1000003ull +
                                uint64_t(order[q]) * 7919ull + uint64_t(k);
          launch_probe(s, n, Bm, seed, wv, vv);
          launch_rcs_matvec(s, g, vv, uu);
          launch_utilde(s, n, Bm, uu, ut);
          launch_loss_grad(s, n, ok, ut, wv, out, dO, loss, k > 0 ? 1 : 0);
        }
        if (train) {
          launch_mlp_backward(s, n, feat, w32, h1, h2, dO, dA2, dA1, grad);
          launch_adam(s, w32, grad, am, av, float(p.learning_rate), ++adam_step,
                      1.0f / float(n * p.probes));
        }
        SGPU_HIP(hipGetLastError());
        SGPU_HIP(hipMemcpyAsync(hloss.data(), loss, size_t(n) * 8, hipMemcpyDeviceToHost, s));
        SGPU_HIP(hipStreamSynchronize(s));
        for (int64_t b = 0; b < n; ++b) tot += hloss[b];
        count += n * p.probes;
      }
      return count ? tot / double(count) : 0.0;
    };
    const double l0 = pass(0, false);
    if (loss_history) loss_history[0] = l0;
    for (int e = 1; e <= p.epochs; ++e) {
      std::shuffle(order.begin(), order.end(), rng);
      const double le = pass(e, true);
      if (loss_history) loss_history[e] = le;
    }
    SGPU_HIP(hipMemcpy(weights_inout, w32, size_t(kPcWeights) * 4, hipMemcpyDeviceToHost));
  });
  cleanup();
  return status;
}

sgpu_status sgpu_precond_forward(int64_t n, const float* features, const float* weights, int32_t use_bf16,
                                 float* outputs) {
  return guarded([&] {
    if (n < 0 || (n > 0 && (!features || !weights || !outputs)))
      throw Error(SGPU_ERR_INVALID_ARGUMENT, "bad arguments");
    if (n == 0) return;
    precond_forward_host(n, features, weights, use_bf16, outputs);
  });
}

sgpu_status sgpu_precond_blocks(int64_t n, const double* diag_blocks, const float* outputs,
                                double* minv_out) {
  return guarded([&] {
    if (n < 0 || (n > 0 && (!diag_blocks || !outputs || !minv_out)))
      throw Error(SGPU_ERR_INVALID_ARGUMENT, "bad arguments");
    if (n == 0) return;
    precond_blocks_host(n, diag_blocks, outputs, minv_out);
  });
}

sgpu_status sgpu_precond_loss_grad(int64_t n, const float* features, const double* diag_blocks,
                                   const double* u, const double* w, const float* weights,
                                   double* loss_out, float* grad_out) {
  return guarded([&] {
    if (n <= 0 || !features || !diag_blocks || !u || !w || !weights || !loss_out || !grad_out)
      throw Error(SGPU_ERR_INVALID_ARGUMENT, "bad arguments");
    precond_loss_grad_host(n, features, diag_blocks, u, w, weights, loss_out, grad_out);
  });
}

}  // extern "C"
