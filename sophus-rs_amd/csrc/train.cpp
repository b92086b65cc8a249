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
                         const float* h2, const float* dO, const float* ones, float* dA2, float* dA1,
                         float* grad);
void launch_fill_f32(hipStream_t s, float* p, int64_t n, float v);
void launch_gather_rows(hipStream_t s, int64_t n, const int64_t* rowmap, int64_t rows_cap, int probes,
                        const float* feat, const int32_t* ok, const double* w, const double* ut,
                        float* feat_out, int32_t* ok_out, double* w_out, double* ut_out, int64_t out_cap);
void launch_adam(hipStream_t s, float* w, const float* g, float* m, float* v, float lr, int step,
                 float scale);

namespace {

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
                               double* loss_history, double* val_history, int64_t* num_snapshots) {
  std::vector<TrainProblem> probs;
  auto destroy = [](TrainProblem& tp) {
    if (tp.opt) sgpu_optimizer_destroy(tp.opt);
    tp.opt = nullptr;
  };
  auto cleanup = [&]() {
    for (auto& tp : probs) destroy(tp);
  };
  const sgpu_status status = guarded([&] {
    if (!params || !weights_inout) throw Error(SGPU_ERR_INVALID_ARGUMENT, "params and weights required");
    const sgpu_precond_train_params p = *params;
    if (p.num_problems < 1 || p.num_cameras < 3 || p.num_points < 10 || p.epochs < 0 || p.probes < 1 ||
        p.lm_iterations < 1 || p.batch_snapshots < 0 || p.holdout_problems < 0 ||
        !(p.learning_rate > 0.0) || !(p.geometric_p > 0.0 && p.geometric_p <= 1.0))
      throw Error(SGPU_ERR_INVALID_ARGUMENT, "bad training parameters");
    SGPU_HIP(hipSetDevice(p.device));
    const int nprob = p.num_problems + p.holdout_problems;
    const int64_t nb_max = p.num_cameras - 1;  // camera 0 is constant
    const int64_t rows_cap = int64_t(nprob) * p.lm_iterations * nb_max;
    const int P = p.probes;
    // 1. The data set, one row per camera block of every RCS snapshot that LM runs with the
    //    plain block-Jacobi PCG produce: the block's features, whether its diagonal block
    //    is positive definite, and `probes` fixed probe pairs (w, u~ = B^T S B w).  Only these
    //    rows are kept, not S itself.
    DeviceArena arena;
    float* d_feat = arena.alloc<float>(size_t(kPcFeat) * rows_cap);
    int32_t* d_ok = arena.alloc<int32_t>(size_t(rows_cap));
    double* d_w = arena.alloc<double>(size_t(6) * P * rows_cap);   // [probe][row][6]
    double* d_ut = arena.alloc<double>(size_t(6) * P * rows_cap);
    double* Bm = arena.alloc<double>(size_t(36) * nb_max);
    double* vv = arena.alloc<double>(size_t(6) * nb_max);
    double* uu = arena.alloc<double>(size_t(6) * nb_max);
    struct Snap {
      int64_t row0, n;
      bool holdout;
    };
    std::vector<Snap> snaps;
    int64_t rows = 0;
    probs.resize(size_t(nprob));
    for (int i = 0; i < nprob; ++i) {
      make_problem(p, i, probs[i]);
      sgpu_optimizer* opt = probs[i].opt;
      hipStream_t s = optimizer_stream(opt);
      for (int k = 0; k < p.lm_iterations; ++k) {
        sgpu_step_info info;
        const sgpu_status st = sgpu_optimizer_step(opt, &info);
        if (st != SGPU_OK) throw Error(st, std::string("training step: ") + sgpu_last_error());
        const DevPcg& g = optimizer_pcg(opt);
        const int64_t n = g.nb;
        if (rows + n > rows_cap) throw Error(SGPU_ERR_INVALID_ARGUMENT, "training rows overflow");
        launch_precond_features(s, g, d_feat + int64_t(kPcFeat) * rows);
        launch_whiten(s, g, Bm, d_ok + rows);
        for (int q = 0; q < P; ++q) {
          const uint64_t seed = p.seed * 0x9E3779B97F4A7C15ull + uint64_t(snaps.size()) * 7919ull + uint64_t(q);
          double* w = d_w + 6 * (int64_t(q) * rows_cap + rows);
          launch_probe(s, n, Bm, seed, w, vv);
          launch_rcs_matvec(s, g, vv, uu);
          launch_utilde(s, n, Bm, uu, d_ut + 6 * (int64_t(q) * rows_cap + rows));
        }
        SGPU_HIP(hipGetLastError());
        SGPU_HIP(hipStreamSynchronize(s));
        snaps.push_back({rows, n, i >= p.num_problems});
        rows += n;
        if (info.termination != SGPU_TERM_NONE) break;
      }
      destroy(probs[i]);  // the data set keeps what training needs
    }
    if (num_snapshots) *num_snapshots = int64_t(snaps.size());
    std::vector<int64_t> train_snaps, val_snaps;
    for (size_t k = 0; k < snaps.size(); ++k) (snaps[k].holdout ? val_snaps : train_snaps).push_back(int64_t(k));
    // 2. training buffers, sized for the largest batch (a batch is whole snapshots)
    const int bs = std::max(1, p.batch_snapshots);
    const int64_t batch_cap = int64_t(bs) * nb_max;
    const int64_t eval_cap = std::max<int64_t>(batch_cap, nb_max);
    float* feat = arena.alloc<float>(size_t(kPcFeat) * eval_cap);
    int32_t* ok = arena.alloc<int32_t>(size_t(eval_cap));
    double* bw = arena.alloc<double>(size_t(6) * P * eval_cap);
    double* but = arena.alloc<double>(size_t(6) * P * eval_cap);
    int64_t* rowmap = arena.alloc<int64_t>(size_t(eval_cap));
    float* h1 = arena.alloc<float>(size_t(kPcHid) * eval_cap);
    float* h2 = arena.alloc<float>(size_t(kPcHid) * eval_cap);
    float* out = arena.alloc<float>(size_t(kPcOut) * eval_cap);
    float* dO = arena.alloc<float>(size_t(kPcOut) * eval_cap);
    float* dA2 = arena.alloc<float>(size_t(kPcHid) * eval_cap);
    float* dA1 = arena.alloc<float>(size_t(kPcHid) * eval_cap);
    float* ones = arena.alloc<float>(size_t(eval_cap));
    double* loss = arena.alloc<double>(size_t(eval_cap));
    float* w32 = arena.alloc<float>(kPcWeights);
    float* grad = arena.alloc<float>(kPcWeights);
    float* am = arena.alloc<float>(kPcWeights);
    float* av = arena.alloc<float>(kPcWeights);
    hipStream_t s = nullptr;
    SGPU_HIP(hipStreamCreate(&s));
    struct StreamGuard {
      hipStream_t s;
      ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{s};
    launch_fill_f32(s, ones, eval_cap, 1.0f);
    SGPU_HIP(hipMemcpyAsync(w32, weights_inout, size_t(kPcWeights) * 4, hipMemcpyHostToDevice, s));
    SGPU_HIP(hipMemsetAsync(am, 0, size_t(kPcWeights) * 4, s));
    SGPU_HIP(hipMemsetAsync(av, 0, size_t(kPcWeights) * 4, s));
    std::vector<double> hloss(static_cast<size_t>(eval_cap));
    std::vector<int64_t> hmap(static_cast<size_t>(eval_cap));
    // gather the rows of `list` (snapshot indices) into the batch buffers; returns the rows
    auto gather = [&](const int64_t* list, int count) -> int64_t {
      int64_t n = 0;
      for (int c = 0; c < count; ++c) {
        const Snap& sn = snaps[size_t(list[c])];
        for (int64_t r = 0; r < sn.n; ++r) hmap[size_t(n++)] = sn.row0 + r;
      }
      SGPU_HIP(hipMemcpyAsync(rowmap, hmap.data(), size_t(n) * 8, hipMemcpyHostToDevice, s));
      launch_gather_rows(s, n, rowmap, rows_cap, P, d_feat, d_ok, d_w, d_ut, feat, ok, bw, but, eval_cap);
      return n;
    };
    // loss of a batch (and its gradient into grad when training)
    auto batch_loss = [&](int64_t n, bool train) -> double {
      launch_mlp_forward(s, n, feat, w32, nullptr, h1, h2, out);
      for (int q = 0; q < P; ++q)
        launch_loss_grad(s, n, ok, but + 6 * int64_t(q) * eval_cap, bw + 6 * int64_t(q) * eval_cap, out, dO,
                         loss, q > 0 ? 1 : 0);
      if (train) launch_mlp_backward(s, n, feat, w32, h1, h2, dO, ones, dA2, dA1, grad);
      SGPU_HIP(hipGetLastError());
      SGPU_HIP(hipMemcpyAsync(hloss.data(), loss, size_t(n) * 8, hipMemcpyDeviceToHost, s));
      SGPU_HIP(hipStreamSynchronize(s));
      double tot = 0.0;
      for (int64_t b = 0; b < n; ++b) tot += hloss[size_t(b)];
      return tot;
    };
    // mean loss per block and probe over a snapshot list, in batches of bs snapshots
    auto evaluate = [&](const std::vector<int64_t>& list) -> double {
      double tot = 0.0;
      int64_t count = 0;
      for (size_t c = 0; c < list.size(); c += size_t(bs)) {
        const int m = int(std::min<size_t>(size_t(bs), list.size() - c));
        const int64_t n = gather(list.data() + c, m);
        tot += batch_loss(n, false);
        count += n * P;
      }
      return count ? tot / double(count) : 0.0;
    };
    std::mt19937_64 rng(p.seed ^ 0x5DEECE66Dull);
    if (loss_history) loss_history[0] = evaluate(train_snaps);
    if (val_history) val_history[0] = evaluate(val_snaps);
    int adam_step = 0;
    for (int e = 1; e <= p.epochs; ++e) {
      std::shuffle(train_snaps.begin(), train_snaps.end(), rng);
      double tot = 0.0;
      int64_t count = 0;
      for (size_t c = 0; c < train_snaps.size(); c += size_t(bs)) {
        const int m = int(std::min<size_t>(size_t(bs), train_snaps.size() - c));
        const int64_t n = gather(train_snaps.data() + c, m);
        tot += batch_loss(n, true);
        count += n * P;
        launch_adam(s, w32, grad, am, av, float(p.learning_rate), ++adam_step, 1.0f / float(n * P));
      }
      if (loss_history) loss_history[e] = count ? tot / double(count) : 0.0;
      if (val_history) val_history[e] = evaluate(val_snaps);
    }
    SGPU_HIP(hipMemcpyAsync(weights_inout, w32, size_t(kPcWeights) * 4, hipMemcpyDeviceToHost, s));
    SGPU_HIP(hipStreamSynchronize(s));
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
