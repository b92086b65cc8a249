// Seeded synthetic bundle-adjustment problems.
//
// Behavioural counterpart of the reference's example problems
// (crates/sophus_opt/src/example_problems/ba_problem.rs:154-381 and
// crates/sophus_opt/src/example_problems/cam_calib.rs:63-151): a pinhole camera
// (fx, fy, cx, cy) shared by all observations, SE(3) world_from_camera poses, 3-D
// points, pixel observations with Gaussian noise and perturbed initial values.
//
// RNG: SplitMix64 seeds a xoshiro256** generator; normals come from Box-Muller.
// The draw order is fixed (ground truth, observations, pose perturbations, point
// perturbations), so one seed gives one bit-identical problem on every machine.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "api_internal.hpp"
#include "lie.hpp"

namespace sgpu {
namespace {

uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

struct Rng {
  uint64_t s[4];
  bool has_spare = false;
  double spare = 0.0;
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto& v : s) v = splitmix64(x);
  }
  uint64_t next() {
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
  }
  // [0, 1)
  double uniform() { return double(next() >> 11) * 0x1.0p-53; }
  double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
  // [0, n)
  int64_t below(int64_t n) { return int64_t(uniform() * double(n)); }
  double normal() {
    if (has_spare) {
      has_spare = false;
      return spare;
    }
    double u1 = 1.0 - uniform();  // (0, 1]
    double u2 = uniform();
    double r = std::sqrt(-2.0 * std::log(u1));
    spare = r * std::sin(2.0 * M_PI * u2);
    has_spare = true;
    return r * std::cos(2.0 * M_PI * u2);
  }
};

void normalize3(double v[3]) {
  double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  v[0] /= n;
  v[1] /= n;
  v[2] /= n;
}

void cross3(const double a[3], const double b[3], double c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// Unit quaternion (w, x, y, z) of a row-major rotation matrix (Shepperd's method).
void quat_from_matrix(const double R[9], double q[4]) {
  const double tr = R[0] + R[4] + R[8];
  if (tr > 0) {
    double s = std::sqrt(tr + 1.0) * 2;
    q[0] = 0.25 * s;
    q[1] = (R[7] - R[5]) / s;
    q[2] = (R[2] - R[6]) / s;
    q[3] = (R[3] - R[1]) / s;
  } else if (R[0] > R[4] && R[0] > R[8]) {
    double s = std::sqrt(1.0 + R[0] - R[4] - R[8]) * 2;
    q[0] = (R[7] - R[5]) / s;
    q[1] = 0.25 * s;
    q[2] = (R[1] + R[3]) / s;
    q[3] = (R[2] + R[6]) / s;
  } else if (R[4] > R[8]) {
    double s = std::sqrt(1.0 + R[4] - R[0] - R[8]) * 2;
    q[0] = (R[2] - R[6]) / s;
    q[1] = (R[1] + R[3]) / s;
    q[2] = 0.25 * s;
    q[3] = (R[5] + R[7]) / s;
  } else {
    double s = std::sqrt(1.0 + R[8] - R[0] - R[4]) * 2;
    q[0] = (R[3] - R[1]) / s;
    q[1] = (R[2] + R[6]) / s;
    q[2] = (R[5] + R[7]) / s;
    q[3] = 0.25 * s;
  }
  double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  for (int i = 0; i < 4; ++i) q[i] /= n;
}

// world_from_camera looking along `forward` with world +z up; camera x right,
// y down, z forward.
void look_along(const double pos[3], const double fwd_in[3], double pose[7]) {
  double f[3] = {fwd_in[0], fwd_in[1], fwd_in[2]};
  normalize3(f);
  const double up[3] = {0, 0, 1};
  double right[3], down[3];
  cross3(f, up, right);
  normalize3(right);
  cross3(f, right, down);
  const double R[9] = {right[0], down[0], f[0], right[1], down[1], f[1],
                       right[2], down[2], f[2]};
  quat_from_matrix(R, pose);
  pose[4] = pos[0];
  pose[5] = pos[1];
  pose[6] = pos[2];
}

struct Problem {
  double intrinsics[4];
  std::vector<double> poses_gt, poses_init, points_gt, points_init, uv;
  std::vector<int64_t> idx;
};

constexpr double kWidth = 640, kHeight = 480;

// Project x_w with camera `pose` (world_from_camera); returns false if behind
// or outside the image.
bool project(const double K[4], const double pose[7], const double xw[3], double uv[2],
             double min_depth) {
  double cw[7];
  se3_inverse(pose, cw);
  double R[9], xc[3];
  so3_matrix(cw, R);
  mat3_vec(R, xw, xc);
  xc[0] += cw[4];
  xc[1] += cw[5];
  xc[2] += cw[6];
  if (xc[2] <= min_depth) return false;
  uv[0] = xc[0] / xc[2] * K[0] + K[2];
  uv[1] = xc[1] / xc[2] * K[1] + K[3];
  return uv[0] >= 0 && uv[0] < kWidth && uv[1] >= 0 && uv[1] < kHeight;
}

// Pose perturbation exp(d) * T_gt (ring: cameras 2 m from the origin) or, for the
// walk, T_gt * exp(d): rotation about the camera's own centre, so that cameras far
// from the world origin get the same image-space perturbation as near ones.
void perturb(Rng& rng, const sgpu_ba_spec& s, Problem& p, int64_t ncam, int64_t npt) {
  p.poses_init = p.poses_gt;
  for (int64_t c = 1; c < ncam; ++c) {
    double d[6];
    for (int k = 0; k < 3; ++k) d[k] = s.rot_noise * rng.normal();
    for (int k = 3; k < 6; ++k) d[k] = s.trans_noise * rng.normal();
    if (s.layout == 0) {
      se3_oplus(&p.poses_gt[7 * c], d, &p.poses_init[7 * c]);
    } else {
      double e[7];
      se3_exp(d, e);
      se3_mul(&p.poses_gt[7 * c], e, &p.poses_init[7 * c]);
    }
  }
  p.points_init = p.points_gt;
  for (int64_t i = 0; i < 3 * npt; ++i) p.points_init[i] += s.point_noise * rng.normal();
}

void gen_ring(const sgpu_ba_spec& s, Problem& p) {
  Rng rng(s.seed);
  const int64_t ncam = s.num_cameras, npt = s.num_points;
  p.poses_gt.resize(7 * ncam);
  for (int64_t j = 0; j < ncam; ++j) {
    const double th = 2.0 * M_PI * double(j) / double(ncam);
    const double pos[3] = {2.0 * std::cos(th), 2.0 * std::sin(th), 0.0};
    const double fwd[3] = {-pos[0], -pos[1], -pos[2]};
    look_along(pos, fwd, &p.poses_gt[7 * j]);
  }
  p.points_gt.resize(3 * npt);
  for (auto& v : p.points_gt) v = rng.uniform(-0.5, 0.5);
  std::vector<int64_t> cams(ncam);
  for (int64_t b = 0; b < npt; ++b) {
    // a fixed count, or (obs_per_point_ring == 0) a uniform draw in [min_obs, max_obs]
    int k = s.obs_per_point_ring;
    if (k == 0) k = int(std::min<int64_t>(ncam, s.min_obs + rng.below(s.max_obs - s.min_obs + 1)));
    for (int64_t c = 0; c < ncam; ++c) cams[c] = c;
    for (int j = 0; j < k; ++j) {  // partial Fisher-Yates: k distinct cameras
      int64_t r = j + rng.below(ncam - j);
      std::swap(cams[j], cams[r]);
      const int64_t c = cams[j];
      double uv[2];
      project(p.intrinsics, &p.poses_gt[7 * c], &p.points_gt[3 * b], uv, -1e300);
      p.idx.insert(p.idx.end(), {0, c, b});
      p.uv.push_back(uv[0] + s.pixel_noise * rng.normal());
      p.uv.push_back(uv[1] + s.pixel_noise * rng.normal());
    }
  }
  perturb(rng, s, p, ncam, npt);
}

void gen_walk(const sgpu_ba_spec& s, Problem& p) {
  Rng rng(s.seed);
  const int64_t ncam = s.num_cameras, npt = s.num_points;
  p.poses_gt.resize(7 * ncam);
  double pos[3] = {0, 0, 0}, psi = 0.0;
  for (int64_t j = 0; j < ncam; ++j) {
    const double fwd[3] = {std::cos(psi), std::sin(psi), 0.0};
    look_along(pos, fwd, &p.poses_gt[7 * j]);
    psi += s.yaw_jitter * rng.normal();
    pos[0] += s.step_length * std::cos(psi);
    pos[1] += s.step_length * std::sin(psi);
  }
  p.points_gt.resize(3 * npt);
  const double log1mp = std::log(1.0 - s.geometric_p);
  std::vector<int64_t> seen;
  std::vector<double> seen_uv;
  for (int64_t b = 0; b < npt; ++b) {
    for (int attempt = 0;; ++attempt) {
      const int64_t a = rng.below(ncam);
      const double u = rng.uniform(0, kWidth), v = rng.uniform(0, kHeight);
      const double depth = rng.uniform(2.0, 10.0);
      double xc[3] = {depth * (u - p.intrinsics[2]) / p.intrinsics[0],
                      depth * (v - p.intrinsics[3]) / p.intrinsics[1], depth};
      const double* T = &p.poses_gt[7 * a];
      double R[9], xw[3];
      so3_matrix(T, R);
      mat3_vec(R, xc, xw);
      xw[0] += T[4];
      xw[1] += T[5];
      xw[2] += T[6];
      // geometric draw on {1, 2, ...}, clipped
      double g = std::ceil(std::log(1.0 - rng.uniform()) / log1mp);
      int64_t want = g < 1 ? 1 : int64_t(g);
      if (want < s.min_obs) want = s.min_obs;
      if (want > s.max_obs) want = s.max_obs;
      seen.clear();
      seen_uv.clear();
      // nearest-in-time cameras: a, a-1, a+1, a-2, a+2, ...
      const int64_t window = 4 * s.max_obs + 16;
      for (int64_t d = 0; d <= window && int64_t(seen.size()) < want; ++d) {
        for (int sgn = -1; sgn <= 1 && int64_t(seen.size()) < want; sgn += 2) {
          if (d == 0 && sgn == 1) break;
          const int64_t c = a + sgn * d;
          if (c < 0 || c >= ncam) continue;
          double uv[2];
          if (project(p.intrinsics, &p.poses_gt[7 * c], xw, uv, 0.5)) {
            seen.push_back(c);
            seen_uv.push_back(uv[0]);
            seen_uv.push_back(uv[1]);
          }
        }
      }
      if (int64_t(seen.size()) >= s.min_obs || attempt >= 64) {
        if (seen.empty()) continue;
        p.points_gt[3 * b] = xw[0];
        p.points_gt[3 * b + 1] = xw[1];
        p.points_gt[3 * b + 2] = xw[2];
        for (size_t j = 0; j < seen.size(); ++j) {
          p.idx.insert(p.idx.end(), {0, seen[j], b});
          p.uv.push_back(seen_uv[2 * j] + s.pixel_noise * rng.normal());
          p.uv.push_back(seen_uv[2 * j + 1] + s.pixel_noise * rng.normal());
        }
        break;
      }
    }
  }
  perturb(rng, s, p, ncam, npt);
}

}  // namespace
}  // namespace sgpu

extern "C" sgpu_status sgpu_generate_ba(const sgpu_ba_spec* spec, sgpu_ba_sizes* sizes,
                                        double* intrinsics, double* poses_gt,
                                        double* poses_init, double* points_gt,
                                        double* points_init, int64_t* obs_idx,
                                        double* obs_uv) {
  using namespace sgpu;
  clear_error();
  if (!spec || !sizes) return set_error(SGPU_ERR_INVALID_ARGUMENT, "spec and sizes required");
  if (spec->num_cameras < 2 || spec->num_points < 1)
    return set_error(SGPU_ERR_INVALID_ARGUMENT, "need >= 2 cameras and >= 1 point");
  if (spec->layout == 0 &&
      (spec->obs_per_point_ring < 0 || spec->obs_per_point_ring > spec->num_cameras))
    return set_error(SGPU_ERR_INVALID_ARGUMENT, "obs_per_point_ring out of range");
  if (spec->layout == 0 && spec->obs_per_point_ring == 0 &&
      (spec->min_obs < 1 || spec->max_obs < spec->min_obs))
    return set_error(SGPU_ERR_INVALID_ARGUMENT, "bad per-point observation range");
  if (spec->layout == 1 && (spec->geometric_p <= 0 || spec->geometric_p >= 1 ||
                            spec->min_obs < 1 || spec->max_obs < spec->min_obs))
    return set_error(SGPU_ERR_INVALID_ARGUMENT, "bad geometric observation spec");
  if (spec->layout != 0 && spec->layout != 1)
    return set_error(SGPU_ERR_INVALID_ARGUMENT, "unknown layout");
  Problem p;
  p.intrinsics[0] = 500.0;
  p.intrinsics[1] = 500.0;
  p.intrinsics[2] = 320.0;
  p.intrinsics[3] = 240.0;
  if (spec->layout == 0)
    gen_ring(*spec, p);
  else
    gen_walk(*spec, p);
  sizes->num_cameras = spec->num_cameras;
  sizes->num_points = spec->num_points;
  sizes->num_obs = int64_t(p.uv.size() / 2);
  auto put = [](double* dst, const std::vector<double>& v) {
    if (dst) std::memcpy(dst, v.data(), v.size() * sizeof(double));
  };
  if (intrinsics) std::memcpy(intrinsics, p.intrinsics, sizeof(p.intrinsics));
  put(poses_gt, p.poses_gt);
  put(poses_init, p.poses_init);
  put(points_gt, p.points_gt);
  put(points_init, p.points_init);
  put(obs_uv, p.uv);
  if (obs_idx) std::memcpy(obs_idx, p.idx.data(), p.idx.size() * sizeof(int64_t));
  return SGPU_OK;
}

// ---- EUCM ring around an open box ------------------------------------------------------------
// The reference's own BaProblem (crates/sophus_opt/src/example_problems/ba_problem.rs:154-345)
// with this file's RNG in place of ChaCha12: num_cameras on a ring of radius 8 in the x-z
// plane at height 0, each looking at (0, 0, 10); points on the five closed faces of the box
// x in [-3, 3], y in [-2, 2], z in [8, 12] (point k on face k % 5); one shared EUCM camera
// (300, 300, 320, 240, 0.5, 1.0) for a 640 x 480 image; a camera sees a point within 8 m
// (any distance below 20 cameras) in front of it (z > 0.1) and inside the image; pixel noise
// 0.15 U(-0.5, 0.5); points seen fewer than twice are dropped.  Initial values: intrinsics
// (297, 297, 321, 241, 0.495, 1.0), poses 2.. perturbed by exp(d) T with d uniform in
// +-(0.04, 0.04, 0.02, 0.03, 0.03, 0.01) (poses 0 and 1 are the gauge), points by
// U(-0.2, 0.2) per coordinate.
namespace sgpu {
namespace {

struct EucmProblem {
  double intr_gt[6] = {300.0, 300.0, 320.0, 240.0, 0.5, 1.0};
  double intr_init[6] = {297.0, 297.0, 321.0, 241.0, 0.495, 1.0};
  std::vector<double> poses_gt, poses_init, points_gt, points_init, uv;
  std::vector<int64_t> idx;
};

void eucm_pixel(const double p[6], const double xc[3], double uv[2]) {
  const double x = xc[0] / xc[2], y = xc[1] / xc[2];
  const double rho = std::sqrt(p[5] * (x * x + y * y) + 1.0);
  const double norm = p[4] * rho + (1.0 - p[4]);
  uv[0] = p[0] * (x / norm) + p[2];
  uv[1] = p[1] * (y / norm) + p[3];
}

void gen_eucm_ring_box(int32_t ncam, int64_t npt_all, uint64_t seed, EucmProblem& p) {
  constexpr double kRadius = 8.0, kCenterZ = 10.0;
  Rng rng(seed);
  p.poses_gt.resize(7 * size_t(ncam));
  for (int32_t j = 0; j < ncam; ++j) {
    const double th = 2.0 * M_PI * double(j) / double(ncam);
    double* T = &p.poses_gt[7 * size_t(j)];
    T[0] = std::cos(-0.5 * th);  // rotation about y by -theta
    T[1] = 0.0;
    T[2] = std::sin(-0.5 * th);
    T[3] = 0.0;
    T[4] = kRadius * std::sin(th);
    T[5] = 0.0;
    T[6] = kCenterZ - kRadius * std::cos(th);
  }
  const double bw = 3.0, bh = 2.0, zn = 8.0, zf = 12.0;
  std::vector<double> pts(3 * size_t(npt_all));
  for (int64_t k = 0; k < npt_all; ++k) {
    double* x = &pts[3 * size_t(k)];
    switch (k % 5) {
      case 0:
        x[0] = (rng.uniform() - 0.5) * 2.0 * bw;
        x[1] = (rng.uniform() - 0.5) * 2.0 * bh;
        x[2] = zf;
        break;
      case 1:
      case 2:
        x[0] = k % 5 == 1 ? -bw : bw;
        x[1] = (rng.uniform() - 0.5) * 2.0 * bh;
        x[2] = zn + rng.uniform() * (zf - zn);
        break;
      default:
        x[0] = (rng.uniform() - 0.5) * 2.0 * bw;
        x[1] = k % 5 == 3 ? bh : -bh;
        x[2] = zn + rng.uniform() * (zf - zn);
        break;
    }
  }
  const double max_dist = ncam >= 20 ? kRadius : INFINITY;
  std::vector<int64_t> raw_idx;
  std::vector<double> raw_uv;
  std::vector<int32_t> count(npt_all, 0);
  for (int32_t j = 0; j < ncam; ++j) {
    const double* T = &p.poses_gt[7 * size_t(j)];
    double cw[7], R[9];
    se3_inverse(T, cw);
    so3_matrix(cw, R);
    for (int64_t k = 0; k < npt_all; ++k) {
      const double* x = &pts[3 * size_t(k)];
      const double dx = T[4] - x[0], dy = T[5] - x[1], dz = T[6] - x[2];
      if (std::sqrt(dx * dx + dy * dy + dz * dz) > max_dist) continue;
      double xc[3];
      mat3_vec(R, x, xc);
      for (int e = 0; e < 3; ++e) xc[e] += cw[4 + e];
      if (xc[2] <= 0.1) continue;
      double uv[2];
      eucm_pixel(p.intr_gt, xc, uv);
      if (uv[0] < 0.0 || uv[0] >= kWidth || uv[1] < 0.0 || uv[1] >= kHeight) continue;
      const double n0 = 0.15 * (rng.uniform() - 0.5), n1 = 0.15 * (rng.uniform() - 0.5);
      raw_idx.insert(raw_idx.end(), {0, j, k});
      raw_uv.insert(raw_uv.end(), {uv[0] + n0, uv[1] + n1});
      ++count[k];
    }
  }
  std::vector<int64_t> remap(npt_all, -1);
  int64_t npt = 0;
  for (int64_t k = 0; k < npt_all; ++k)
    if (count[k] >= 2) {
      remap[k] = npt++;
      p.points_gt.insert(p.points_gt.end(), &pts[3 * size_t(k)], &pts[3 * size_t(k)] + 3);
    }
  for (size_t o = 0; o < raw_uv.size() / 2; ++o) {
    const int64_t nk = remap[raw_idx[3 * o + 2]];
    if (nk < 0) continue;
    p.idx.insert(p.idx.end(), {0, raw_idx[3 * o + 1], nk});
    p.uv.insert(p.uv.end(), {raw_uv[2 * o], raw_uv[2 * o + 1]});
  }
  p.poses_init = p.poses_gt;
  const double amp[6] = {0.08, 0.08, 0.04, 0.06, 0.06, 0.02};
  for (int32_t j = 2; j < ncam; ++j) {
    double d[6], e[7];
    for (int k = 0; k < 6; ++k) d[k] = (rng.uniform() - 0.5) * amp[k];
    se3_exp(d, e);
    se3_mul(e, &p.poses_gt[7 * size_t(j)], &p.poses_init[7 * size_t(j)]);
  }
  p.points_init = p.points_gt;
  for (double& v : p.points_init) v += (rng.uniform() - 0.5) * 0.4;
}

}  // namespace
}  // namespace sgpu

extern "C" sgpu_status sgpu_generate_ba_eucm(int32_t num_cameras, int64_t num_points, uint64_t seed,
                                             sgpu_ba_sizes* sizes, double* intrinsics_gt,
                                             double* intrinsics_init, double* poses_gt,
                                             double* poses_init, double* points_gt,
                                             double* points_init, int64_t* obs_idx,
                                             double* obs_uv) {
  using namespace sgpu;
  return guarded([&] {
    if (!sizes || num_cameras < 2 || num_points < 1 || num_points > (int64_t(1) << 31))
      throw Error(SGPU_ERR_INVALID_ARGUMENT, "bad EUCM problem size");
    EucmProblem p;
    gen_eucm_ring_box(num_cameras, num_points, seed, p);
    sizes->num_cameras = num_cameras;
    sizes->num_points = int64_t(p.points_gt.size() / 3);
    sizes->num_obs = int64_t(p.uv.size() / 2);
    if (!intrinsics_gt) return;
    if (!intrinsics_init || !poses_gt || !poses_init || !points_gt || !points_init || !obs_idx || !obs_uv)
      throw Error(SGPU_ERR_INVALID_ARGUMENT, "output arrays required");
    std::memcpy(intrinsics_gt, p.intr_gt, sizeof(p.intr_gt));
    std::memcpy(intrinsics_init, p.intr_init, sizeof(p.intr_init));
    std::memcpy(poses_gt, p.poses_gt.data(), p.poses_gt.size() * 8);
    std::memcpy(poses_init, p.poses_init.data(), p.poses_init.size() * 8);
    std::memcpy(points_gt, p.points_gt.data(), p.points_gt.size() * 8);
    std::memcpy(points_init, p.points_init.data(), p.points_init.size() * 8);
    std::memcpy(obs_idx, p.idx.data(), p.idx.size() * 8);
    std::memcpy(obs_uv, p.uv.data(), p.uv.size() * 8);
  });
}
