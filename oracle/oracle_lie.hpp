// ORACLE — test infrastructure only.  CPU FP64 restatement of the reference's
// SO(3)/SE(3), pinhole and forward-mode dual-number math.  Only tests/, the smoke
// check and the benchmark's CPU-baseline leg may load the oracle.
//
// Every function is generic over the scalar S (double, or Dual<N> for forward-mode
// Jacobians, the counterpart of DualScalar in
// crates/sophus_autodiff/src/dual/dual_scalar.rs:62).  Branch conditions compare
// real parts, as the reference's `select` does for dual scalars.
#pragma once

#include <cmath>
#include <cstdint>

namespace oracle {

constexpr double kEpsF64 = 1e-6;  // EPS_F64, crates/sophus_autodiff/src/linalg.rs:75

// ---- forward-mode dual numbers ---------------------------------------------------
template <int N>
struct Dual {
  double re = 0;
  double d[N] = {};
  Dual() = default;
  Dual(double r) : re(r) {}  // NOLINT: implicit from constants
  static Dual var(double r, int k) {
    Dual x(r);
    x.d[k] = 1.0;
    return x;
  }
};

template <int N> Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.re + b.re);
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int N> Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.re - b.re);
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <int N> Dual<N> operator-(const Dual<N>& a) {
  Dual<N> r(-a.re);
  for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
  return r;
}
template <int N> Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.re * b.re);
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.re + a.re * b.d[i];
  return r;
}
template <int N> Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(a.re / b.re);
  const double inv = 1.0 / b.re;
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.re * b.d[i]) * inv;
  return r;
}
template <int N> Dual<N> operator+(const Dual<N>& a, double b) { return a + Dual<N>(b); }
template <int N> Dual<N> operator+(double a, const Dual<N>& b) { return Dual<N>(a) + b; }
template <int N> Dual<N> operator-(const Dual<N>& a, double b) { return a - Dual<N>(b); }
template <int N> Dual<N> operator-(double a, const Dual<N>& b) { return Dual<N>(a) - b; }
template <int N> Dual<N> operator*(const Dual<N>& a, double b) { return a * Dual<N>(b); }
template <int N> Dual<N> operator*(double a, const Dual<N>& b) { return Dual<N>(a) * b; }
template <int N> Dual<N> operator/(const Dual<N>& a, double b) { return a / Dual<N>(b); }
template <int N> Dual<N> operator/(double a, const Dual<N>& b) { return Dual<N>(a) / b; }

template <int N> Dual<N> sqrt(const Dual<N>& a) {
  Dual<N> r(std::sqrt(a.re));
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] / (2.0 * r.re);
  return r;
}
template <int N> Dual<N> sin(const Dual<N>& a) {
  Dual<N> r(std::sin(a.re));
  const double c = std::cos(a.re);
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * c;
  return r;
}
template <int N> Dual<N> cos(const Dual<N>& a) {
  Dual<N> r(std::cos(a.re));
  const double s = -std::sin(a.re);
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * s;
  return r;
}
template <int N> Dual<N> atan2(const Dual<N>& y, const Dual<N>& x) {
  Dual<N> r(std::atan2(y.re, x.re));
  const double den = x.re * x.re + y.re * y.re;
  for (int i = 0; i < N; ++i) r.d[i] = (x.re * y.d[i] - y.re * x.d[i]) / den;
  return r;
}

inline double real(double x) { return x; }
template <int N> double real(const Dual<N>& x) { return x.re; }
using std::atan2;
using std::cos;
using std::sin;
using std::sqrt;

// ---- SO(3) -------------------------------------------------------------------------
// Row-major 3x3 matrices, quaternions (w, x, y, z).

template <typename S> void hat(const S w[3], S H[9]) {
  H[0] = S(0.0); H[1] = -w[2];   H[2] = w[1];
  H[3] = w[2];   H[4] = S(0.0);  H[5] = -w[0];
  H[6] = -w[1];  H[7] = w[0];    H[8] = S(0.0);
}

template <typename S> void mat3_mul(const S A[9], const S B[9], S C[9]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      S acc = A[3 * i] * B[j];
      acc = acc + A[3 * i + 1] * B[3 + j];
      acc = acc + A[3 * i + 2] * B[6 + j];
      C[3 * i + j] = acc;
    }
}

template <typename S> void mat3_vec(const S A[9], const S v[3], S out[3]) {
  for (int i = 0; i < 3; ++i) out[i] = A[3 * i] * v[0] + A[3 * i + 1] * v[1] + A[3 * i + 2] * v[2];
}

// Rotation3Impl::exp, crates/sophus_lie/src/groups/rotation3.rs:363-387
template <typename S> void so3_exp(const S w[3], S q[4]) {
  const S theta_sq = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  const S theta_po4 = theta_sq * theta_sq;
  const S theta = sqrt(theta_sq);
  const S half_theta = S(0.5) * theta;
  S imag, re;
  if (real(theta_sq) <= kEpsF64 * kEpsF64) {
    imag = S(0.5) - S(1.0 / 48.0) * theta_sq + S(1.0 / 3840.0) * theta_po4;
    re = S(1.0) - S(1.0 / 8.0) * theta_sq + S(1.0 / 384.0) * theta_po4;
  } else {
    imag = sin(half_theta) / theta;
    re = cos(half_theta);
  }
  q[0] = re;
  q[1] = imag * w[0];
  q[2] = imag * w[1];
  q[3] = imag * w[2];
}

// Rotation3Impl::log, rotation3.rs:389-411
template <typename S> void so3_log(const S q[4], S w[3]) {
  const S sn = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
  const S qw = q[0];
  S f;
  if (real(sn) <= kEpsF64 * kEpsF64) {
    const S w_sq = qw * qw;
    f = S(2.0) / qw - S(2.0 / 3.0) * sn / (w_sq * qw);
  } else {
    const S n = sqrt(sn);
    const S sign = real(qw) <= 0.0 ? S(-1.0) : S(1.0);
    const S atan_nbyw = sign * atan2(n, sign * qw);
    f = S(2.0) * atan_nbyw / n;
  }
  w[0] = q[1] * f;
  w[1] = q[2] * f;
  w[2] = q[3] * f;
}

// Rotation3Impl::matrix, rotation3.rs:449-489 (no normalization)
template <typename S> void so3_matrix(const S q[4], S R[9]) {
  const S r = q[0], x = q[1], y = q[2], z = q[3];
  const S two(2.0), one(1.0);
  const S two_x = two * x, two_y = two * y, two_z = two * z, two_r = two * r;
  const S two_xx = two_x * x, two_yy = two_y * y, two_zz = two_z * z;
  const S two_xy = two_x * y, two_xz = two_x * z, two_yz = two_y * z;
  const S two_rx = two_r * x, two_ry = two_r * y, two_rz = two_r * z;
  R[0] = one - (two_yy + two_zz);
  R[1] = two_xy - two_rz;
  R[2] = two_xz + two_ry;
  R[3] = two_xy + two_rz;
  R[4] = one - (two_xx + two_zz);
  R[5] = two_yz - two_rx;
  R[6] = two_xz - two_ry;
  R[7] = two_yz + two_rx;
  R[8] = one - (two_xx + two_yy);
}

// QuaternionImpl::mult, crates/sophus_lie/src/quaternion.rs:212-224
template <typename S> void quat_mult(const S a[4], const S b[4], S out[4]) {
  const S dot = a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
  const S re = a[0] * b[0] - dot;
  const S c0 = a[2] * b[3] - a[3] * b[2];
  const S c1 = a[3] * b[1] - a[1] * b[3];
  const S c2 = a[1] * b[2] - a[2] * b[1];
  out[0] = re;
  out[1] = b[1] * a[0] + a[1] * b[0] + c0;
  out[2] = b[2] * a[0] + a[2] * b[0] + c1;
  out[3] = b[3] * a[0] + a[3] * b[0] + c2;
}

// Rotation3Impl::mat_v, rotation3.rs:761-776
template <typename S> void so3_mat_v(const S w[3], S V[9]) {
  const S theta_sq = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  S H[9], H2[9];
  hat(w, H);
  mat3_mul(H, H, H2);
  if (real(theta_sq) <= kEpsF64) {
    for (int i = 0; i < 9; ++i) V[i] = H[i] * S(0.5);
  } else {
    const S theta = sqrt(theta_sq);
    const S a = (S(1.0) - cos(theta)) / theta_sq;
    const S b = (theta - sin(theta)) / (theta_sq * theta);
    for (int i = 0; i < 9; ++i) V[i] = H[i] * a + H2[i] * b;
  }
  for (int i = 0; i < 3; ++i) V[4 * i] = V[4 * i] + S(1.0);
}

// Rotation3Impl::mat_v_inverse, rotation3.rs:778-798
template <typename S> void so3_mat_v_inverse(const S w[3], S Vi[9]) {
  const S theta_sq = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  S H[9], H2[9];
  hat(w, H);
  mat3_mul(H, H, H2);
  S c;
  if (real(theta_sq) <= kEpsF64) {
    c = S(1.0 / 12.0);
  } else {
    const S theta = sqrt(theta_sq);
    const S half = S(0.5) * theta;
    c = (S(1.0) - (S(0.5) * theta * cos(half)) / sin(half)) / (theta * theta);
  }
  for (int i = 0; i < 9; ++i) Vi[i] = -(H[i] * S(0.5)) + H2[i] * c;
  for (int i = 0; i < 3; ++i) Vi[4 * i] = Vi[4 * i] + S(1.0);
}

// Rotation3Impl::inv_left_jacobian, rotation3.rs:640-656 (near-zero branch I + hat/2
// as written in the reference).
template <typename S> void so3_inv_left_jacobian(const S w[3], S J[9]) {
  const S theta_sq = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  S H[9], H2[9];
  hat(w, H);
  mat3_mul(H, H, H2);
  if (real(theta_sq) <= kEpsF64) {
    for (int i = 0; i < 9; ++i) J[i] = H[i] * S(0.5);
  } else {
    const S theta = sqrt(theta_sq);
    const S c = (S(1.0) / theta_sq) - (S(1.0) + cos(theta)) / (S(2.0) * theta * sin(theta));
    for (int i = 0; i < 9; ++i) J[i] = -(H[i] * S(0.5)) + H2[i] * c;
  }
  for (int i = 0; i < 3; ++i) J[4 * i] = J[4 * i] + S(1.0);
}

// Rotation3Impl::left_jacobian_of_translation, rotation3.rs:986-1031
template <typename S> void so3_left_jacobian_of_translation(const S w[3], const S u[3], S Q[9]) {
  const S theta_sq = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  S Uh[9], Wh[9], WU[9];
  hat(u, Uh);
  hat(w, Wh);
  mat3_mul(Wh, Uh, WU);
  if (real(theta_sq) <= kEpsF64) {
    for (int i = 0; i < 9; ++i) Q[i] = Uh[i] * S(0.5) + WU[i] * S(1.0 / 6.0);
    return;
  }
  S UW[9], WUW[9], WW[9], WWU[9], UWW[9], WUWW[9], WWUW[9];
  mat3_mul(Uh, Wh, UW);
  mat3_mul(WU, Wh, WUW);
  mat3_mul(Wh, Wh, WW);
  mat3_mul(WW, Uh, WWU);
  mat3_mul(UW, Wh, UWW);
  mat3_mul(WUW, Wh, WUWW);
  mat3_mul(WWU, Wh, WWUW);
  const S theta = sqrt(theta_sq);
  const S s = sin(theta), c = cos(theta);
  const S b = (theta - s) / (theta * theta * theta);
  const S cc = (theta * theta + S(2.0) * c - S(2.0)) / (S(2.0) * (theta * theta * theta * theta));
  const S d = (S(2.0) * theta - S(3.0) * s + theta * c) /
              (S(2.0) * (theta * theta * theta * theta * theta));
  for (int i = 0; i < 9; ++i)
    Q[i] = Uh[i] * S(0.5) + (WU[i] + UW[i] + WUW[i]) * b +
           (WWU[i] + UWW[i] - WUW[i] * S(3.0)) * cc + (WUWW[i] + WWUW[i]) * d;
}

// ---- SE(3): params (qw, qx, qy, qz, tx, ty, tz); tangent [omega; nu] ---------------

// AffineGroupTemplateImpl::inverse, crates/sophus_lie/src/groups/affine_group_template.rs:398-404
template <typename S> void se3_inverse(const S a[7], S out[7]) {
  S qc[4] = {a[0], -a[1], -a[2], -a[3]};
  S R[9], Rt[3];
  so3_matrix(qc, R);
  mat3_vec(R, a + 4, Rt);
  for (int i = 0; i < 4; ++i) out[i] = qc[i];
  for (int i = 0; i < 3; ++i) out[4 + i] = -Rt[i];
}

// AffineGroupTemplateImpl::group_mul, affine_group_template.rs:387-396
template <typename S> void se3_mul(const S a[7], const S b[7], S out[7]) {
  S q[4], R[9], f[3];
  quat_mult(a, b, q);
  so3_matrix(a, R);
  mat3_vec(R, b + 4, f);
  for (int i = 0; i < 4; ++i) out[i] = q[i];
  for (int i = 0; i < 3; ++i) out[4 + i] = f[i] + a[4 + i];
}

// AffineGroupTemplateImpl::transform, affine_group_template.rs:406-410
template <typename S> void se3_transform(const S a[7], const S p[3], S out[3]) {
  S R[9];
  so3_matrix(a, R);
  mat3_vec(R, p, out);
  for (int i = 0; i < 3; ++i) out[i] = out[i] + a[4 + i];
}

// AffineGroupTemplateImpl::exp, affine_group_template.rs:354-360
template <typename S> void se3_exp(const S t[6], S out[7]) {
  S q[4], V[9], tr[3];
  so3_exp(t, q);
  so3_mat_v(t, V);
  mat3_vec(V, t + 3, tr);
  for (int i = 0; i < 4; ++i) out[i] = q[i];
  for (int i = 0; i < 3; ++i) out[4 + i] = tr[i];
}

// AffineGroupTemplateImpl::log, affine_group_template.rs:362-369
template <typename S> void se3_log(const S a[7], S t[6]) {
  so3_log(a, t);
  S Vi[9];
  so3_mat_v_inverse(t, Vi);
  mat3_vec(Vi, a + 4, t + 3);
}

// AffineGroupTemplateImpl::adj, affine_group_template.rs:330-340 (row-major 6x6)
template <typename S> void se3_adj(const S a[7], S A[36]) {
  S R[9], H[9], HR[9];
  so3_matrix(a, R);
  hat(a + 4, H);
  mat3_mul(H, R, HR);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      A[6 * i + j] = R[3 * i + j];
      A[6 * i + j + 3] = S(0.0);
      A[6 * (i + 3) + j] = HR[3 * i + j];
      A[6 * (i + 3) + j + 3] = R[3 * i + j];
    }
}

// AffineGroupTemplateImpl::inv_left_jacobian, affine_group_template.rs:601-618
template <typename S> void se3_inv_left_jacobian(const S t[6], S J[36]) {
  S Ji[9], Vi[9], Q[9], VQ[9], B[9];
  so3_inv_left_jacobian(t, Ji);
  so3_mat_v_inverse(t, Vi);
  so3_left_jacobian_of_translation(t, t + 3, Q);
  mat3_mul(Vi, Q, VQ);
  mat3_mul(VQ, Ji, B);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      J[6 * i + j] = Ji[3 * i + j];
      J[6 * i + j + 3] = S(0.0);
      J[6 * (i + 3) + j] = -B[3 * i + j];
      J[6 * (i + 3) + j + 3] = Vi[3 * i + j];
    }
}

// ---- residuals ----------------------------------------------------------------------

// PinholeCameraReprojectionCostTerm::residual, pinhole_reprojection.rs:42-50:
// uv - K * proj(world_from_camera^-1 * x_w); the affine "distortion" of
// crates/sophus_sensor/src/distortions/affine.rs:50-61 on perspective
// crates/sophus_sensor/src/projections/perspective.rs:30-39.
template <typename S>
void pinhole_residual(const S intr[4], const S pose[7], const S xw[3], const double uv[2],
                      S r[2]) {
  S cw[7], xc[3];
  se3_inverse(pose, cw);
  se3_transform(cw, xw, xc);
  const S px = xc[0] / xc[2], py = xc[1] / xc[2];
  r[0] = S(uv[0]) - (px * intr[0] + intr[2]);
  r[1] = S(uv[1]) - (py * intr[1] + intr[3]);
}

// EnhancedUnifiedDistortionImpl::distort, crates/sophus_sensor/src/distortions/
// enhanced_unified.rs:60-95: (x, y) on the z = 1 plane, rho = sqrt(beta r^2 + 1),
// norm = alpha rho + 1 - alpha, pixel = (fx x / norm + cx, fy y / norm + cy).
template <typename S>
void eucm_distort(const S p[6], const S xy[2], S uv[2]) {
  const S r2 = xy[0] * xy[0] + xy[1] * xy[1];
  const S rho = sqrt(p[5] * r2 + S(1.0));
  const S norm = p[4] * rho + (S(1.0) - p[4]);
  uv[0] = p[0] * (xy[0] / norm) + p[2];
  uv[1] = p[1] * (xy[1] / norm) + p[3];
}

// EnhancedUnifiedDistortionImpl::undistort, enhanced_unified.rs:97-124.
template <typename S>
void eucm_undistort(const S p[6], const S uv[2], S xy[2]) {
  const S mx = (uv[0] - p[2]) / p[0], my = (uv[1] - p[3]) / p[1];
  const S r2 = mx * mx + my * my;
  const S gamma = S(1.0) - p[4];
  const S num = S(1.0) - p[4] * p[4] * p[5] * r2;
  const S den = p[4] * sqrt(S(1.0) - (p[4] - gamma) * p[5] * r2) + gamma;
  const S k = num / den;
  xy[0] = mx / k;
  xy[1] = my / k;
}

// EucmCameraReprojectionCostTerm::residual, example_problems/ba_problem.rs:58-66:
// uv - distort(params, proj(world_from_camera^-1 * x_w)) (camera.rs:94-99).
template <typename S>
void eucm_residual(const S intr[6], const S pose[7], const S xw[3], const double uv[2], S r[2]) {
  S cw[7], xc[3], xy[2], px[2];
  se3_inverse(pose, cw);
  se3_transform(cw, xw, xc);
  xy[0] = xc[0] / xc[2];
  xy[1] = xc[1] / xc[2];
  eucm_distort(intr, xy, px);
  r[0] = S(uv[0]) - px[0];
  r[1] = S(uv[1]) - px[1];
}

// PoseGraph3CostTerm::residual, pose_graph3.rs:33-39
template <typename S>
void pose_graph3_residual(const S m[7], const S n[7], const S mn[7], S r[6]) {
  S mi[7], mni[7], t1[7], t2[7];
  se3_inverse(m, mi);
  se3_inverse(mn, mni);
  se3_mul(mi, n, t1);
  se3_mul(t1, mni, t2);
  se3_log(t2, r);
}

}  // namespace oracle
