// SO(3) / SE(3) math used inside the kernels and by the host-side generator.
// Templated on the scalar (double for the parity path, float for the mixed path).
//
// Parameter layout follows the reference: SO(3) = unit-ish quaternion (w, x, y, z),
// SE(3) = (qw, qx, qy, qz, tx, ty, tz), tangent = [omega; nu] (rotation first).
// Quaternions are never renormalized (crates/sophus_lie/src/lie_group.rs:139-141),
// so the matrix and inverse formulas below are the reference's non-normalizing ones.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#define SGPU_HD __host__ __device__ __forceinline__

namespace sgpu {

// EPS_F64 (crates/sophus_autodiff/src/linalg.rs:75).
constexpr double kEps = 1e-6;

template <typename T>
SGPU_HD T sq(T x) { return x * x; }

// Rotation3Impl::exp (crates/sophus_lie/src/groups/rotation3.rs:363-387).
template <typename T>
SGPU_HD void so3_exp(const T w[3], T q[4]) {
  const T theta_sq = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  const T theta_po4 = theta_sq * theta_sq;
  const T theta = sqrt(theta_sq);
  const T half_theta = T(0.5) * theta;
  T imag, real;
  if (theta_sq <= T(kEps * kEps)) {
    imag = T(0.5) - T(1.0 / 48.0) * theta_sq + T(1.0 / 3840.0) * theta_po4;
    real = T(1.0) - T(1.0 / 8.0) * theta_sq + T(1.0 / 384.0) * theta_po4;
  } else {
    imag = sin(half_theta) / theta;
    real = cos(half_theta);
  }
  q[0] = real;
  q[1] = imag * w[0];
  q[2] = imag * w[1];
  q[3] = imag * w[2];
}

// Rotation3Impl::matrix (rotation3.rs:449-489), row-major R[3][3].
template <typename T>
SGPU_HD void so3_matrix(const T q[4], T R[9]) {
  const T r = q[0], x = q[1], y = q[2], z = q[3];
  const T tx = T(2) * x, ty = T(2) * y, tz = T(2) * z, tr = T(2) * r;
  const T txx = tx * x, tyy = ty * y, tzz = tz * z;
  const T txy = tx * y, txz = tx * z, tyz = ty * z;
  const T trx = tr * x, try_ = tr * y, trz = tr * z;
  R[0] = T(1) - (tyy + tzz);
  R[1] = txy - trz;
  R[2] = txz + try_;
  R[3] = txy + trz;
  R[4] = T(1) - (txx + tzz);
  R[5] = tyz - trx;
  R[6] = txz - try_;
  R[7] = tyz + trx;
  R[8] = T(1) - (txx + tyy);
}

// QuaternionImpl::mult (crates/sophus_lie/src/quaternion.rs:212-224).
template <typename T>
SGPU_HD void quat_mult(const T a[4], const T b[4], T out[4]) {
  const T re = a[0] * b[0] - (a[1] * b[1] + a[2] * b[2] + a[3] * b[3]);
  const T cx = a[2] * b[3] - a[3] * b[2];
  const T cy = a[3] * b[1] - a[1] * b[3];
  const T cz = a[1] * b[2] - a[2] * b[1];
  out[0] = re;
  out[1] = b[1] * a[0] + a[1] * b[0] + cx;
  out[2] = b[2] * a[0] + a[2] * b[0] + cy;
  out[3] = b[3] * a[0] + a[3] * b[0] + cz;
}

template <typename T>
SGPU_HD void mat3_vec(const T R[9], const T v[3], T out[3]) {
  out[0] = R[0] * v[0] + R[1] * v[1] + R[2] * v[2];
  out[1] = R[3] * v[0] + R[4] * v[1] + R[5] * v[2];
  out[2] = R[6] * v[0] + R[7] * v[1] + R[8] * v[2];
}

template <typename T>
SGPU_HD void mat3_mul(const T A[9], const T B[9], T C[9]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}

template <typename T>
SGPU_HD void hat3(const T w[3], T H[9]) {
  H[0] = T(0);  H[1] = -w[2]; H[2] = w[1];
  H[3] = w[2];  H[4] = T(0);  H[5] = -w[0];
  H[6] = -w[1]; H[7] = w[0];  H[8] = T(0);
}

// Rotation3Impl::mat_v (rotation3.rs:761-776).
template <typename T>
SGPU_HD void so3_mat_v(const T w[3], T V[9]) {
  const T theta_sq = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  T H[9], H2[9];
  hat3(w, H);
  mat3_mul(H, H, H2);
  if (theta_sq <= T(kEps)) {
    for (int i = 0; i < 9; ++i) V[i] = T(0.5) * H[i];
  } else {
    const T theta = sqrt(theta_sq);
    const T a = (T(1) - cos(theta)) / theta_sq;
    const T b = (theta - sin(theta)) / (theta_sq * theta);
    for (int i = 0; i < 9; ++i) V[i] = H[i] * a + H2[i] * b;
  }
  V[0] += T(1); V[4] += T(1); V[8] += T(1);
}

// Rotation3Impl::mat_v_inverse (rotation3.rs:778-798).
template <typename T>
SGPU_HD void so3_mat_v_inverse(const T w[3], T Vi[9]) {
  const T theta_sq = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  T H[9], H2[9];
  hat3(w, H);
  mat3_mul(H, H, H2);
  T c;
  if (theta_sq <= T(kEps)) {
    c = T(1.0 / 12.0);
  } else {
    const T theta = sqrt(theta_sq);
    const T half = T(0.5) * theta;
    c = (T(1) - (T(0.5) * theta * cos(half)) / sin(half)) / (theta * theta);
  }
  for (int i = 0; i < 9; ++i) Vi[i] = -H[i] * T(0.5) + H2[i] * c;
  Vi[0] += T(1); Vi[4] += T(1); Vi[8] += T(1);
}

// Rotation3Impl::log (rotation3.rs:389-411).
template <typename T>
SGPU_HD void so3_log(const T q[4], T w[3]) {
  const T sn = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
  const T qw = q[0];
  T f;
  if (sn <= T(kEps * kEps)) {
    const T w_sq = qw * qw;
    f = T(2) / qw - T(2.0 / 3.0) * sn / (w_sq * qw);
  } else {
    const T n = sqrt(sn);
    const T sign = qw <= T(0) ? T(-1) : T(1);
    const T atan_nbyw = sign * atan2(n, sign * qw);
    f = T(2) * atan_nbyw / n;
  }
  w[0] = q[1] * f;
  w[1] = q[2] * f;
  w[2] = q[3] * f;
}

// Rotation3Impl::inv_left_jacobian (rotation3.rs:640-656).  The near-zero branch
// is I + hat/2, exactly as the reference writes it.
template <typename T>
SGPU_HD void so3_inv_left_jacobian(const T w[3], T J[9]) {
  const T theta_sq = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  T H[9], H2[9];
  hat3(w, H);
  mat3_mul(H, H, H2);
  if (theta_sq <= T(kEps)) {
    for (int i = 0; i < 9; ++i) J[i] = H[i] * T(0.5);
  } else {
    const T theta = sqrt(theta_sq);
    const T c = (T(1) / theta_sq) - (T(1) + cos(theta)) / (T(2) * theta * sin(theta));
    for (int i = 0; i < 9; ++i) J[i] = -H[i] * T(0.5) + H2[i] * c;
  }
  J[0] += T(1); J[4] += T(1); J[8] += T(1);
}

// Rotation3Impl::left_jacobian_of_translation (rotation3.rs:986-1031).
template <typename T>
SGPU_HD void so3_left_jacobian_of_translation(const T w[3], const T u[3], T Q[9]) {
  const T theta_sq = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  T Uh[9], Wh[9], WU[9];
  hat3(u, Uh);
  hat3(w, Wh);
  mat3_mul(Wh, Uh, WU);
  if (theta_sq <= T(kEps)) {
    for (int i = 0; i < 9; ++i) Q[i] = Uh[i] * T(0.5) + WU[i] * T(1.0 / 6.0);
    return;
  }
  T UW[9], WUW[9], WW[9], WWU[9], UWW[9], WUWW[9], WWUW[9];
  mat3_mul(Uh, Wh, UW);
  mat3_mul(WU, Wh, WUW);
  mat3_mul(Wh, Wh, WW);
  mat3_mul(WW, Uh, WWU);
  mat3_mul(UW, Wh, UWW);
  mat3_mul(WUW, Wh, WUWW);
  mat3_mul(WWU, Wh, WWUW);
  const T theta = sqrt(theta_sq);
  const T s = sin(theta), c = cos(theta);
  const T a = T(0.5);
  const T b = (theta - s) / (theta * theta * theta);
  const T cc = (theta * theta + T(2) * c - T(2)) / (T(2) * (theta * theta * theta * theta));
  const T d = (T(2) * theta - T(3) * s + theta * c) /
              (T(2) * (theta * theta * theta * theta * theta));
  for (int i = 0; i < 9; ++i)
    Q[i] = Uh[i] * a + (WU[i] + UW[i] + WUW[i]) * b + (WWU[i] + UWW[i] - WUW[i] * T(3)) * cc +
           (WUWW[i] + WWUW[i]) * d;
}

// ---- SE(3) ------------------------------------------------------------------------

// AffineGroupTemplateImpl::inverse (affine_group_template.rs:398-404).
template <typename T>
SGPU_HD void se3_inverse(const T a[7], T out[7]) {
  T qc[4] = {a[0], -a[1], -a[2], -a[3]};
  T R[9];
  so3_matrix(qc, R);
  T t[3] = {a[4], a[5], a[6]}, Rt[3];
  mat3_vec(R, t, Rt);
  out[0] = qc[0]; out[1] = qc[1]; out[2] = qc[2]; out[3] = qc[3];
  out[4] = -Rt[0]; out[5] = -Rt[1]; out[6] = -Rt[2];
}

// AffineGroupTemplateImpl::group_mul (affine_group_template.rs:387-396).
template <typename T>
SGPU_HD void se3_mul(const T a[7], const T b[7], T out[7]) {
  T q[4];
  quat_mult(a, b, q);
  T R[9];
  so3_matrix(a, R);
  T t2[3] = {b[4], b[5], b[6]}, f[3];
  mat3_vec(R, t2, f);
  out[0] = q[0]; out[1] = q[1]; out[2] = q[2]; out[3] = q[3];
  out[4] = f[0] + a[4]; out[5] = f[1] + a[5]; out[6] = f[2] + a[6];
}

// AffineGroupTemplateImpl::exp (affine_group_template.rs:354-360).
template <typename T>
SGPU_HD void se3_exp(const T tangent[6], T out[7]) {
  T q[4], V[9], t[3];
  so3_exp(tangent, q);
  so3_mat_v(tangent, V);
  mat3_vec(V, tangent + 3, t);
  out[0] = q[0]; out[1] = q[1]; out[2] = q[2]; out[3] = q[3];
  out[4] = t[0]; out[5] = t[1]; out[6] = t[2];
}

// AffineGroupTemplateImpl::log (affine_group_template.rs:362-369).
template <typename T>
SGPU_HD void se3_log(const T a[7], T tangent[6]) {
  so3_log(a, tangent);
  T Vi[9];
  so3_mat_v_inverse(tangent, Vi);
  mat3_vec(Vi, a + 4, tangent + 3);
}

// LieGroup::oplus = exp(delta) * self (lie_group_manifold.rs:33-35).
template <typename T>
SGPU_HD void se3_oplus(const T a[7], const T delta[6], T out[7]) {
  T e[7];
  se3_exp(delta, e);
  se3_mul(e, a, out);
}

// AffineGroupTemplateImpl::adj (affine_group_template.rs:330-340), row-major 6x6.
template <typename T>
SGPU_HD void se3_adj(const T a[7], T A[36]) {
  T R[9], H[9], HR[9];
  so3_matrix(a, R);
  hat3(a + 4, H);
  mat3_mul(H, R, HR);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      A[6 * i + j] = R[3 * i + j];
      A[6 * i + j + 3] = T(0);
      A[6 * (i + 3) + j] = HR[3 * i + j];
      A[6 * (i + 3) + j + 3] = R[3 * i + j];
    }
}

// AffineGroupTemplateImpl::inv_left_jacobian (affine_group_template.rs:601-618).
template <typename T>
SGPU_HD void se3_inv_left_jacobian(const T tangent[6], T J[36]) {
  T Ji[9], Vi[9], Q[9], VQ[9], B[9];
  so3_inv_left_jacobian(tangent, Ji);
  so3_mat_v_inverse(tangent, Vi);
  so3_left_jacobian_of_translation(tangent, tangent + 3, Q);
  mat3_mul(Vi, Q, VQ);
  mat3_mul(VQ, Ji, B);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      J[6 * i + j] = Ji[3 * i + j];
      J[6 * i + j + 3] = T(0);
      J[6 * (i + 3) + j] = -B[3 * i + j];
      J[6 * (i + 3) + j + 3] = Vi[3 * i + j];
    }
}

}  // namespace sgpu
