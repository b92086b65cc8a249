// gfx950 kernels of one Schur-complement LM iteration.
//
// Per observation o of point b seen by camera c (pinhole_reprojection.rs:60-95):
//   A_o = dx_distort_x * dx_proj_x(x_c) * R_cw       (2x3)
//   J_pose = A_o [hat(-x_b) | I] = A_o P_b,   J_point = -A_o
// so every per-term block of the reference (evaluated_term.rs:286-327) is a 3-space
// quantity lifted by the per-point P_b:
//   M_o = A_o^T A_o,   v_o = A_o^T r~_o
//   H_pp += P^T M_o P,  H_pl = -M_o P,  H_ll += M_o,  neg_grad_p -= P^T v_o,  neg_grad_l += v_o.
// The Schur complement (schur_ldlt.rs:82-182) then reads
//   S(c_i, c_j) -= P^T (M_i H_ll^-1 M_j) P,   rhs(c_i) += P^T (M_i H_ll^-1 s_b),
// and back-substitution (schur_ldlt.rs:785-806) dx_b = H_ll^-1 (s_b + sum_i M_i P dx_{c_i}).
// Kernels only ever store 3-space data per point; the 6x6 lifts are formed in registers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "api_internal.hpp"
#include "block6.hpp"
#include "kernels.hpp"
#include "lie.hpp"
#include "dual.hpp"

namespace sgpu {

namespace {

constexpr int kThreads = 256;
constexpr int kMaxRowSlots = 112;  // S row slots accumulated in LDS (else global atomics)

__device__ __forceinline__ void atomic_add_d(double* p, double v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add_f(float* p, float v) { unsafeAtomicAdd(p, v); }
template <typename T>
__device__ __forceinline__ void lds_add(T* p, T v) {
  unsafeAtomicAdd(p, v);
}

// max |v| over the wavefront, then one atomic per wavefront (every lane of the wavefront
// calls it); a per-lane atomic on one address would serialise one instruction per lane group
__device__ __forceinline__ void wave_atomic_max_abs(unsigned long long* p, double v) {
  v = fabs(v);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  if ((threadIdx.x & 63) == 0 && v > 0.0) atomicMax(p, (unsigned long long)__double_as_longlong(v));
}

// Huber weight sqrt(rho(|r|)) / |r| (robust_kernel.rs:13-35); 1 without a kernel.
__device__ __forceinline__ double robust_weight(int robust, double delta, double nrm) {
  if (robust != 1) return 1.0;
  if (nrm == 0.0) return 0.0;
  const double a = fabs(nrm) <= delta ? nrm * nrm : 2.0 * delta * fabs(nrm) - delta * delta;
  return sqrt(a) / nrm;
}

// base[i] when ok, else 0; the load is unconditional (from base[0] when !ok, which the
// caller keeps valid) so a panel's loads issue back to back without branches
__device__ __forceinline__ double load_or_zero(const double* base, int64_t i, bool ok) {
  const double v = base[ok ? i : 0];
  return ok ? v : 0.0;
}

// Copies n (<= N * blockDim.x) values to LDS with every global load of the thread issued
// before its first LDS store, so the copy waits for memory once instead of once per
// element: load(e, ok) returns element e (0 when !ok, from a valid address), put(e, v)
// stores it.
template <int N, typename Load, typename Put>
__device__ __forceinline__ void stage_lds(int n, Load load, Put put) {
  double v[N];
#pragma unroll
  for (int it = 0; it < N; ++it) {
    const int e = int(threadIdx.x) + it * int(blockDim.x);
    v[it] = load(e < n ? e : 0, e < n);
  }
#pragma unroll
  for (int it = 0; it < N; ++it) {
    const int e = int(threadIdx.x) + it * int(blockDim.x);
    if (e < n) put(e, v[it]);
  }
}

// stage_lds for two arrays of n values each, with the loads of both issued first.
template <int N, typename LoadA, typename PutA, typename LoadB, typename PutB>
__device__ __forceinline__ void stage_lds2(int n, LoadA load_a, PutA put_a, LoadB load_b,
                                           PutB put_b) {
  double va[N], vb[N];
#pragma unroll
  for (int it = 0; it < N; ++it) {
    const int e = int(threadIdx.x) + it * int(blockDim.x);
    va[it] = load_a(e < n ? e : 0, e < n);
    vb[it] = load_b(e < n ? e : 0, e < n);
  }
#pragma unroll
  for (int it = 0; it < N; ++it) {
    const int e = int(threadIdx.x) + it * int(blockDim.x);
    if (e < n) {
      put_a(e, va[it]);
      put_b(e, vb[it]);
    }
  }
}

// camera_from_world cache: R_cw (row-major) + t_cw.
__global__ void k_cam_cache(int npose, const double* __restrict__ poses, double* __restrict__ cw) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npose) return;
  double p[7], inv[7], R[9];
  for (int k = 0; k < 7; ++k) p[k] = poses[7 * i + k];
  se3_inverse(p, inv);
  so3_matrix(inv, R);
  for (int k = 0; k < 9; ++k) cw[12 * i + k] = R[k];
  cw[12 * i + 9] = inv[4];
  cw[12 * i + 10] = inv[5];
  cw[12 * i + 11] = inv[6];
}

__device__ __forceinline__ void to_camera(const double* __restrict__ cw, const double x[3],
                                          double xc[3]) {
#pragma clang fp contract(off)
  xc[0] = cw[0] * x[0] + cw[1] * x[1] + cw[2] * x[2] + cw[9];
  xc[1] = cw[3] * x[0] + cw[4] * x[1] + cw[5] * x[2] + cw[10];
  xc[2] = cw[6] * x[0] + cw[7] * x[1] + cw[8] * x[2] + cw[11];
}

// A = diag(fx, fy) * dx_proj_x(x_c) * R_cw  (perspective.rs:54-72, affine.rs:77-85).
// Contraction is off here and in gram(): the point pass, the camera pass and the
// back-substitution each evaluate the same observation's A and M = A^T A, and the mixed
// path relies on those values being bitwise equal in all three (H_ll - M_o cancels
// exactly only then).
template <typename T>
__device__ __forceinline__ void obs_jacobian(const double* __restrict__ cw,
                                             const double* __restrict__ in, const double xc[3],
                                             T A[6]) {
#pragma clang fp contract(off)
  const T z = T(xc[2]);
  const T iz = T(1) / z;
  const T zz = z * z;
  const T d00 = T(in[0]) * iz, d02 = T(in[0]) * (-T(xc[0]) / zz);
  const T d11 = T(in[1]) * iz, d12 = T(in[1]) * (-T(xc[1]) / zz);
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    A[j] = d00 * T(cw[j]) + d02 * T(cw[6 + j]);
    A[3 + j] = d11 * T(cw[3 + j]) + d12 * T(cw[6 + j]);
  }
}

// M = A^T A (symmetric 3x3, stored full row-major)
template <typename T>
__device__ __forceinline__ void gram(const T A[6], T M[9]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) M[3 * i + j] = A[i] * A[j] + A[3 + i] * A[3 + j];
}

// residual r = uv - K proj(x_c) in FP64 and the weighted v = A^T (w r).
template <typename T>
__device__ __forceinline__ void obs_residual(const double* __restrict__ in,
                                             const double* __restrict__ uv, const double xc[3],
                                             int robust, double delta, const T A[6], T v[3]) {
  const double px = xc[0] / xc[2], py = xc[1] / xc[2];
  const double r0 = uv[0] - (px * in[0] + in[2]);
  const double r1 = uv[1] - (py * in[1] + in[3]);
  const double w = robust_weight(robust, delta, sqrt(r0 * r0 + r1 * r1));
  const T rw0 = T(robust == 1 ? w * r0 : r0), rw1 = T(robust == 1 ? w * r1 : r1);
#pragma unroll
  for (int j = 0; j < 3; ++j) v[j] = A[j] * rw0 + A[3 + j] * rw1;
}

// ---- camera models ----------------------------------------------------------------------
// The reprojection kernels are templated on the camera model: the pinhole (above) or the
// enhanced unified camera model (EUCM).  Only A_o = d_pi R_cw and the predicted pixel depend
// on it; everything downstream is the same 3-space algebra.
constexpr int kPinhole = 0, kEucm = 1;
template <int kModel>
__host__ __device__ constexpr int intr_stride() {
  return kModel == kEucm ? 6 : 4;
}

// EUCM on the perspective projection (enhanced_unified.rs:60-215, perspective.rs:30-72,
// ba_problem.rs:76-101): the pixel of x_c, D = dx_distort_x(proj) dx_proj_x(x_c) (2x3
// row-major) and, with E non-null, dx_distort_params(proj) (2x6 row-major; J_intr = -E).
// Contraction is off for the reason given at obs_jacobian.
__device__ __forceinline__ void eucm_project(const double* __restrict__ in, const double xc[3],
                                             double px[2], double D[6], double* E) {
#pragma clang fp contract(off)
  const double x = xc[0] / xc[2], y = xc[1] / xc[2];
  const double fx = in[0], fy = in[1], alpha = in[4], beta = in[5];
  const double r2 = x * x + y * y;
  const double rho = sqrt(beta * r2 + 1.0);
  const double norm = alpha * rho + (1.0 - alpha);
  const double mx = x / norm, my = y / norm;
  px[0] = fx * mx + in[2];
  px[1] = fy * my + in[3];
  const double nn = norm * norm;
  const double dnx = alpha * ((2.0 * beta * x) / (2.0 * rho));
  const double dny = alpha * ((2.0 * beta * y) / (2.0 * rho));
  const double d00 = fx * ((norm - x * dnx) / nn), d01 = fx * (-x * dny / nn);
  const double d10 = fy * (-y * dnx / nn), d11 = fy * ((norm - y * dny) / nn);
  const double iz = 1.0 / xc[2], zz = xc[2] * xc[2];
  const double p02 = -xc[0] / zz, p12 = -xc[1] / zz;
  D[0] = d00 * iz;
  D[1] = d01 * iz;
  D[2] = d00 * p02 + d01 * p12;
  D[3] = d10 * iz;
  D[4] = d11 * iz;
  D[5] = d10 * p02 + d11 * p12;
  if (E) {
    const double dna = rho - 1.0, dnb = alpha * (r2 / (2.0 * rho));
    E[0] = mx;
    E[1] = 0.0;
    E[2] = 1.0;
    E[3] = 0.0;
    E[4] = fx * (-x * dna / nn);
    E[5] = fx * (-x * dnb / nn);
    E[6] = 0.0;
    E[7] = my;
    E[8] = 0.0;
    E[9] = 1.0;
    E[10] = fy * (-y * dna / nn);
    E[11] = fy * (-y * dnb / nn);
  }
}

// A = D R_cw (2x3) from the projection derivative D and the cached rotation.
__device__ __forceinline__ void lift_rotation(const double* __restrict__ cw, const double D[6],
                                              double A[6]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      A[3 * i + j] = D[3 * i] * cw[j] + D[3 * i + 1] * cw[3 + j] + D[3 * i + 2] * cw[6 + j];
}

template <int kModel, typename T>
__device__ __forceinline__ void model_jacobian(const double* __restrict__ cw,
                                               const double* __restrict__ in, const double xc[3],
                                               T A[6]) {
  if constexpr (kModel == kPinhole) {
    obs_jacobian<T>(cw, in, xc, A);
  } else {
    double px[2], D[6], Ad[6];
    eucm_project(in, xc, px, D, nullptr);
    lift_rotation(cw, D, Ad);
#pragma unroll
    for (int k = 0; k < 6; ++k) A[k] = T(Ad[k]);
  }
}

// predicted pixel of x_c
template <int kModel>
__device__ __forceinline__ void model_pixel(const double* __restrict__ in, const double xc[3],
                                            double px[2]) {
  if constexpr (kModel == kPinhole) {
    px[0] = xc[0] / xc[2] * in[0] + in[2];
    px[1] = xc[1] / xc[2] * in[1] + in[3];
  } else {
    double D[6];
    eucm_project(in, xc, px, D, nullptr);
  }
}

// residual r = uv - pi(x_c) in FP64 and v = A^T (w r)
template <int kModel, typename T>
__device__ __forceinline__ void model_residual(const double* __restrict__ in,
                                               const double* __restrict__ uv, const double xc[3],
                                               int robust, double delta, const T A[6], T v[3]) {
  if constexpr (kModel == kPinhole) {
    obs_residual<T>(in, uv, xc, robust, delta, A, v);
  } else {
    double px[2];
    model_pixel<kModel>(in, xc, px);
    const double r0 = uv[0] - px[0], r1 = uv[1] - px[1];
    const double w = robust_weight(robust, delta, sqrt(r0 * r0 + r1 * r1));
    const T rw0 = T(robust == 1 ? w * r0 : r0), rw1 = T(robust == 1 ? w * r1 : r1);
#pragma unroll
    for (int j = 0; j < 3; ++j) v[j] = A[j] * rw0 + A[3 + j] * rw1;
  }
}

// ---- kernel: per-point H_ll, its inverse and the point gradient --------------------------
template <typename T, int kModel>
__global__ void __launch_bounds__(kThreads)
    k_point_prep(DevProblem p, const double* __restrict__ points, const double* __restrict__ intr,
                 const double* __restrict__ cw, double nu, double* __restrict__ s_out,
                 double* __restrict__ prec, DevScalars* sc) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = gid < p.npt;
  const int b = live ? gid : p.npt - 1;  // trailing lanes stay for the wavefront max
  const double x[3] = {points[3 * b], points[3 * b + 1], points[3 * b + 2]};
  const bool active = p.pt_block[b] >= 0;
  // H_ll and s are summed in FP64 in both precisions (of FP32 terms in the mixed path), so
  // that H_ll - M_o in the camera pass is exactly the other terms' sum plus the damping and
  // each point's Schur contribution stays positive semi-definite however small nu gets
  double H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double s[3] = {0, 0, 0};
  for (int o = p.pt_ptr[b]; live && active && o < p.pt_ptr[b + 1]; ++o) {
    const double* c = cw + 12 * p.obs_pose[o];
    const double* in = intr + intr_stride<kModel>() * p.obs_intr[o];
    double xc[3];
    to_camera(c, x, xc);
    T A[6], v[3], M[9];
    model_jacobian<kModel, T>(c, in, xc, A);
    model_residual<kModel, T>(in, p.obs_uv + 2 * o, xc, p.robust, p.delta, A, v);
    gram(A, M);
#pragma unroll
    for (int k = 0; k < 9; ++k) H[k] += double(M[k]);
#pragma unroll
    for (int k = 0; k < 3; ++k) s[k] += double(v[k]);
  }
  {
    double g = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) g = fmax(g, fabs(s[k]));
    wave_atomic_max_abs(&sc->max_grad_bits, live && active ? g : 0.0);
  }
  if (!live) return;
  double* rec = prec + kPtRec * int64_t(b);
  if (!active) {  // constant point: H_ll^-1, y and H_ll are 0 in its record
#pragma unroll
    for (int k = 0; k < kPtRec; ++k) rec[k] = k < 3 ? x[k] : 0.0;
    return;
  }
  const double damp = nu * double(p.pt_nu[b]);
  H[0] += damp;
  H[4] += damp;
  H[8] += damp;
  // closed-form 3x3 inverse; det == 0 is a singular marginal block (schur_ldlt.rs:101)
  const double m11 = H[0], m12 = H[1], m13 = H[2], m21 = H[3], m22 = H[4], m23 = H[5],
               m31 = H[6], m32 = H[7], m33 = H[8];
  const double a = m22 * m33 - m32 * m23, bb = m21 * m33 - m31 * m23, c = m21 * m32 - m31 * m22;
  const double det = m11 * a - m12 * bb + m13 * c;
  double inv[9];
  if (det == 0.0 || !isfinite(det)) {
    atomicOr(&sc->singular_marg, 1);
    for (int k = 0; k < 9; ++k) inv[k] = 0.0;
  } else {
    const double rdet = 1.0 / det;  // one division, then products
    inv[0] = a * rdet;
    inv[1] = (m13 * m32 - m33 * m12) * rdet;
    inv[2] = (m12 * m23 - m22 * m13) * rdet;
    inv[3] = -bb * rdet;
    inv[4] = (m11 * m33 - m31 * m13) * rdet;
    inv[5] = (m13 * m21 - m23 * m11) * rdet;
    inv[6] = c * rdet;
    inv[7] = (m12 * m31 - m32 * m11) * rdet;
    inv[8] = (m11 * m22 - m21 * m12) * rdet;
  }
  double yv[3];
  for (int k = 0; k < 3; ++k) yv[k] = inv[3 * k] * s[0] + inv[3 * k + 1] * s[1] + inv[3 * k + 2] * s[2];
  const double rv[kPtRec] = {x[0],   x[1],   x[2],   inv[0], inv[1], inv[2], inv[4], inv[5], inv[8],
                             yv[0],  yv[1],  yv[2],  m11,    m12,    m13,    m22,    m23,    m33};
#pragma unroll
  for (int k = 0; k < kPtRec; ++k) rec[k] = rv[k];
  for (int k = 0; k < 3; ++k) s_out[3 * b + k] = s[k];
}

// P_i^T N P_j for P = [hat(-x) | I] and a 3x3 N:
//   [[H_i N H_j^T, H_i N], [N H_j^T, N]]  with H = hat(x).
// Adds sign * (P_i^T N P_j) to a 6x6 row-major accumulator via `add(idx, value)`.
template <typename T, typename Add>
__device__ __forceinline__ void add_lift(const T xi[3], const T xj[3], const T N[9], T sign,
                                         Add add) {
  // HN = hat(x_i) N ; NHt = N hat(x_j)^T ; HNHt = hat(x_i) N hat(x_j)^T
  T HN[9], NHt[9], HNHt[9];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    // hat(x) rows: [0,-x2,x1], [x2,0,-x0], [-x1,x0,0]
    HN[j] = -xi[2] * N[3 + j] + xi[1] * N[6 + j];
    HN[3 + j] = xi[2] * N[j] - xi[0] * N[6 + j];
    HN[6 + j] = -xi[1] * N[j] + xi[0] * N[3 + j];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    // (N hat(x_j)^T)[i][j] = sum_t N[i][t] hat(x_j)[j][t]
    NHt[3 * i + 0] = -N[3 * i + 1] * xj[2] + N[3 * i + 2] * xj[1];
    NHt[3 * i + 1] = N[3 * i + 0] * xj[2] - N[3 * i + 2] * xj[0];
    NHt[3 * i + 2] = -N[3 * i + 0] * xj[1] + N[3 * i + 1] * xj[0];
    HNHt[3 * i + 0] = -HN[3 * i + 1] * xj[2] + HN[3 * i + 2] * xj[1];
    HNHt[3 * i + 1] = HN[3 * i + 0] * xj[2] - HN[3 * i + 2] * xj[0];
    HNHt[3 * i + 2] = -HN[3 * i + 0] * xj[1] + HN[3 * i + 1] * xj[0];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      add(6 * i + j, sign * HNHt[3 * i + j]);
      add(6 * i + 3 + j, sign * HN[3 * i + j]);
      add(6 * (i + 3) + j, sign * NHt[3 * i + j]);
      add(6 * (i + 3) + 3 + j, sign * N[3 * i + j]);
    }
}

// World position of a camera from its camera_from_world cache: c = -R_cw^T t_cw.
__device__ __forceinline__ void cam_center(const double* __restrict__ cw, double c[3]) {
#pragma unroll
  for (int k = 0; k < 3; ++k) c[k] = -(cw[k] * cw[9] + cw[3 + k] * cw[10] + cw[6 + k] * cw[11]);
}

// (G^T G)[a][b] for G = [[I, 0], [hat(c), I]]:  [[|c|^2 I - c c^T + I, hat(c)^T], [hat(c), I]].
__device__ __forceinline__ double centered_gram(const double c[3], int a, int b) {
  if (a < 3 && b < 3) return (a == b ? 1.0 + c[0] * c[0] + c[1] * c[1] + c[2] * c[2] : 0.0) - c[a] * c[b];
  if (a >= 3 && b >= 3) return a == b ? 1.0 : 0.0;
  // hat(c)[r][q]
  auto hatc = [&](int r, int q) -> double {
    const double h[9] = {0.0, -c[2], c[1], c[2], 0.0, -c[0], -c[1], c[0], 0.0};
    return h[3 * r + q];
  };
  return a < 3 ? hatc(b - 3, a) : hatc(a - 3, b);
}

template <typename T>
__device__ __forceinline__ void mat3_mul_t(const T A[9], const T B[9], T C[9]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}

// ---- the reduced camera system ----------------------------------------------------------
// Block (c1, c2) of S sums over the (obs_i, obs_j) pairs of the points seen by both
// cameras (cam(obs_i) = c1, cam(obs_j) = c2):
//   i != j:  -P_i^T (M_i H_ll^-1 M_j) P_j
//   i == j:  +P^T (M_i H_ll^-1 (H_ll - M_i)) P      (= P^T (M_i - M_i H_ll^-1 M_i) P, written
//                                                   without the cancellation; M_i alone
//                                                   for a constant point)
// and the (o, o) terms also carry the camera's gradient -P^T v_o and its reduced rhs
// P^T (M_o y_b - v_o).
// With H_ll^-1 = L L^T (3x3 Cholesky of the point's inverse) and U_o = L^T M_o,
//   M_i H_ll^-1 M_j = U_i^T U_j,
// so a cross pair needs one 12-value record per observation (U_o and its lift point)
// instead of both observations' Jacobians.  k_cam_obs forms the records (FP64, stored
// as T) for every observation of a free camera in camera-major order and sums each
// camera's (o, o) terms per chunk of 256 observations; k_schur_blocks then reduces each
// block's cross pairs and, on the diagonal, the camera's chunk sums.  Every sum runs in
// a fixed order (no atomics).
// With `centered`, each camera's tangent is taken about its own centre c
// (delta = G delta', G = [[I, 0], [hat(c), I]]): P becomes [hat(c - x) | I], whose entries
// are scene-depth sized instead of world-coordinate sized, so FP32 products keep their
// relative accuracy in S' = G^T S G and rhs' = G^T rhs.  The gradient stays in the
// original coordinates.

constexpr int kObsRec = 12;  // per-observation record: U_o (3x3 row-major), lift point

// Sums v[0..N) over the wavefront.  Halving steps (while the live length is even) leave
// each lane with a slice of the totals; the remaining masks are a plain butterfly.
// Returns the slice offset: lane holds the totals of v[off .. off + len) in v[0 .. len).
template <int N>
__device__ __forceinline__ int wave_sum_scatter(double (&v)[N], int lane) {
  int len = N, off = 0;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    if (len % 2 == 0) {
      const int h = len / 2;
      const bool up = (lane & m) != 0;
#pragma unroll
      for (int k = 0; k < N / 2; ++k) {
        if (k < h) {
          const double send = up ? v[k] : v[k + h];
          const double keep = up ? v[k + h] : v[k];
          v[k] = keep + __shfl_xor(send, m, 64);
        }
      }
      len = h;
      off += up ? h : 0;
    } else {
#pragma unroll
      for (int k = 0; k < N; ++k)
        if (k < len) v[k] += __shfl_xor(v[k], m, 64);
    }
  }
  return off;
}

template <typename T, int kModel>
__global__ void __launch_bounds__(kThreads)
    k_cam_obs(DevProblem p, const double* __restrict__ points, const double* __restrict__ intr,
              const double* __restrict__ cw, const double* __restrict__ prec, int centered,
              T* __restrict__ vobs,
              double* __restrict__ own_part) {
  constexpr int kOwn = 48;  // 36 block entries, 6 reduced rhs, 6 gradient
  __shared__ T sv[kThreads * kObsRec];
  __shared__ double red[kThreads / 64][kOwn];
  const int ch = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int cam = p.own_chunks[3 * ch], q0 = p.own_chunks[3 * ch + 1], q1 = p.own_chunks[3 * ch + 2];
  const int q = q0 + tid;
  double own[kOwn];
#pragma unroll
  for (int e = 0; e < kOwn; ++e) own[e] = 0.0;
  if (q < q1) {
    // camera-major observation fields; one camera per chunk
    const int b = p.cm_point[q];
    const double* c = cw + 12 * p.cam_member[cam];
    const double* ini = intr + intr_stride<kModel>() * p.cm_intr[q];
    const double* uvq = p.cm_uv + 2 * int64_t(q);
    T* vs = sv + kObsRec * tid;
    double xd[3], xl[3], xc[3], Ad[6], Md[9], v[3], t[3];
    if constexpr (std::is_same<T, double>::value) {
      // packed point record: x, H_ll^-1 upper, y, H_ll upper (zeros but x for a constant
      // point)
      double rc[kPtRec];
#pragma unroll
      for (int k = 0; k < kPtRec; ++k) rc[k] = prec[kPtRec * int64_t(b) + k];
#pragma unroll
      for (int e = 0; e < 3; ++e) xd[e] = xl[e] = rc[e];
      if (centered) {
        double cc[3];
        cam_center(c, cc);
#pragma unroll
        for (int e = 0; e < 3; ++e) xl[e] -= cc[e];
      }
      to_camera(c, xd, xc);
      model_jacobian<kModel, double>(c, ini, xc, Ad);
      gram(Ad, Md);
      model_residual<kModel, double>(ini, uvq, xc, p.robust, p.delta, Ad, v);
      // U = L^T M with H_ll^-1 = L L^T (zero columns past a non-positive pivot)
      const double l00 = sqrt(fmax(rc[3], 0.0)), i00 = l00 > 0.0 ? 1.0 / l00 : 0.0;
      const double l10 = rc[4] * i00, l20 = rc[5] * i00;
      const double l11 = sqrt(fmax(rc[6] - l10 * l10, 0.0)), i11 = l11 > 0.0 ? 1.0 / l11 : 0.0;
      const double l21 = (rc[7] - l20 * l10) * i11;
      const double l22 = sqrt(fmax(rc[8] - l20 * l20 - l21 * l21, 0.0));
      double U[9];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        U[j] = l00 * Md[j] + l10 * Md[3 + j] + l20 * Md[6 + j];
        U[3 + j] = l11 * Md[3 + j] + l21 * Md[6 + j];
        U[6 + j] = l22 * Md[6 + j];
      }
#pragma unroll
      for (int k = 0; k < 9; ++k) vs[k] = U[k];
#pragma unroll
      for (int e = 0; e < 3; ++e) vs[9 + e] = xl[e];
      // own term: M - M H_ll^-1 M = M H_ll^-1 (H_ll - M), without the cancellation (M
      // alone for a constant point);  t = M y - v
      double N[9];
      if (p.pt_block[b] >= 0) {
        const double* Hu = rc + 12;
        const double Hf[9] = {Hu[0], Hu[1], Hu[2], Hu[1], Hu[3], Hu[4], Hu[2], Hu[4], Hu[5]};
        const double Hv[9] = {rc[3], rc[4], rc[5], rc[4], rc[6], rc[7], rc[5], rc[7], rc[8]};
        double R[9], Z[9];
#pragma unroll
        for (int e = 0; e < 9; ++e) R[e] = Hf[e] - Md[e];
        mat3_mul_t(Md, Hv, Z);
        mat3_mul_t(Z, R, N);
      } else {
#pragma unroll
        for (int e = 0; e < 9; ++e) N[e] = Md[e];
      }
#pragma unroll
      for (int r = 0; r < 3; ++r) t[r] = Md[3 * r] * rc[9] + Md[3 * r + 1] * rc[10] + Md[3 * r + 2] * rc[11] - v[r];
      add_lift(xl, xl, N, 1.0, [&](int e, double val) { own[e] = val; });
    } else {
      const bool active = p.pt_block[b] >= 0;
#pragma unroll
      for (int e = 0; e < 3; ++e) xd[e] = xl[e] = points[3 * b + e];
      if (centered) {
        double cc[3];
        cam_center(c, cc);
#pragma unroll
        for (int e = 0; e < 3; ++e) xl[e] -= cc[e];
      }
      to_camera(c, xd, xc);
      model_jacobian<kModel, double>(c, ini, xc, Ad);
      model_residual<kModel, double>(ini, uvq, xc, p.robust, p.delta, Ad, v);
      // M_o as the point pass and the back-substitution form it (FP32 Jacobian and Gram),
      // held in FP64: H_ll sums exactly these terms, so R = H_ll - M_o below is the other
      // terms plus the damping, and U, the reduced rhs and the own term all describe one
      // matrix
      T Ai[6], Mi[9];
      model_jacobian<kModel, T>(c, ini, xc, Ai);
      gram(Ai, Mi);
#pragma unroll
      for (int e = 0; e < 9; ++e) Md[e] = double(Mi[e]);
      // record: U_o row-major, then the lift point (zero for a constant point: it has no
      // cross pairs)
      // the point's packed record: H_ll^-1 upper at 3, y at 9, damped H_ll upper at 12
      const double* rb = prec + kPtRec * int64_t(b);
      if (active) {
        const double* h = rb + 3;
        const double l00 = sqrt(fmax(h[0], 0.0)), i00 = l00 > 0.0 ? 1.0 / l00 : 0.0;
        const double l10 = h[1] * i00, l20 = h[2] * i00;
        const double l11 = sqrt(fmax(h[3] - l10 * l10, 0.0)), i11 = l11 > 0.0 ? 1.0 / l11 : 0.0;
        const double l21 = (h[4] - l20 * l10) * i11;
        const double l22 = sqrt(fmax(h[5] - l20 * l20 - l21 * l21, 0.0));
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          vs[j] = T(l00 * Md[j] + l10 * Md[3 + j] + l20 * Md[6 + j]);
          vs[3 + j] = T(l11 * Md[3 + j] + l21 * Md[6 + j]);
          vs[6 + j] = T(l22 * Md[6 + j]);
          vs[9 + j] = T(xl[j]);
        }
      } else {
#pragma unroll
        for (int e = 0; e < kObsRec; ++e) vs[e] = T(0);
      }
      // own (o, o) term in T with the FP64 difference R = H_ll - M_o:
      // N = M_o H_ll^-1 R (M_o alone for a constant point)
      const T x[3] = {T(xl[0]), T(xl[1]), T(xl[2])};
      T N[9];
      if (active) {
        const double* yb = rb + 9;
#pragma unroll
        for (int r = 0; r < 3; ++r) t[r] = Md[3 * r] * yb[0] + Md[3 * r + 1] * yb[1] + Md[3 * r + 2] * yb[2] - v[r];
        const double* H = rb + 12;
        const double Hf[9] = {H[0], H[1], H[2], H[1], H[3], H[4], H[2], H[4], H[5]};
        const double* h = rb + 3;
        const double Hv[9] = {h[0], h[1], h[2], h[1], h[3], h[4], h[2], h[4], h[5]};
        T R[9], Hi[9], Z[9];
#pragma unroll
        for (int e = 0; e < 9; ++e) {
          R[e] = T(Hf[e] - Md[e]);
          Hi[e] = T(Hv[e]);
        }
        mat3_mul_t(Mi, Hi, Z);
        mat3_mul_t(Z, R, N);
      } else {
#pragma unroll
        for (int r = 0; r < 3; ++r) t[r] = -v[r];
#pragma unroll
        for (int e = 0; e < 9; ++e) N[e] = Mi[e];
      }
      add_lift(x, x, N, T(1), [&](int e, T val) { own[e] = double(val); });
    }
    // P^T t = [x cross t; t]
    own[36] = xl[1] * t[2] - xl[2] * t[1];
    own[37] = xl[2] * t[0] - xl[0] * t[2];
    own[38] = xl[0] * t[1] - xl[1] * t[0];
    own[39] = t[0];
    own[40] = t[1];
    own[41] = t[2];
    own[42] = -(xd[1] * v[2] - xd[2] * v[1]);
    own[43] = -(xd[2] * v[0] - xd[0] * v[2]);
    own[44] = -(xd[0] * v[1] - xd[1] * v[0]);
    own[45] = -v[0];
    own[46] = -v[1];
    own[47] = -v[2];
  }
  __syncthreads();
  // the chunk's factors are contiguous in camera-major order: coalesced stores
  T* dst = vobs + kObsRec * int64_t(q0);
  for (int e = tid; e < kObsRec * (q1 - q0); e += kThreads) dst[e] = sv[e];
  const int off = wave_sum_scatter(own, lane);
  // 48 = 16 * 3: after four halvings each lane holds 3 totals, shared by 4 lanes
  if ((lane & 3) == 0)
#pragma unroll
    for (int k = 0; k < 3; ++k) red[wid][off + k] = own[k];
  __syncthreads();
  if (tid < kOwn) {
    double sum = 0.0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) sum += red[w][tid];
    own_part[kOwn * int64_t(ch) + tid] = sum;
  }
}

// acc -= P_i^T (U_i^T U_j) P_j (6x6 row-major) from two observation records.  With one lift
// point (the records' points are the same point, not camera-centred), only U is read from
// the second record.
template <bool kOneLift, typename T>
__device__ __forceinline__ void sub_pair(const T* __restrict__ ri, const T* __restrict__ rj,
                                         double acc[36]) {
  constexpr int kB = kOneLift ? 9 : kObsRec;
  T a[kObsRec], b[kObsRec];
#pragma unroll
  for (int e = 0; e < kObsRec; ++e) a[e] = ri[e];
#pragma unroll
  for (int e = 0; e < kB; ++e) b[e] = rj[e];
  T N[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int q = 0; q < 3; ++q) N[3 * r + q] = a[r] * b[q] + a[3 + r] * b[3 + q] + a[6 + r] * b[6 + q];
  add_lift(a + 9, (kOneLift ? a : b) + 9, N, T(-1), [&](int e, T val) { acc[e] += double(val); });
}

// Workgroups are dispatched round-robin over the 8 XCDs; this remaps them so each XCD
// takes one contiguous range of logical workgroups, and neighbouring work items that
// share inputs meet in one XCD's L2.
__device__ __forceinline__ int xcd_contiguous(int bid, int n) {
  const int x = bid & 7, i = bid >> 3, q = n >> 3, r = n & 7;
  return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// One wavefront per S block: lanes stride over the block's cross pairs (point order),
// FP64 accumulation, a fixed-order wavefront reduction; on the diagonal the camera's
// chunk sums, the damping and the rhs / gradient rows.
template <typename T, bool kCentered>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(3)))
    k_schur_blocks(DevProblem p, const T* __restrict__ vobs, const double* __restrict__ own_part,
                   const double* __restrict__ cw, double nu,
                   double* __restrict__ front, double* __restrict__ rhs_nat,
                   double* __restrict__ grad_cam, double* __restrict__ bsr) {
  const int lane = threadIdx.x & 63;
  const int64_t k = int64_t(xcd_contiguous(blockIdx.x, gridDim.x)) * (blockDim.x >> 6) +
                    (threadIdx.x >> 6);
  if (k >= p.nslot) return;  // whole wavefront leaves together
  const int c1 = p.s_row[k];
  const bool diag = p.s_col[k] == c1;
  double acc[36];
#pragma unroll
  for (int e = 0; e < 36; ++e) acc[e] = 0.0;
  // a pair's two records hold the same point, so without centring one lift point serves both
  if constexpr (kCentered) {
#pragma unroll 2
    for (int64_t q = p.pair_ptr[k] + lane; q < p.pair_ptr[k + 1]; q += 64) {
      const int i = p.pairs[2 * q], j = p.pairs[2 * q + 1];
      sub_pair<false>(vobs + kObsRec * int64_t(i), vobs + kObsRec * int64_t(j), acc);
    }
  } else {
#pragma unroll 2
    for (int64_t q = p.pair_ptr[k] + lane; q < p.pair_ptr[k + 1]; q += 64) {
      const int i = p.pairs[2 * q], j = p.pairs[2 * q + 1];
      sub_pair<true>(vobs + kObsRec * int64_t(i), vobs + kObsRec * int64_t(j), acc);
    }
  }
  // 36 = 4 * 9: after two halvings lanes 16 m .. 16 m + 15 hold totals 9 m .. 9 m + 8
  wave_sum_scatter(acc, lane);
  double mine = 0.0;
  const int src = 16 * (min(lane, 35) / 9), want = min(lane, 35) % 9;
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    const double t = __shfl(acc[e], src, 64);
    if (want == e) mine = t;
  }
  if (lane >= 36) mine = 0.0;
  if (diag && lane < 48) {
    for (int ch = p.cam_chunk_ptr[c1]; ch < p.cam_chunk_ptr[c1 + 1]; ++ch)
      mine += own_part[48 * int64_t(ch) + lane];
  }
  if (lane < 36) {
    const int a = lane / 6, bb = lane % 6;
    if (diag && kCentered) {
      double c[3];
      cam_center(cw + 12 * p.cam_member[c1], c);
      mine += nu * double(p.cam_nu[c1]) * centered_gram(c, a, bb);
    } else if (diag && a == bb) {
      mine += nu * double(p.cam_nu[c1]);
    }
    if (bsr) {  // row-major lower block in slot order (iterative solvers)
      bsr[36 * k + lane] = mine;
    } else {
      const int ldc = p.s_ld[k], ld = ldc >> 1;
      front[p.s_off[k] + ((ldc & 1) ? a * ld + bb : bb * ld + a)] = mine;
    }
  } else if (diag && lane < 42) {
    rhs_nat[6 * c1 + lane - 36] = mine;
  } else if (diag && lane < 48) {
    grad_cam[6 * c1 + lane - 42] = mine;
  }
}

__global__ void k_fill(double* p, int64_t n, double v) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

__global__ void k_zero_scalars(DevScalars* sc) {
  if (threadIdx.x == 0) {
    sc->cost = 0;
    sc->dd = 0;
    sc->dg = 0;
    sc->max_delta_bits = 0;
    sc->max_grad_bits = 0;
    sc->singular_marg = 0;
    sc->bad_pivot = 0;
    sc->pcg_iterations = 0;
  }
}


// ---- supernodal multifrontal LDLt ---------------------------------------------------------
// Fronts are dense column-major F x F (F = 6 * front rows); the first K = 6 * node columns
// are eliminated with 6x6 pivots (the reference's block LDLt, block_sparse_ldlt.rs:614-747):
// for pivot j: (L_jj, d_j) = ldlt6(C_jj); L_ij = C_ij L_jj^-T D^-1 L_jj^-1 for rows i > j;
// trailing C_ik -= L_ij C_kj^T.  Panels of kPanelBlocks pivots are factored by one
// workgroup per node; the rank-24 trailing update runs as 64x64 tiles over many workgroups.

// x <- x L^-T D^-1 L^-1 for a 6x6 pivot (unit lower L, dinv = reciprocal pivots, 0 for a
// zero pivot)
__device__ __forceinline__ void right_solve6(const double* L, const double* dinv, double x[6]) {
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    double v = x[c];
    for (int t = 0; t < c; ++t) v -= x[t] * L[6 * c + t];
    x[c] = v;
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) x[c] *= dinv[c];
#pragma unroll
  for (int c = 5; c >= 0; --c) {
    double v = x[c];
    for (int t = c + 1; t < 6; ++t) v -= x[t] * L[6 * t + c];
    x[c] = v;
  }
}

constexpr int kPW = 6 * kPanelBlocks;

// View a [kPW][kTile + 1] LDS tile as a [kPW][kPW + 1] square (kPW <= kTile).
static_assert(kPW <= kTile, "the next diagonal block lies inside tile (0, 0)");
__device__ __forceinline__ double (*as_square(double (*t)[kTile + 1]))[kPW + 1] {
  return reinterpret_cast<double (*)[kPW + 1]>(&t[0][0]);
}

// Trailing update of pivot JB of a diagonal block: the rows below the next pivot block,
// four threads per row (thread t3 of 192), each forming the row's L from its W and
// updating every fourth column; the updated values of the next pivot's columns are also
// kept in Wd (the next pivot's W).  The pivot index is a template argument so that the
// column loop has a fixed length and its LDS reads can be issued together.
template <int JB>
__device__ __forceinline__ void diag_trailing(int t3, int np, const double* cL, const double* cdi,
                                              double (*D)[kPW + 1], double (*Wd)[kPW + 1]) {
  constexpr int J0 = 6 * JB, J1 = J0 + 6, kSplit = 4;
  const int rr = t3 / kSplit, cg = t3 % kSplit;
  const int r = J1 + 6 + rr;
  if (J1 + 6 >= kPW || r >= np) return;
  double x[6];
#pragma unroll
  for (int t = 0; t < 6; ++t) x[t] = Wd[r][J0 + t];
  right_solve6(cL, cdi, x);
  if (cg == 0)
#pragma unroll
    for (int t = 0; t < 6; ++t) D[r][J0 + t] = x[t];
  constexpr int kCols = J1 < kPW ? (kPW - 1 - J1) / kSplit + 1 : 0;
#pragma unroll
  for (int i = 0; i < kCols; ++i) {
    const int c = J1 + cg + kSplit * i;
    if (c <= r) {
      double acc = 0.0;
#pragma unroll
      for (int t = 0; t < 6; ++t) acc += x[t] * Wd[c][J0 + t];
      const double v = D[r][c] - acc;
      D[r][c] = v;
      if (i < 2 && c < J1 + 6) Wd[r][c] = v;
    }
  }
}

// Row block JB of T = L^-1 for the unit block-lower L held below the pivot blocks of D
// (its column blocks < JB are final), kept transposed in D's free upper triangle:
// D[c][r] = T[r][c] for block(c) < block(r).  Entries slot, slot + stride, ...; each a
// fixed-length sum (terms outside the entry's range masked) in four partial sums.
template <int JB>
__device__ __forceinline__ void diag_trows(int slot, int stride, double (*D)[kPW + 1]) {
  constexpr int J0 = 6 * JB;
  for (int e = slot; e < 6 * J0; e += stride) {
    const int a = e % 6, c = e / 6, cb = c / 6;
    double acc[4] = {-D[J0 + a][c], 0.0, 0.0, 0.0};
#pragma unroll
    for (int kb = 1; kb < JB; ++kb)
#pragma unroll
      for (int t = 0; t < 6; ++t) {
        const double tv = D[c][6 * kb + t];
        acc[(6 * kb + t) & 3] -= D[J0 + a][6 * kb + t] * (kb > cb ? tv : 0.0);
      }
    D[c][J0 + a] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
  }
}

// The per-node values a diagonal factorization needs, read at kernel entry (ahead of the
// barriers, which would otherwise keep these loads from being hoisted).
struct DiagNode {
  int c0, K;                        // first column block, node columns
  int64_t loff, toff, poff, uoff;   // L panel, row-major L_KK copy, L_PP^-T slots, RHS
};
__device__ __forceinline__ DiagNode diag_node(const DevFactor& f, int X) {
  return DiagNode{f.sn_c0[X],   6 * f.sn_nc[X], f.sn_loff[X],
                  f.sn_toff[X], f.sn_poff[X],   f.sn_uoff[X]};
}

// Factor the np x np diagonal block of the panel at column k0 of one front in LDS (block
// LDLt with 6x6 pivots and the PSD rule), then store L_PP (strictly below the pivot
// blocks) into the factor and its row-major copy, the pivots (L_jj, d_j, M_j) into
// piv_L / piv_d / piv_M and L_PP^-T into the panel's slot of tinv.  Called by all threads
// of one workgroup.  D holds the block's lower triangle already when `loaded` (else it is
// read from the front).  sL / sdi hold each pivot's factor (L_jj, then d_j in
// sL[j][36..41]; 1 / d_j in sdi[j]).
//
// Per pivot j (one barrier each):
//   wavefront 0   the next pivot block: lane (a, b), a >= b, of 21 forms row j1 + a of L
//                 from its W (the row's values before the solve) and subtracts L_a W_b^T;
//                 thread 0 gathers the 21 entries and factors the block, so its factor is
//                 ready at the next barrier;
//   wavefronts 1-3  the trailing update of the rows below the next pivot block (four
//                 consecutive threads per row, each row's L formed from its W; the updated
//                 values of the next pivot's columns are also kept in Wd, the next
//                 pivot's W) and row block j of T = L_PP^-1 (its L columns are final),
//                 spread from the last thread down so that it falls on the threads whose
//                 rows have no trailing work.
// The pivots' inverses M_j and the stores follow the loop.
__device__ void factor_panel_diag(const DevFactor& f, const DiagNode& dn, double* A, int F, int k0,
                                  int np, DevScalars* sc, double (*D)[kPW + 1],
                                  double (*Wd)[kPW + 1], double (*sL)[42], double (*sdi)[6],
                                  bool loaded, double* __restrict__ z) {
  const int tid = threadIdx.x;
  if (!loaded) {
    stage_lds<(kPW * kPW + kThreads - 1) / kThreads>(
        np * np,
        [&](int e, bool ok) {
          const int r = e % np, c = e / np;
          return load_or_zero(A, int64_t(k0 + c) * F + k0 + r, ok && r >= c);
        },
        [&](int e, double v) { D[e % np][e / np] = v; });
    __syncthreads();
  }
  // W of the first pivot: its columns below its block
  for (int e = 36 + tid; e < 6 * np; e += blockDim.x) Wd[e / 6][e % 6] = D[e / 6][e % 6];
  // L_PP^-T (row-major kPW x kPW) for the panel rows: its unit diagonal blocks and zeros
  // here, the blocks above the diagonal as T's row blocks are formed
  double* Tg = f.tinv + dn.poff + int64_t(k0 / kPW) * kPW * kPW;
  for (int e = tid - 64; e >= 0 && e < kPW * kPW; e += blockDim.x - 64) {  // beside pivot 0
    const int r = e / kPW, c = e % kPW;
    if (r / 6 >= c / 6 || c >= np) Tg[e] = r == c && r < np ? 1.0 : 0.0;
  }
  double* Lout = f.lbuf + dn.loff;
  double* Lt = f.ltbuf + dn.toff;
  const int K = dn.K;
  // thread 0 factors the 6x6 block a6 into slot s
  auto factor_pivot = [&](const double* a6, int s) {
    double L[36], d[6], di[6];
    const bool ok = f.scalar_pivots ? ldlt6<true>(a6, L, d, di) : ldlt6<false>(a6, L, d, di);
    if (!ok) {
      atomicOr(&sc->bad_pivot, 1);
      for (int k = 0; k < 36; ++k) L[k] = (k % 7 == 0) ? 1.0 : 0.0;
      for (int k = 0; k < 6; ++k) d[k] = di[k] = 0.0;
    }
    for (int k = 0; k < 36; ++k) sL[s][k] = L[k];
    for (int k = 0; k < 6; ++k) {
      sL[s][36 + k] = d[k];
      sdi[s][k] = di[k];
    }
  };
  if (tid == 0) {
    double a6[36];
    for (int a = 0; a < 6; ++a)
      for (int b = 0; b < 6; ++b) a6[6 * a + b] = a >= b ? D[a][b] : D[b][a];
    factor_pivot(a6, 0);
  }
  for (int j0 = 0; j0 < np; j0 += 6) {
    __syncthreads();  // pivot j0 is in slot j0 / 6; the previous pivot's work is done
    const int jb = j0 / 6;
    const double* cL = sL[jb];
    const double* cdi = sdi[jb];
    const int j1 = j0 + 6, m = np - j1;  // the next pivot; rows of the block below this one
    if (tid < 64) {
      if (m > 0) {
        const int e = tid < 21 ? tid : 0;
        int a = 0;
        while ((a + 1) * (a + 2) / 2 <= e) ++a;
        const int b = e - a * (a + 1) / 2;
        double x[6];
#pragma unroll
        for (int t = 0; t < 6; ++t) x[t] = Wd[j1 + a][j0 + t];
        right_solve6(cL, cdi, x);
        if (tid < 21 && b == 0)
#pragma unroll
          for (int t = 0; t < 6; ++t) D[j1 + a][j0 + t] = x[t];
        double acc = D[j1 + a][j1 + b];
#pragma unroll
        for (int t = 0; t < 6; ++t) acc -= x[t] * Wd[j1 + b][j0 + t];
        // gather the 21 entries into thread 0 (uniform reads of lanes 0..20)
        const int2 bits = __builtin_bit_cast(int2, acc);
        double a6[36];
#pragma unroll
        for (int q = 0; q < 21; ++q) {
          int qa = 0;
          while ((qa + 1) * (qa + 2) / 2 <= q) ++qa;
          const int qb = q - qa * (qa + 1) / 2;
          const int2 vb = make_int2(__builtin_amdgcn_readlane(bits.x, q),
                                    __builtin_amdgcn_readlane(bits.y, q));
          const double v = __builtin_bit_cast(double, vb);
          a6[6 * qa + qb] = v;
          a6[6 * qb + qa] = v;
        }
        if (tid == 0) factor_pivot(a6, jb + 1);
      }
    } else {
      static_assert(kThreads - 64 == 4 * kPW && kPanelBlocks == 8, "four threads per block row");
      switch (jb) {
#define SGPU_TRAIL(J) \
  case J: diag_trailing<J>(tid - 64, np, cL, cdi, D, Wd); break;
        SGPU_TRAIL(0) SGPU_TRAIL(1) SGPU_TRAIL(2) SGPU_TRAIL(3)
        SGPU_TRAIL(4) SGPU_TRAIL(5) SGPU_TRAIL(6) SGPU_TRAIL(7)
#undef SGPU_TRAIL
      }
    }
    // Beside the look-ahead: row block jb of T = L_PP^-1 (its L columns are final), kept
    // transposed in the free upper triangle, D[c][r] = T[r][c] for block(c) < block(r),
    // four partial sums per entry; then the outputs of pivot jb - 1 (M_j, L_jj, d_j) and
    // the stores of L's column block jb - 1 and T's row block jb - 1.  The entries of T go
    // first to the wavefronts with the least trailing work (wavefront 1, then 3 and 2 while
    // wavefront 2's rows are updated, 2 and 3 after); the last pivot has no look-ahead,
    // so there wavefront 0 joins and also handles that pivot's outputs.
    const bool look = m > 0;
    const int wv = tid >> 6, lane = tid & 63;
    if (!look || wv > 0) {
      const int pos = !look ? wv : (jb >= 4 ? wv - 1 : (wv == 1 ? 0 : (wv == 3 ? 1 : 2)));
      const int slot = 64 * pos + lane, stride = look ? kThreads - 64 : kThreads;
      switch (jb) {
#define SGPU_TROWS(J) \
  case J: diag_trows<J>(slot, stride, D); break;
        SGPU_TROWS(1) SGPU_TROWS(2) SGPU_TROWS(3) SGPU_TROWS(4)
        SGPU_TROWS(5) SGPU_TROWS(6) SGPU_TROWS(7)
#undef SGPU_TROWS
      }
      // M_j = L^-T D^+ L^-1 (lane (a, b) of 36, from the columns of L^-1), L_jj and d_j
      auto pivot_outputs = [&](int j) {
        const int gb = dn.c0 + k0 / 6 + j;
        if (lane < 36) {
          const int a = lane / 6, b = lane % 6;
          double ua[6], ub[6];
#pragma unroll
          for (int k = 0; k < 6; ++k) {
            ua[k] = k == a ? 1.0 : 0.0;
            ub[k] = k == b ? 1.0 : 0.0;
          }
#pragma unroll
          for (int k = 1; k < 6; ++k)
#pragma unroll
            for (int t = 0; t < k; ++t) {
              ua[k] -= sL[j][6 * k + t] * ua[t];
              ub[k] -= sL[j][6 * k + t] * ub[t];
            }
          double mm = 0.0;
#pragma unroll
          for (int k = 0; k < 6; ++k) mm += ua[k] * sdi[j][k] * ub[k];
          f.piv_M[36 * int64_t(gb) + lane] = mm;
          f.piv_L[36 * int64_t(gb) + lane] = sL[j][lane];
        } else if (lane < 42) {
          f.piv_d[6 * int64_t(gb) + lane - 36] = sL[j][lane];
        }
      };
      if (jb > 0 && wv == (jb >= 4 ? 2 : 3)) {
        pivot_outputs(jb - 1);
        const int p0 = j0 - 6, nr = np - j0;
        for (int e = lane; e < 6 * nr; e += 64) {
          const int rr2 = e % nr, t = e / nr;
          Lout[int64_t(k0 + p0 + t) * F + k0 + j0 + rr2] = D[j0 + rr2][p0 + t];
        }
        for (int e = lane; e < 6 * nr; e += 64) {  // row-major copy
          const int rr2 = e / 6, t = e % 6;
          Lt[int64_t(k0 + j0 + rr2) * K + k0 + p0 + t] = D[j0 + rr2][p0 + t];
        }
        for (int e = lane; e < 6 * p0; e += 64) {
          const int cp = e / 6, a = e % 6;
          Tg[cp * kPW + p0 + a] = D[cp][p0 + a];
        }
      }
      if (!look && wv == 0) pivot_outputs(jb);
    }
  }
  __syncthreads();  // T's last row block is in D
  const int pl = np - 6;
  for (int e = tid; e < 6 * pl; e += blockDim.x) {
    const int cp = e / 6, a = e % 6;
    Tg[cp * kPW + pl + a] = D[cp][pl + a];
  }
  if (k0 + np == F) {
    // no rows below the panel (the last panel of a root), so no panel step follows: its
    // right-hand side is final here; z_P = D^-1 L_PP^-1 b_P for the panel's columns
    const double* bx = f.usolve + dn.uoff + k0;
    double* yv = Wd[0];  // the block's W copy is done with
    if (tid < np) {
      double y = bx[tid];
      for (int r = 0; r < 6 * (tid / 6); ++r) y += D[r][tid] * bx[r];  // T[c][r] = D[r][c]
      yv[tid] = y;
    }
    __syncthreads();
    if (tid < np / 6) {
      const int j = tid;
      double v[6];
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        double t = yv[6 * j + i];
        for (int k = 0; k < i; ++k) t -= sL[j][6 * i + k] * v[k];
        v[i] = t;
      }
#pragma unroll
      for (int i = 0; i < 6; ++i) v[i] *= sdi[j][i];
#pragma unroll
      for (int i = 5; i >= 0; --i) {
        double t = v[i];
        for (int k = i + 1; k < 6; ++k) t -= sL[j][6 * k + i] * v[k];
        v[i] = t;
      }
#pragma unroll
      for (int i = 0; i < 6; ++i) z[6 * int64_t(dn.c0) + k0 + 6 * j + i] = v[i];
    }
  }
}

// First panel diagonal of the leaves (whose fronts are assembled before the factorization
// starts; every other node's is factored by the last step of its last child).
__global__ void __launch_bounds__(kThreads)
    k_mf_diag0(DevFactor f, const int32_t* __restrict__ nodes, DevScalars* sc, double* __restrict__ z) {
  const int X = nodes[blockIdx.x];
  const int F = 6 * f.sn_nf[X], K = 6 * f.sn_nc[X];
  __shared__ double D[kPW][kPW + 1], Wd[kPW][kPW + 1], sL[kPanelBlocks][42], sdi[kPanelBlocks][6];
  factor_panel_diag(f, diag_node(f, X), f.front + f.sn_off[X], F, 0, min(K, kPW), sc, D, Wd, sL,
                    sdi, false, z);
}

typedef double mfma_d4 __attribute__((ext_vector_type(4)));

// Position in the parent's front of row u of node X's update matrix (front row K + u):
// update row block b is the parent's row block upos[b].
__device__ __forceinline__ int parent_row(const DevFactor& f, int X, int u) {
  return 6 * f.sn_upos[f.sn_upos_ptr[X] + u / 6] + u % 6;
}

// Extend-add of one lane's results of a trailing-update tile on node X's last panel step
// (the tile lies in its update matrix, rows and columns >= K): accumulator (mi, nj)
// component q is the front element (rl + 16 mi, cl + 16 nj + 4 q).  Each is added into
// the parent P's front at the mapped row and column (update rows keep their order in the
// parent, so the lower triangle maps to the lower triangle); every load first.
__device__ __forceinline__ void extend_tile(const DevFactor& f, int X, int P, int K, int F,
                                            int rl, int cl, const mfma_d4 (&acc)[2][2]) {
  double* B = f.front + f.sn_off[P];
  const int64_t FP = 6 * int64_t(f.sn_nf[P]);
  int pr[2];
  int64_t pc[2][4];
#pragma unroll
  for (int mi = 0; mi < 2; ++mi) {
    const int r = rl + 16 * mi;
    pr[mi] = r < F ? parent_row(f, X, r - K) : 0;
  }
#pragma unroll
  for (int nj = 0; nj < 2; ++nj)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = cl + 16 * nj + 4 * q;
      pc[nj][q] = c < F ? parent_row(f, X, c - K) * FP : 0;
    }
  double v[2][2][4];
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int nj = 0; nj < 2; ++nj)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = rl + 16 * mi, c = cl + 16 * nj + 4 * q;
        v[mi][nj][q] = load_or_zero(B, pc[nj][q] + pr[mi], r < F && c < F && r >= c);
      }
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int nj = 0; nj < 2; ++nj)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = rl + 16 * mi, c = cl + 16 * nj + 4 * q;
        if (r < F && c < F && r >= c) B[pc[nj][q] + pr[mi]] = v[mi][nj][q] + acc[mi][nj][q];
      }
}

// extend_tile for the tile (0, 0) of a parent's last child on its last step, which then
// factors the parent's first diagonal block (np x np): besides the adds into the parent's
// front, the block's values after them are built in LDS (D2, which the caller's update no
// longer reads) from the block's values before them, loaded with the adds' own loads, so
// the factorization does not wait for the adds to reach memory and load them again.
// Called by all threads; ends with the block complete in D2.
__device__ __forceinline__ void extend_tile_hand(const DevFactor& f, int X, int P, int K, int F,
                                                 int rl, int cl, const mfma_d4 (&acc)[2][2],
                                                 double (*D2)[kPW + 1], int np) {
  double* B = f.front + f.sn_off[P];
  const int64_t FP = 6 * int64_t(f.sn_nf[P]);
  int pr[2], pc[2][4];
#pragma unroll
  for (int mi = 0; mi < 2; ++mi) {
    const int r = rl + 16 * mi;
    pr[mi] = r < F ? parent_row(f, X, r - K) : 0;
  }
#pragma unroll
  for (int nj = 0; nj < 2; ++nj)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = cl + 16 * nj + 4 * q;
      pc[nj][q] = c < F ? parent_row(f, X, c - K) : 0;
    }
  double v[2][2][4];
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int nj = 0; nj < 2; ++nj)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = rl + 16 * mi, c = cl + 16 * nj + 4 * q;
        v[mi][nj][q] = load_or_zero(B, pc[nj][q] * FP + pr[mi], r < F && c < F && r >= c);
      }
  constexpr int kNB = (kPW * kPW + kThreads - 1) / kThreads;
  double bv[kNB];
#pragma unroll
  for (int it = 0; it < kNB; ++it) {
    const int e = threadIdx.x + kThreads * it, r = e % np, c = e / np;
    bv[it] = load_or_zero(B, int64_t(c) * FP + r, e < np * np && r >= c);
  }
  __syncthreads();  // every wavefront is done with D2's memory
#pragma unroll
  for (int it = 0; it < kNB; ++it) {
    const int e = threadIdx.x + kThreads * it;
    if (e < np * np) D2[e % np][e / np] = bv[it];
  }
  __syncthreads();  // the block's values before the adds are in place
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int nj = 0; nj < 2; ++nj)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = rl + 16 * mi, c = cl + 16 * nj + 4 * q;
        if (r < F && c < F && r >= c) {
          const double nv = v[mi][nj][q] + acc[mi][nj][q];
          B[pc[nj][q] * FP + pr[mi]] = nv;
          if (pr[mi] < np && pc[nj][q] < np) D2[pr[mi]][pc[nj][q]] = nv;
        }
      }
  __syncthreads();
}

// Outputs of one wavefront's 16 rows r0.. of a panel step's rows below the diagonal block,
// from their L in registers (component q of L4[ct] is row r0 + lc + 4 q, column
// 16 ct + lr of the panel, lane = 16 lc + lr): the L rows to the factor (and the row-major
// copy for the node's own rows), the rows' right-hand side b_r -= L_r y_P with
// y_P = L_PP^-1 b_P (sb: b_P, bt: the rows' b, sy: the wavefront's copy of y_P), into the
// parent's right-hand side on the node's last step (P >= 0); with zp, z_P = D^-1 y_P for
// the panel's columns.
__device__ __forceinline__ void panel_rows_out(const DevFactor& f, const DiagNode& dn, int P, int X,
                                               int F, int K, int k0, int np, int r0,
                                               const mfma_d4 (&L4)[3], const double (*sT)[kPW + 1],
                                               const double* sb, const double (*pM)[36],
                                               const double* bt, double* sy, double* __restrict__ z,
                                               int lane, bool zp) {
  const int lr = lane & 15, lc = lane >> 4;
  if (lane < kPW) {
    double a4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int t = 0; t < kPW; ++t) a4[t & 3] += sT[t][lane] * sb[t];  // L_PP^-T (t, c) = L_PP^-1 (c, t)
    sy[lane] = (a4[0] + a4[1]) + (a4[2] + a4[3]);
  }
  double* Lo = f.lbuf + dn.loff;
  double* Lt = f.ltbuf + dn.toff;
#pragma unroll
  for (int ct = 0; ct < 3; ++ct) {
    const int c = 16 * ct + lr;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = r0 + lc + 4 * q;
      if (c < np && r < F) Lo[int64_t(k0 + c) * F + r] = L4[ct][q];
      if (c < np && r < K) Lt[int64_t(r) * K + k0 + c] = L4[ct][q];
    }
  }
  double part[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double acc = 0.0;
#pragma unroll
    for (int ct = 0; ct < 3; ++ct) acc += L4[ct][q] * sy[16 * ct + lr];  // L is 0 past np
    acc += __shfl_xor(acc, 1, 64);
    acc += __shfl_xor(acc, 2, 64);
    acc += __shfl_xor(acc, 4, 64);
    acc += __shfl_xor(acc, 8, 64);
    part[q] = acc;
  }
  if (lr == 0)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = lc + 4 * q, r = r0 + i;
      if (r < F) {
        if (P < 0)
          f.usolve[dn.uoff + r] = bt[i] - part[q];
        else  // a finished update row: into the parent's right-hand side
          f.usolve[f.sn_uoff[P] + parent_row(f, X, r - K)] += bt[i] - part[q];
      }
    }
  if (zp && lane < np) {  // z_P = D^-1 y_P (blockwise M_j)
    const int jb = lane / 6, a = lane % 6;
    double zz = 0.0;
#pragma unroll
    for (int t = 0; t < 6; ++t) zz += pM[jb][6 * a + t] * sy[6 * jb + t];
    z[6 * int64_t(dn.c0) + k0 + lane] = zz;
  }
}

// One round of panel steps, one workgroup per item (node, panel, tr, tc): the 64x64 tile
// (tr, tc) of the node's lower trailing region below the panel:
//   1. the tile's values, the panel columns of its row block tr and column block tc (C),
//      L_PP^-T and the pivot inverses M_j are loaded (every global load first);
//   2. W = C L_PP^-T for both blocks and L = W blockdiag(M_j) for block tr on the matrix
//      cores (v_mfma_f64_16x16x4; each wavefront its own 16 rows of a 64-row block, so
//      W and L need no barrier between them);
//   3. A -= L W^T on the matrix cores: four wavefronts x 2 x 2 tiles of 16x16, (-W) L^T so
//      that each lane's results run down 16 rows of the column-major front, accumulators
//      seeded with the front's values;
//   4. each row block's output tile stores the block's L rows (and the row-major copy of
//      the node's own rows) and eliminates its rows of the node's right-hand side,
//      b_r -= L_r y_P with y_P = L_PP^-1 b_P; the output tile is the diagonal tile (tr, tr),
//      but for row block 0 it is tile (1, 0) when the step has one (it forms the block's L
//      from its column block's W), and it also writes z_P = D^-1 y_P for the panel's
//      columns.  Tile (0, 0) factors the next panel's diagonal block (the top-left of its
//      tile, handed over in LDS);
//   5. on the node's last panel step the trailing region is its update matrix: the tile
//      adds its results into the parent's front instead (extend-add; child update row
//      block b is the parent's row block upos[b]), and diagonal tiles add their finished
//      update rows of the right-hand side into the parent's.  No two children of one
//      parent end in one round and a child's targets are distinct, so the adds are plain;
//      the last child's tile (0, 0) then factors the parent's first diagonal block.
// Each tile forms the L and W rows it needs itself (a row block is formed again by every
// tile of its tile row and column), so a panel step is one launch with no panel-row pass
// before it.
__global__ void __launch_bounds__(kThreads)
    k_mf_step(DevFactor f, const int32_t* __restrict__ tiles, DevScalars* sc,
              double* __restrict__ z) {
  static_assert(kTile == 64 && kThreads == 256 && kPW == 48, "4 wavefronts, 16-row slices");
  // consecutive tiles (one tile row shares its L rows) run on one XCD
  const int bid = xcd_contiguous(blockIdx.x, gridDim.x);
  const int X = tiles[4 * bid], panel = tiles[4 * bid + 1];
  const int tr = tiles[4 * bid + 2], tc = tiles[4 * bid + 3];
  const int F = 6 * f.sn_nf[X], K = 6 * f.sn_nc[X];
  double* A = f.front + f.sn_off[X];
  const DiagNode dn = diag_node(f, X);
  const double* Tg = f.tinv + dn.poff + int64_t(panel) * kPW * kPW;
  double* bx = f.usolve + dn.uoff;  // the node's right-hand side (front row order)
  const int k0 = kPW * panel;
  const int kend = min(K, k0 + kPW);
  const int np = kend - k0;
  const int r0 = kend + kTile * tr, c0 = kend + kTile * tc;
  const bool dg = tr == tc;
  // The L rows and right-hand side of row block tr are output by the diagonal tile (tr, tr),
  // except row block 0's: tile (1, 0), when the step has it, forms that block's L from its
  // column block's W and outputs it, so that tile (0, 0) goes from its update straight to
  // the next diagonal block.
  const int nt = (F - kend + kTile - 1) / kTile;  // tile rows of the step
  const bool own0 = tr == 1 && tc == 0;
  const bool out = own0 || (dg && !(tr == 0 && nt > 1));
  const int ro = own0 ? c0 : r0;  // the output row block's first row
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int lr = lane & 15, lc = lane >> 4;
  const int rb = 32 * (wv >> 1), cb = 32 * (wv & 1);  // the wavefront's quadrant
  const int rw = 16 * wv;                              // the wavefront's 16-row slice
  __shared__ double sA[kPW][kTile + 1], sB[kPW][kTile + 1];  // C^T, then L^T / W^T
  __shared__ double sT[kPW][kPW + 1];                          // L_PP^-T
  __shared__ double pM[kPanelBlocks][36];
  __shared__ double sL[kPanelBlocks][42], sdi[kPanelBlocks][6];
  __shared__ double sb[kPW], sy[kThreads / 64][kPW];
  __shared__ double sbt[kTile];  // the output row block's rows of the right-hand side
  // the node's last panel step: its update matrix goes to the parent (-1: none)
  const int P = kend == K ? f.sn_parent[X] : -1;
  // the last child of its parent (in round order) factors the parent's first diagonal
  // block after its last step's tile (0, 0): update rows keep their order in the parent, so
  // that block gets the children's update only from their tiles (0, 0), and the other
  // children ended in earlier rounds
  const bool hand = P >= 0 && tr == 0 && tc == 0 && f.sn_hand[X];
  bool diag = false;  // this workgroup factors the node's next diagonal block
  {
    const int gb0 = dn.c0 + k0 / 6, nbp = np / 6;
    constexpr int kNC = kPW * kTile / kThreads, kNT = kPW * kPW / kThreads;
    constexpr int kNM = (kPanelBlocks * 36 + kThreads - 1) / kThreads;
    static_assert(kPW * kTile % kThreads == 0 && kPW * kPW % kThreads == 0, "whole rounds");
    double va[kNC], vb[kNC], vt[kNT], vm[kNM];
#pragma unroll
    for (int it = 0; it < kNC; ++it) {
      const int e = tid + kThreads * it, t = e / kTile, i = e % kTile;
      va[it] = load_or_zero(A, int64_t(k0 + t) * F + r0 + i, t < np && r0 + i < F);
      vb[it] = load_or_zero(A, int64_t(k0 + t) * F + c0 + i, !dg && t < np && c0 + i < F);
    }
#pragma unroll
    for (int it = 0; it < kNT; ++it) vt[it] = Tg[tid + kThreads * it];
#pragma unroll
    for (int it = 0; it < kNM; ++it) {
      const int e = tid + kThreads * it;
      vm[it] = load_or_zero(f.piv_M, int64_t(36) * gb0 + e, e < kPanelBlocks * 36 && e / 36 < nbp);
    }
    const double vbb = load_or_zero(bx, k0 + tid, out && tid < np);
    const double vbt = load_or_zero(bx, ro + tid - kPW, out && tid >= kPW && tid < kPW + kTile &&
                                                            ro + tid - kPW < F);
#pragma unroll
    for (int it = 0; it < kNC; ++it) {
      const int e = tid + kThreads * it;
      sA[e / kTile][e % kTile] = va[it];
      sB[e / kTile][e % kTile] = vb[it];
    }
#pragma unroll
    for (int it = 0; it < kNT; ++it) {
      const int e = tid + kThreads * it;
      sT[e / kPW][e % kPW] = vt[it];
    }
#pragma unroll
    for (int it = 0; it < kNM; ++it) {
      const int e = tid + kThreads * it;
      if (e < kPanelBlocks * 36) pM[e / 36][e % 36] = vm[it];
    }
    if (tid < kPW) sb[tid] = vbb;
    if (tid >= kPW && tid < kPW + kTile) sbt[tid - kPW] = vbt;
  }
  __syncthreads();
  // accumulator (mi, nj) component q is the tile element
  //   row rb + 16 mi + lr, column cb + 16 nj + lc + 4 q;
  // it starts from the element's current value, loaded while W and L are formed
  mfma_d4 acc[2][2];
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int nj = 0; nj < 2; ++nj)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = r0 + rb + 16 * mi + lr, c = c0 + cb + 16 * nj + lc + 4 * q;
        acc[mi][nj][q] = load_or_zero(A, int64_t(c) * F + r, r < F && c < F && r >= c);
      }
  // W = C L_PP^-T for the wavefront's slice of both row blocks; L_PP^-T is block upper
  // triangular, so column tile ct stops at row kW[ct] (C and L_PP^-T are zero past np)
  constexpr int kW[3] = {18, 36, 48};
  mfma_d4 wa[3], wb[3];
#pragma unroll
  for (int ct = 0; ct < 3; ++ct) wa[ct] = wb[ct] = mfma_d4{0.0, 0.0, 0.0, 0.0};
  if (dg) {  // the column block is the row block
#pragma unroll
    for (int k4 = 0; k4 < kPW; k4 += 4) {
      const double av = sA[k4 + lc][rw + lr];
#pragma unroll
      for (int ct = 0; ct < 3; ++ct)
        if (k4 < kW[ct])
          wa[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, sT[k4 + lc][16 * ct + lr], wa[ct], 0, 0, 0);
    }
  } else {
#pragma unroll
    for (int k4 = 0; k4 < kPW; k4 += 4) {
      const double av = sA[k4 + lc][rw + lr], bv = sB[k4 + lc][rw + lr];
#pragma unroll
      for (int ct = 0; ct < 3; ++ct)
        if (k4 < kW[ct]) {
          const double tv = sT[k4 + lc][16 * ct + lr];
          wa[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, tv, wa[ct], 0, 0, 0);
          wb[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv, tv, wb[ct], 0, 0, 0);
        }
    }
  }
  __syncthreads();  // every wavefront is done with C
  // W^T of both blocks (a diagonal tile's column block is its row block)
#pragma unroll
  for (int ct = 0; ct < 3; ++ct)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      sA[16 * ct + lr][rw + lc + 4 * q] = wa[ct][q];
      sB[16 * ct + lr][rw + lc + 4 * q] = dg ? wa[ct][q] : wb[ct][q];
    }
  // L = W blockdiag(M_j) for the slice of row block tr: column tile ct meets the pivot
  // blocks with rows [kL0[ct], kL1[ct]); the block-diagonal operand is read from the M_j
  constexpr int kL0[3] = {0, 12, 30}, kL1[3] = {18, 36, 48};
#pragma unroll
  for (int ct = 0; ct < 3; ++ct) wa[ct] = mfma_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k4 = 0; k4 < kPW; k4 += 4) {
    const int k = k4 + lc;
    const double av = sA[k][rw + lr];
#pragma unroll
    for (int ct = 0; ct < 3; ++ct)
      if (k4 + 4 > kL0[ct] && k4 < kL1[ct]) {
        const int n = 16 * ct + lr;
        const double mv = k / 6 == n / 6 ? pM[k / 6][6 * (k % 6) + n % 6] : 0.0;
        wa[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, mv, wa[ct], 0, 0, 0);
      }
  }
  if (own0) {
    // row block 0's outputs, from L of the block (the column block's W times the M_j) in
    // registers: component q of wb[ct] is L row rw + lc + 4 q, column 16 ct + lr
#pragma unroll
    for (int ct = 0; ct < 3; ++ct) wb[ct] = mfma_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k4 = 0; k4 < kPW; k4 += 4) {
      const int k = k4 + lc;
      const double av = sB[k][rw + lr];
#pragma unroll
      for (int ct = 0; ct < 3; ++ct)
        if (k4 + 4 > kL0[ct] && k4 < kL1[ct]) {
          const int n = 16 * ct + lr;
          const double mv = k / 6 == n / 6 ? pM[k / 6][6 * (k % 6) + n % 6] : 0.0;
          wb[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, mv, wb[ct], 0, 0, 0);
        }
    }
    panel_rows_out(f, dn, P, X, F, K, k0, np, ro + rw, wb, sT, sb, pM, sbt + rw, sy[wv], z,
                   lane, wv == 0);
  }
  // L^T replaces the slice's W^T in sA
#pragma unroll
  for (int ct = 0; ct < 3; ++ct)
#pragma unroll
    for (int q = 0; q < 4; ++q) sA[16 * ct + lr][rw + lc + 4 * q] = wa[ct][q];
  __syncthreads();  // the update reads every row of L^T and W^T
#pragma unroll
  for (int s = 0; s < kPW / 4; ++s) {
    const int t = 4 * s + lc;  // this lane's k index
    double a[2], b[2];
#pragma unroll
    for (int nj = 0; nj < 2; ++nj) a[nj] = -sB[t][cb + 16 * nj + lr];  // A operand: -W (16 x 4)
#pragma unroll
    for (int mi = 0; mi < 2; ++mi) b[mi] = sA[t][rb + 16 * mi + lr];   // B operand: L^T (4 x 16)
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
      for (int nj = 0; nj < 2; ++nj)
        acc[mi][nj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[nj], b[mi], acc[mi][nj], 0, 0, 0);
  }
  if (P < 0) {
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
      for (int nj = 0; nj < 2; ++nj)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = r0 + rb + 16 * mi + lr, c = c0 + cb + 16 * nj + lc + 4 * q;
          if (r < F && c < F && r >= c) A[int64_t(c) * F + r] = acc[mi][nj][q];
        }
  } else if (!hand) {
    extend_tile(f, X, P, K, F, r0 + rb + lr, c0 + cb + lc, acc);
  } else {
    extend_tile_hand(f, X, P, K, F, r0 + rb + lr, c0 + cb + lc, acc, as_square(sB),
                     min(6 * f.sn_nc[P], kPW));
  }
  if (out && !own0) {
    // the slice's L rows to the factor (and the row-major copy of the node's own rows)
    double (*Lr)[kTile + 1] = sA;  // L^T of the row block
    double* Lo = f.lbuf + dn.loff;
    double* Lt = f.ltbuf + dn.toff;
    for (int e = lane; e < 16 * kPW; e += 64) {
      const int i = rw + e % 16, c = e / 16, r = ro + i;
      if (c < np && r < F) Lo[int64_t(k0 + c) * F + r] = Lr[c][i];
    }
    if (ro + rw < K)
      for (int e = lane; e < 16 * kPW; e += 64) {
        const int c = e % kPW, i = rw + e / kPW, r = ro + i;
        if (c < np && r < K) Lt[int64_t(r) * K + k0 + c] = Lr[c][i];
      }
    // forward elimination of the right-hand side: y_P = L_PP^-1 b_P (each wavefront its
    // own copy, lane c < kPW entry c), then b_r -= L_r y_P for the slice's rows (four
    // lanes per row, every fourth column, two xor shuffles)
    if (lane < kPW) {
      double a4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int t = 0; t < kPW; ++t) a4[t & 3] += sT[t][lane] * sb[t];  // L_PP^-T (t, c) = L_PP^-1 (c, t)
      sy[wv][lane] = (a4[0] + a4[1]) + (a4[2] + a4[3]);
    }
    {
      const int i = rw + (lane >> 2), part = lane & 3, r = ro + i;
      double s4 = 0.0;
#pragma unroll
      for (int u = 0; u < kPW / 4; ++u) s4 += Lr[part + 4 * u][i] * sy[wv][part + 4 * u];
      s4 += __shfl_xor(s4, 1, 64);
      s4 += __shfl_xor(s4, 2, 64);
      if (part == 0 && r < F) {
        if (P < 0)
          bx[r] = sbt[i] - s4;
        else  // a finished update row: into the parent's right-hand side
          f.usolve[f.sn_uoff[P] + parent_row(f, X, r - K)] += sbt[i] - s4;
      }
    }
    if (ro == kend && wv == 0 && lane < np) {  // row block 0: z_P = D^-1 y_P (blockwise M_j)
      const int jb = lane / 6, a = lane % 6;
      double zz = 0.0;
#pragma unroll
      for (int t = 0; t < 6; ++t) zz += pM[jb][6 * a + t] * sy[0][6 * jb + t];
      z[6 * int64_t(dn.c0) + k0 + lane] = zz;
    }
  }
  if (tr == 0 && tc == 0 && kend < K) {
    // the next panel's diagonal block is the top-left of this tile: hand the updated
    // values to LDS directly
    const int npn = min(K, kend + kPW) - kend;
    double (*D)[kPW + 1] = as_square(sA);
    __syncthreads();  // every read of sA / sB is done
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
      for (int nj = 0; nj < 2; ++nj)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int rl = rb + 16 * mi + lr, cl = cb + 16 * nj + lc + 4 * q;
          if (rl < npn && cl < npn) D[rl][cl] = rl >= cl ? acc[mi][nj][q] : 0.0;
        }
    __syncthreads();
    diag = true;
  }
  if (diag) {
    factor_panel_diag(f, dn, A, F, kend, min(K, kend + kPW) - kend, sc, as_square(sA),
                      as_square(sB), sL, sdi, true, z);
  } else if (hand) {
    __syncthreads();  // the work above that reads sA is done
    factor_panel_diag(f, diag_node(f, P), f.front + f.sn_off[P], 6 * f.sn_nf[P], 0,
                      min(6 * f.sn_nc[P], kPW), sc, as_square(sB), as_square(sA), sL, sdi, true, z);
  }
}

// The split form of a panel step, for launches with many tiles (where forming a row
// block's L and W once, instead of in every tile of its tile row and column, saves more
// matrix-core time than the extra launch costs).
//
// Panel rows of one round of panel steps (items = (node, panel, kPanelRows-row chunk)
// below the panel's diagonal block), as two products instead of a substitution:
//   W_r = C_r L_PP^-T            (the rows' values after the panel's own eliminations)
//   L_r,j = W_r,j M_j,  M_j = L_jj^-T D_j^-1 L_jj^-1
// with L_PP^-T from the diagonal factorization.  Both run on the matrix cores as 4 x 3
// tiles of v_mfma_f64_16x16x4, three per wavefront.  W goes to the node's W scratch, L to
// the factor and, for the node's own rows, to its row-major copy.
__global__ void __launch_bounds__(kThreads)
    k_mf_panel(DevFactor f, const int32_t* __restrict__ items, double* __restrict__ z) {
  static_assert(kPanelRows == 64 && kPW == 48 && kThreads == 256, "4 x 3 tiles, 3 per wavefront");
  static_assert(kPW * kPanelRows % kThreads == 0 && kPW * kPW % kThreads == 0, "whole staging rounds");
  const int X = items[3 * blockIdx.x], panel = items[3 * blockIdx.x + 1];
  const int chunk = items[3 * blockIdx.x + 2];
  const int F = 6 * f.sn_nf[X], K = 6 * f.sn_nc[X];
  const double* A = f.front + f.sn_off[X];
  double* Lo = f.lbuf + f.sn_loff[X];
  double* Wo = f.wbuf + f.sn_woff[X];
  double* Lt = f.ltbuf + f.sn_toff[X];
  const double* Tg = f.tinv + f.sn_poff[X] + int64_t(panel) * kPW * kPW;
  const int k0 = kPW * panel, kend = min(K, k0 + kPW), np = kend - k0;
  const int rbase = kend + kPanelRows * chunk;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, lr = lane & 15, lk = lane >> 4;
  __shared__ double sC[kPW][kPanelRows + 1];  // the rows' panel columns, C^T
  __shared__ double sT[kPW][kPW + 1];         // L_PP^-T
  __shared__ double sW[kPanelRows][kPW + 1];
  __shared__ double pM[kPanelBlocks][36];
  __shared__ double sb[kPW], sy[kThreads / 64][kPW];  // the panel's RHS; y_P per wavefront
  double* bx = f.usolve + f.sn_uoff[X];  // the node's right-hand side (front row order)
  const int gb0 = f.sn_c0[X] + k0 / 6, nbp = np / 6;
  // every global load of the thread first, then the LDS stores
  constexpr int kNC = kPW * kPanelRows / kThreads, kNT = kPW * kPW / kThreads;
  constexpr int kNM = (kPanelBlocks * 36 + kThreads - 1) / kThreads;
  double vc[kNC], vt[kNT], vm[kNM];
#pragma unroll
  for (int it = 0; it < kNC; ++it) {
    const int e = tid + kThreads * it, t = e / kPanelRows, r = rbase + e % kPanelRows;
    vc[it] = load_or_zero(A, int64_t(k0 + t) * F + r, t < np && r < F);
  }
#pragma unroll
  for (int it = 0; it < kNT; ++it) vt[it] = Tg[tid + kThreads * it];
#pragma unroll
  for (int it = 0; it < kNM; ++it) {
    const int e = tid + kThreads * it;
    vm[it] = load_or_zero(f.piv_M, int64_t(36) * gb0 + e, e < kPanelBlocks * 36 && e / 36 < nbp);
  }
  const double vb = load_or_zero(bx, k0 + tid, tid < np);
#pragma unroll
  for (int it = 0; it < kNC; ++it) {
    const int e = tid + kThreads * it;
    sC[e / kPanelRows][e % kPanelRows] = vc[it];
  }
#pragma unroll
  for (int it = 0; it < kNT; ++it) {
    const int e = tid + kThreads * it;
    sT[e / kPW][e % kPW] = vt[it];
  }
#pragma unroll
  for (int it = 0; it < kNM; ++it) {
    const int e = tid + kThreads * it;
    if (e < kPanelBlocks * 36) pM[e / 36][e % 36] = vm[it];
  }
  if (tid < kPW) sb[tid] = vb;
  __syncthreads();
  // forward elimination of the right-hand side: y_P = L_PP^-1 b_P (each wavefront forms
  // its own copy; lane c < kPW entry c), and z_P = D^-1 y_P to the solve vector
  if (lane < kPW) {
    double a4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int t = 0; t < kPW; ++t) a4[t & 3] += sT[t][lane] * sb[t];  // L_PP^-T (t, c) = L_PP^-1 (c, t)
    sy[wv][lane] = (a4[0] + a4[1]) + (a4[2] + a4[3]);
  }
  if (chunk == 0 && wv == 0 && lane < np) {
    const int jb = lane / 6, a = lane % 6;
    double zz = 0.0;
#pragma unroll
    for (int t = 0; t < 6; ++t) zz += pM[jb][6 * a + t] * sy[0][6 * jb + t];
    z[6 * int64_t(f.sn_c0[X]) + k0 + lane] = zz;
  }
  // Wavefront wv owns row tile wv (16 rows) and its three 16-column tiles; it reads only
  // its own rows of C, W and L after the barrier above, so it needs no further barrier.
  // W = C L_PP^-T: L_PP^-T is block upper triangular, so column tile ct stops at row
  // kW[ct] (C and L_PP^-T are zero past np).
  constexpr int kW[3] = {18, 36, 48};
  const int rw = 16 * wv;
  mfma_d4 acc[3];
#pragma unroll
  for (int ct = 0; ct < 3; ++ct) acc[ct] = mfma_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k4 = 0; k4 < kPW; k4 += 4) {
    const double av = sC[k4 + lk][rw + lr];
#pragma unroll
    for (int ct = 0; ct < 3; ++ct)
      if (k4 < kW[ct])
        acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, sT[k4 + lk][16 * ct + lr], acc[ct], 0, 0, 0);
  }
#pragma unroll
  for (int ct = 0; ct < 3; ++ct)
#pragma unroll
    for (int q = 0; q < 4; ++q) sW[rw + lk + 4 * q][16 * ct + lr] = acc[ct][q];
  // L = W blockdiag(M_j): column tile ct meets the pivot blocks with rows
  // [kL0[ct], kL1[ct]); the block-diagonal operand is read from the M_j directly
  constexpr int kL0[3] = {0, 12, 30}, kL1[3] = {18, 36, 48};
#pragma unroll
  for (int ct = 0; ct < 3; ++ct) acc[ct] = mfma_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k4 = 0; k4 < kPW; k4 += 4) {
    const int k = k4 + lk;
    const double av = sW[rw + lr][k];
#pragma unroll
    for (int ct = 0; ct < 3; ++ct)
      if (k4 + 4 > kL0[ct] && k4 < kL1[ct]) {
        const int n = 16 * ct + lr;
        const double bv = k / 6 == n / 6 ? pM[k / 6][6 * (k % 6) + n % 6] : 0.0;
        acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[ct], 0, 0, 0);
      }
  }
  // L^T into this wavefront's columns of sC
#pragma unroll
  for (int ct = 0; ct < 3; ++ct)
#pragma unroll
    for (int q = 0; q < 4; ++q) sC[16 * ct + lr][rw + lk + 4 * q] = acc[ct][q];
  // the right-hand side of the wavefront's 16 rows: b_r -= L_r y_P (four lanes per row,
  // every fourth column, combined by two xor shuffles)
  {
    const int i = rw + (lane >> 2), part = lane & 3, r = rbase + i;
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < kPW / 4; ++u) acc += sC[part + 4 * u][i] * sy[wv][part + 4 * u];
    acc += __shfl_xor(acc, 1, 64);
    acc += __shfl_xor(acc, 2, 64);
    if (part == 0 && r < F) {
      if (kend < K || f.sn_parent[X] < 0)
        bx[r] -= acc;
      else  // the node's last step: a finished update row, into the parent's right-hand side
        f.usolve[f.sn_uoff[f.sn_parent[X]] + parent_row(f, X, r - K)] += bx[r] - acc;
    }
  }
  // the wavefront's 16 rows to memory: W and L column-major, and for the node's own rows
  // the row-major copy of L
  for (int e = lane; e < 16 * kPW; e += 64) {
    const int i = rw + e % 16, c = e / 16, r = rbase + i;
    if (c < np && r < F) {
      Wo[int64_t(c) * F + r] = sW[i][c];
      Lo[int64_t(k0 + c) * F + r] = sC[c][i];
    }
  }
  if (rbase + rw < K)
    for (int e = lane; e < 16 * kPW; e += 64) {
      const int c = e % kPW, i = rw + e / kPW, r = rbase + i;
      if (c < np && r < K) Lt[int64_t(r) * K + k0 + c] = sC[c][i];
    }
}

// Trailing updates of one round of split panel steps, one workgroup per 64x64 tile of a
// node's lower trailing region (items as in k_mf_step): A[r, c] -= L_r W_c^T on the
// matrix cores.  The tile's L rows and W columns are staged in
// LDS (zero past the panel's width); each wavefront owns a 32x32 quadrant as 2 x 2 tiles
// of v_mfma_f64_16x16x4, computing (-W) L^T so that each lane's results run down 16
// consecutive rows of the column-major front (coalesced read-modify-write).  The
// accumulators start from the tile's current values.  On the node's last step the results
// go into the parent's front (extend-add).  The tile (0, 0), which holds the next panel's
// diagonal block, then factors that block (on the last child's last step: the parent's
// first one).  Two workgroups per CU (the tiles are many, and short).
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2)))
    k_mf_update(DevFactor f, const int32_t* __restrict__ tiles, DevScalars* sc,
                double* __restrict__ z) {
  static_assert(kTile == 64 && kThreads == 256 && kPW % 4 == 0, "4 wavefronts of 32x32");
  // consecutive tiles (one tile row shares its L rows) run on one XCD
  const int bid = xcd_contiguous(blockIdx.x, gridDim.x);
  const int X = tiles[4 * bid], panel = tiles[4 * bid + 1];
  const int tr = tiles[4 * bid + 2], tc = tiles[4 * bid + 3];
  const int F = 6 * f.sn_nf[X], K = 6 * f.sn_nc[X];
  double* A = f.front + f.sn_off[X];
  const double* Lo = f.lbuf + f.sn_loff[X];
  const double* Wo = f.wbuf + f.sn_woff[X];
  const DiagNode dn = diag_node(f, X);
  const int k0 = kPW * panel;
  const int kend = min(K, k0 + kPW);
  const int np = kend - k0;
  const int r0 = kend + kTile * tr, c0 = kend + kTile * tc;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int lr = lane & 15, lc = lane >> 4;
  const int rb = 32 * (wv >> 1), cb = 32 * (wv & 1);  // the wavefront's quadrant
  __shared__ double sA[kPW][kTile + 1], sB[kPW][kTile + 1];
  __shared__ double sL[kPanelBlocks][42], sdi[kPanelBlocks][6];
  // the node's last panel step: its update matrix goes to the parent (-1: none)
  const int P = kend == K ? f.sn_parent[X] : -1;
  // the last child of its parent (in round order) factors the parent's first diagonal
  // block after its last step's tile (0, 0): update rows keep their order in the parent, so
  // that block gets the children's update only from their tiles (0, 0), and the other
  // children ended in earlier rounds
  const bool hand = P >= 0 && tr == 0 && tc == 0 && f.sn_hand[X];
  bool diag = false;  // this workgroup factors the node's next diagonal block
  // accumulator (mi, nj) component q is the tile element
  //   row rb + 16 mi + lr, column cb + 16 nj + lc + 4 q;
  // it starts from the element's current value (loaded first)
  mfma_d4 acc[2][2];
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int nj = 0; nj < 2; ++nj)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = r0 + rb + 16 * mi + lr, c = c0 + cb + 16 * nj + lc + 4 * q;
        acc[mi][nj][q] = load_or_zero(A, int64_t(c) * F + r, r < F && c < F && r >= c);
      }
  constexpr int kStage = (kPW * kTile + kThreads - 1) / kThreads;
  stage_lds2<kStage>(
      kPW * kTile,
      [&](int e, bool ok) {
        const int t = e / kTile, r = r0 + e % kTile;
        return load_or_zero(Lo, int64_t(k0 + t) * F + r, ok && t < np && r < F);
      },
      [&](int e, double v) { sA[e / kTile][e % kTile] = v; },
      [&](int e, bool ok) {
        const int t = e / kTile, c = c0 + e % kTile;
        return load_or_zero(Wo, int64_t(t) * F + c, ok && t < np && c < F);
      },
      [&](int e, double v) { sB[e / kTile][e % kTile] = v; });
  __syncthreads();
#pragma unroll
  for (int s = 0; s < kPW / 4; ++s) {
    const int t = 4 * s + lc;  // this lane's k index
    double a[2], b[2];
#pragma unroll
    for (int nj = 0; nj < 2; ++nj) a[nj] = -sB[t][cb + 16 * nj + lr];  // A operand: -W (16 x 4)
#pragma unroll
    for (int mi = 0; mi < 2; ++mi) b[mi] = sA[t][rb + 16 * mi + lr];   // B operand: L^T (4 x 16)
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
      for (int nj = 0; nj < 2; ++nj)
        acc[mi][nj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[nj], b[mi], acc[mi][nj], 0, 0, 0);
  }
  if (P < 0) {
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
      for (int nj = 0; nj < 2; ++nj)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = r0 + rb + 16 * mi + lr, c = c0 + cb + 16 * nj + lc + 4 * q;
          if (r < F && c < F && r >= c) A[int64_t(c) * F + r] = acc[mi][nj][q];
        }
  } else if (!hand) {
    extend_tile(f, X, P, K, F, r0 + rb + lr, c0 + cb + lc, acc);
  } else {
    extend_tile_hand(f, X, P, K, F, r0 + rb + lr, c0 + cb + lc, acc, as_square(sB),
                     min(6 * f.sn_nc[P], kPW));
  }
  if (tr == 0 && tc == 0 && kend < K) {
    // the next panel's diagonal block is the top-left of this tile: hand the updated
    // values to LDS directly
    const int npn = min(K, kend + kPW) - kend;
    double (*D)[kPW + 1] = as_square(sA);
    __syncthreads();  // every MFMA operand read of sA / sB is done
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
      for (int nj = 0; nj < 2; ++nj)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int rl = rb + 16 * mi + lr, cl = cb + 16 * nj + lc + 4 * q;
          if (rl < npn && cl < npn) D[rl][cl] = rl >= cl ? acc[mi][nj][q] : 0.0;
        }
    __syncthreads();
    diag = true;
  }
  if (diag) {
    factor_panel_diag(f, dn, A, F, kend, min(K, kend + kPW) - kend, sc, as_square(sA),
                      as_square(sB), sL, sdi, true, z);
  } else if (hand) {
    __syncthreads();  // the work above that reads sA is done
    factor_panel_diag(f, diag_node(f, P), f.front + f.sn_off[P], 6 * f.sn_nf[P], 0,
                      min(6 * f.sn_nc[P], kPW), sc, as_square(sB), as_square(sA), sL, sdi, true, z);
  }
}

// Each node's right-hand side in front row order: its own rows from the permuted RHS, its
// update rows zero (they gather the node's and its descendants' eliminations).
__global__ void __launch_bounds__(kThreads) k_mf_rhs_init(DevFactor f, const double* __restrict__ b) {
  const int X = blockIdx.x;
  const int F = 6 * f.sn_nf[X], K = 6 * f.sn_nc[X], c0 = f.sn_c0[X];
  double* bx = f.usolve + f.sn_uoff[X];
  for (int r = threadIdx.x; r < F; r += blockDim.x) bx[r] = r < K ? b[6 * int64_t(c0) + r] : 0.0;
}

// ---- solves with the supernodal factor ----------------------------------------------------
// The forward solve runs inside the factorization: each node carries its right-hand side
// in front row order (k_mf_rhs_init), every panel step eliminates it with the panel's
// L (k_mf_panel, which also writes z_P = D^-1 y_P for the panel's columns) and the
// extend-add carries the update rows to the parent.  The backward solve, parents first:
// t_K = z_K - L_rect^T x_rows (column tiles), then x_K = L_KK^-T t_K (one workgroup per
// node, one thread per row, K <= kTriMax, blockDim a multiple of 64 covering K), panel by
// panel from the last of the factor's kPW-column panels: the panel's rows publish their
// values and apply the panel's stored L_PP^-T to them (a dot product instead of a
// substitution), then every earlier row applies the panel's L entries to itself.  Each
// thread holds one kPW-wide operand row in registers, loaded a panel ahead: the row of
// L_PP^-T while its row is in the panel, the panel rows' L entries of its column before.
constexpr int kTriMax = 6 * 96;  // largest supernode (96 column blocks)

// t_K = z_K - L_rect^T x_rows for kSolveCols node columns per workgroup (4 columns per
// wavefront, accumulated side by side; lanes stride over the rows below the node);
// items = (node, column tile).
__global__ void __launch_bounds__(kThreads)
    k_mf_bwd_rect(DevFactor f, const int32_t* __restrict__ items, double* __restrict__ y) {
  static_assert(kSolveCols == 4 * (kThreads / 64), "4 columns per wavefront");
  const int X = items[2 * blockIdx.x], tile = items[2 * blockIdx.x + 1];
  const int F = 6 * f.sn_nf[X], K = 6 * f.sn_nc[X], c0 = f.sn_c0[X];
  const double* A = f.lbuf + f.sn_loff[X];
  const int32_t* rows = f.sn_rows + f.sn_rows_ptr[X];
  constexpr int kMaxLds = 6 * 512;
  __shared__ double xr[kMaxLds];
  const int m = F - K;
  const bool in_lds = m <= kMaxLds;
  if (in_lds)
    stage_lds<kMaxLds / kThreads>(
        m,
        [&](int e, bool ok) { return load_or_zero(y, 6 * rows[(K + e) / 6] + (K + e) % 6, ok); },
        [&](int e, double v) { xr[e] = v; });
  __syncthreads();
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int cb = kSolveCols * tile + 4 * wid;
  const double* col[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) col[j] = A + int64_t(min(cb + j, K - 1)) * F + K;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int i = lane; i < m; i += 64) {
    const double xv = in_lds ? xr[i] : y[6 * rows[(K + i) / 6] + (K + i) % 6];
#pragma unroll
    for (int j = 0; j < 4; ++j) s[j] += col[j][i] * xv;
  }
#pragma unroll
  for (int j = 0; j < 4; ++j)
    for (int off = 32; off > 0; off >>= 1) s[j] += __shfl_down(s[j], off, 64);
  if (lane == 0)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (cb + j < K) y[6 * c0 + cb + j] -= s[j];
}

// x_K = L_KK^-T t_K: one thread per node column, panels last to first, reading the
// row-major copy of L_KK (thread c takes L[k0 + t, c]; lanes read consecutive columns).
__global__ void __launch_bounds__(kTriMax)
    k_mf_bwd_tri(DevFactor f, const int32_t* __restrict__ nodes, double* __restrict__ y) {
  const int X = nodes[blockIdx.x];
  const int K = 6 * f.sn_nc[X], c0 = f.sn_c0[X];
  const double* Lt = f.ltbuf + f.sn_toff[X];
  const double* Tn = f.tinv + f.sn_poff[X];
  __shared__ double xl[kPW], xs[kTriMax];
  const int c = threadIdx.x;
  const bool own = c < K;
  double x = own ? y[6 * c0 + c] : 0.0;
  const int npan = (K + kPW - 1) / kPW;
  for (int e = c; e < kTriMax; e += blockDim.x) xs[e] = 0.0;
  // the operand row for panel p: row c - k0 of L_PP^-T in the panel, the panel rows'
  // entries of this column (row-major copy of L) above it
  double cur[kPW];
  auto load_panel = [&](int p) {
    const int k0 = kPW * p, npp = min(kPW, K - k0);
    const bool in_panel = c >= k0 && c < k0 + npp;
    const double* Tg = Tn + int64_t(p) * kPW * kPW;
    // wavefronts with a column in or above the panel: one load per element from a
    // selected (always valid) address and no select on the value, so the loads stay in
    // flight across the next barriers; the entries past the panel's width meet zeros
    // (L_PP^-T is zero there, xs past K)
    if (__any(in_panel || c < k0)) {
#pragma unroll
      for (int i = 0; i < kPW; ++i) {
        const uintptr_t at = reinterpret_cast<uintptr_t>(Tg + (in_panel ? (c - k0) * kPW : 0) + i);
        const uintptr_t al =
            reinterpret_cast<uintptr_t>(Lt + int64_t(k0 + min(i, npp - 1)) * K + (own ? c : 0));
        cur[i] = *reinterpret_cast<const double*>(in_panel ? at : al);
      }
    }
  };
  load_panel(npan - 1);
  for (int p = npan - 1; p >= 0; --p) {
    const int k0 = kPW * p, npp = min(kPW, K - k0);
    const bool in_panel = c >= k0 && c < k0 + npp;
    if (in_panel) xl[c - k0] = x;
    __syncthreads();
    if (in_panel) {
      double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int i = 0; i < kPW; ++i) s[i & 3] += cur[i] * xl[min(i, npp - 1)];  // 0 past npp
      x = (s[0] + s[1]) + (s[2] + s[3]);
      xs[c] = x;
    }
    __syncthreads();
    if (c < k0) {
      double s[4] = {x, 0.0, 0.0, 0.0};
#pragma unroll
      for (int i = 0; i < kPW; ++i) s[i & 3] -= cur[i] * xs[k0 + i];  // xs is 0 past K
      x = (s[0] + s[1]) + (s[2] + s[3]);
    }
    if (p > 0) load_panel(p - 1);
  }
  if (own) y[6 * c0 + c] = x;
}

// Forward solve of one level's nodes (one workgroup each) for a right-hand side already in
// the nodes' front row order: first the children's finished update rows are added into the
// node's (children one after another, so the adds are plain), then panel by panel y_P =
// L_PP^-1 b_P (the stored L_PP^-T), z_P = blockdiag(M_j) y_P and b_r -= L_r y_P for every
// row below the panel, the same eliminations the factorization applies to the step's own
// right-hand side.
__global__ void __launch_bounds__(kThreads)
    k_mf_fwd(DevFactor f, const int32_t* __restrict__ nodes, double* __restrict__ z) {
  const int X = nodes[blockIdx.x];
  const int F = 6 * f.sn_nf[X], K = 6 * f.sn_nc[X], c0 = f.sn_c0[X];
  double* b = f.usolve + f.sn_uoff[X];
  const double* L = f.lbuf + f.sn_loff[X];
  const double* Tn = f.tinv + f.sn_poff[X];
  const int tid = threadIdx.x;
  __shared__ double sb[kPW], sy[kPW];
  for (int ci = f.sn_child_ptr[X]; ci < f.sn_child_ptr[X + 1]; ++ci) {
    const int C = f.sn_child[ci];
    const int FC = 6 * f.sn_nf[C], KC = 6 * f.sn_nc[C];
    const double* bc = f.usolve + f.sn_uoff[C];
    for (int u = tid; u < FC - KC; u += blockDim.x) b[parent_row(f, C, u)] += bc[KC + u];
    __syncthreads();
  }
  const int npan = (K + kPW - 1) / kPW;
  for (int p = 0; p < npan; ++p) {
    const int k0 = kPW * p, np = min(kPW, K - k0);
    const double* Tg = Tn + int64_t(p) * kPW * kPW;
    if (tid < np) sb[tid] = b[k0 + tid];
    __syncthreads();
    if (tid < np) {
      double y = 0.0;
      for (int t = 0; t < np; ++t) y += Tg[t * kPW + tid] * sb[t];  // L_PP^-T (t, c) = L_PP^-1 (c, t)
      sy[tid] = y;
    }
    __syncthreads();
    if (tid < np) {
      const int jb = tid / 6, a = tid % 6;
      const double* M = f.piv_M + 36 * int64_t(f.sn_c0[X] + k0 / 6 + jb);
      double zz = 0.0;
#pragma unroll
      for (int t = 0; t < 6; ++t) zz += M[6 * a + t] * sy[6 * jb + t];
      z[6 * int64_t(c0) + k0 + tid] = zz;
    }
    for (int r = k0 + np + tid; r < F; r += blockDim.x) {
      double acc = 0.0;
      for (int t = 0; t < np; ++t) acc += L[int64_t(k0 + t) * F + r] * sy[t];
      b[r] -= acc;
    }
    __syncthreads();
  }
}

// Coupling blocks of one point with its cameras (covariance queries): one thread per
// observation.  H_{c,b} = (A P)^T (-A) = -P^T M with P = [hat(-x) | I], M = A^T A.
template <int kModel>
__global__ void k_point_coupling(DevProblem p, const double* __restrict__ points,
                                 const double* __restrict__ intr, const double* __restrict__ cw,
                                 const double* __restrict__ prec, int b, double* __restrict__ out,
                                 int32_t* __restrict__ cam, double* __restrict__ hinv) {
  if (threadIdx.x == 0) {
    const double* h = prec + kPtRec * int64_t(b) + 3;
    const double hv[9] = {h[0], h[1], h[2], h[1], h[3], h[4], h[2], h[4], h[5]};
    for (int e = 0; e < 9; ++e) hinv[e] = hv[e];
  }
  for (int q = threadIdx.x; p.pt_ptr[b] + q < p.pt_ptr[b + 1]; q += blockDim.x) {
    const int o = p.pt_ptr[b] + q;
    const int pm = p.obs_pose[o];
    cam[q] = p.pose_block[pm];
    const double x[3] = {points[3 * b], points[3 * b + 1], points[3 * b + 2]};
    const double* c = cw + 12 * pm;
    double xc[3], A[6], M[9];
    to_camera(c, x, xc);
    model_jacobian<kModel, double>(c, intr + intr_stride<kModel>() * p.obs_intr[o], xc, A);
    gram(A, M);
    // P^T M: rows 0-2 hat(-x)^T M = hat(x) M, rows 3-5 M
    double* H = out + 18 * q;
    for (int j = 0; j < 3; ++j) {
      H[0 * 3 + j] = -(-x[2] * M[3 + j] + x[1] * M[6 + j]);
      H[1 * 3 + j] = -(x[2] * M[j] - x[0] * M[6 + j]);
      H[2 * 3 + j] = -(-x[1] * M[j] + x[0] * M[3 + j]);
      H[9 + j] = -M[j];
      H[12 + j] = -M[3 + j];
      H[15 + j] = -M[6 + j];
    }
  }
}

__global__ void k_permute_in(int nb, const int32_t* __restrict__ perm, const double* __restrict__ src,
                             double* __restrict__ dst) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < 6 * nb) dst[e] = src[6 * perm[e / 6] + e % 6];
}

__global__ void k_permute_out(int nb, const int32_t* __restrict__ perm, const double* __restrict__ src,
                              double* __restrict__ dst) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < 6 * nb) dst[6 * perm[e / 6] + e % 6] = src[e];
}

// k_permute_out for SparseLdlt's pivot rule.  Its solve divides by every pivot
// (sparse_ldlt.rs:283-286), so a zero pivot's component is y / 0 there; it is written as
// NaN here.  That is the reference's result for a pivot of a block nothing couples to (a
// free pose no term touches: y = 0, and 0 / 0 is NaN, while its empty rows and columns of L
// keep it out of every other component).  Where a coupled variable's pivot is zero, the
// reference's y / 0 may be +-inf and its back-substitution carries it into the components
// coupled to it; those stay finite here.
__global__ void k_permute_out_scalar_rule(int nb, const int32_t* __restrict__ perm,
                                          const double* __restrict__ piv_d,
                                          const double* __restrict__ src, double* __restrict__ dst) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < 6 * nb)
    dst[6 * perm[e / 6] + e % 6] = piv_d[e] == 0.0 ? __builtin_nan("") : src[e];
}

__global__ void k_scatter_s(DevProblem p, int64_t nslots, const double* __restrict__ blocks,
                            double* __restrict__ front) {
  const int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (e >= nslots * 36) return;
  const int64_t k = e / 36;
  const int q = int(e % 36), a = q / 6, b = q % 6;
  const int ldc = p.s_ld[k], ld = ldc >> 1;
  front[p.s_off[k] + ((ldc & 1) ? a * ld + b : b * ld + a)] += blocks[e];
}

// block reduction of two doubles into partial[2 * blockIdx.x + {0, 1}]
__device__ void block_reduce2(double a, double b, double* partial) {
  __shared__ double sa[kThreads / 64], sb[kThreads / 64];
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    b += __shfl_down(b, off, 64);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    sa[wid] = a;
    sb[wid] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ta = 0, tb = 0;
    for (int w = 0; w < int(blockDim.x / 64); ++w) {
      ta += sa[w];
      tb += sb[w];
    }
    partial[2 * blockIdx.x] = ta;
    partial[2 * blockIdx.x + 1] = tb;
  }
}

// ---- back-substitution of the points and their update ------------------------------------
// kFreeIntr: free shared intrinsics, dx_b -= V_b^T dx_intr (H_ll^-1 H_lc dx_intr, H_lc = W_b^T)
template <typename T, int kModel, bool kFreeIntr>
__global__ void __launch_bounds__(kThreads)
    k_backsub(DevProblem p, const double* __restrict__ points, const double* __restrict__ intr,
              const double* __restrict__ cw, const double* __restrict__ prec,
              const double* __restrict__ pt_s, const double* __restrict__ dx_cam,
              const double* __restrict__ pt_vw, double* __restrict__ points_new,
              double* __restrict__ partial, DevScalars* sc) {
  double dd = 0, dg = 0, mx = 0;
  for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < p.npt; b += gridDim.x * blockDim.x) {
    const double xd[3] = {points[3 * b], points[3 * b + 1], points[3 * b + 2]};
    if (p.pt_block[b] < 0) {
      for (int k = 0; k < 3; ++k) points_new[3 * b + k] = xd[k];
      continue;
    }
    T acc[3] = {T(pt_s[3 * b]), T(pt_s[3 * b + 1]), T(pt_s[3 * b + 2])};
    for (int o = p.pt_ptr[b]; o < p.pt_ptr[b + 1]; ++o) {
      const int pm = p.obs_pose[o];
      const int c = p.pose_block[pm];
      if (c < 0) continue;
      const double* cwo = cw + 12 * pm;
      double xc[3];
      to_camera(cwo, xd, xc);
      T A[6], M[9];
      model_jacobian<kModel, T>(cwo, intr + intr_stride<kModel>() * p.obs_intr[o], xc, A);
      gram(A, M);
      const double* dx = dx_cam + 6 * c;
      // P dx = omega x point + nu, in FP64 (omega x point and nu nearly cancel far from
      // the world origin)
      const T q0 = T(dx[1] * xd[2] - dx[2] * xd[1] + dx[3]);
      const T q1 = T(dx[2] * xd[0] - dx[0] * xd[2] + dx[4]);
      const T q2 = T(dx[0] * xd[1] - dx[1] * xd[0] + dx[5]);
#pragma unroll
      for (int k = 0; k < 3; ++k) acc[k] += M[3 * k] * q0 + M[3 * k + 1] * q1 + M[3 * k + 2] * q2;
    }
    // H_ll^-1 from the point's packed record (upper triangle 00, 01, 02, 11, 12, 22)
    const double* hu = prec + kPtRec * int64_t(b) + 3;
    const double Hi[9] = {hu[0], hu[1], hu[2], hu[1], hu[3], hu[4], hu[2], hu[4], hu[5]};
    for (int k = 0; k < 3; ++k) {
      double dk = Hi[3 * k] * double(acc[0]) + Hi[3 * k + 1] * double(acc[1]) +
                  Hi[3 * k + 2] * double(acc[2]);
      if constexpr (kFreeIntr) {
        const double* V = pt_vw + 36 * int64_t(b);
        const double* dxi = dx_cam + 6 * p.intr_block;
        for (int a = 0; a < 6; ++a) dk -= V[3 * a + k] * dxi[a];
      }
      points_new[3 * b + k] = xd[k] + dk;
      dd += dk * dk;
      dg += dk * pt_s[3 * b + k];
      mx = fmax(mx, fabs(dk));
    }
  }
  wave_atomic_max_abs(&sc->max_delta_bits, mx);
  block_reduce2(dd, dg, partial);
}

// ---- free shared intrinsics (EUCM) -----------------------------------------------------------
// One active intrinsics member c shared by the observations adds the free block c to the RCS.
// With E_o = dx_distort_params (J_intr = -E_o), J_point = -A_o and J_pose = A_o P_b:
//   H_cc = sum E^T E,  H_cl = W_b = sum_o E_o^T A_o (6x3),  H_cp = -E^T A P,
//   neg_grad_c = sum E^T r~.
// With V_b = W_b H_ll^-1 the Schur complement's intrinsics rows are
//   S_cc = H_cc + nu n_c I - sum_b V_b W_b^T,
//   S_cp = sum_{o of camera p} (V_b M_o - E_o^T A_o) P_b,
//   rhs_c = neg_grad_c - sum_b V_b s_b,
// and the back-substitution subtracts V_b^T dx_c from each point's step.  Terms whose
// intrinsics member is not the free one are conditioned for it.
constexpr int kIntrPart = 33;  // per workgroup: S_cc upper (21), rhs_c (6), neg_grad_c (6)

// deterministic workgroup sum of N per-thread values: wavefront butterfly, then the
// wavefronts in order; thread e < N of the workgroup receives total e in out
template <int N>
__device__ __forceinline__ void block_sum_n(double (&v)[N], double (*red)[N], double* out) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int e = 0; e < N; ++e) {
    double x = v[e];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane == 0) red[wid][e] = x;
  }
  __syncthreads();
  if (int(threadIdx.x) < N) {
    double t = 0.0;
    for (int w = 0; w < int(blockDim.x >> 6); ++w) t += red[w][threadIdx.x];
    *out = t;
  }
}

// Per point: W_b and V_b (pt_vw: V row-major 6x3, then W), and the point's share of S_cc,
// rhs_c and neg_grad_c summed over the workgroup into intr_part[kIntrPart * block + e].
// G lanes per point (G divides 64) stride over its observations; their sums are combined by
// a butterfly over the G lanes, after which every lane of the group holds the same totals
// (each addition's operands are the same pair on both lanes), and the group's first lane
// finishes the point.
template <int G>
__global__ void __launch_bounds__(kThreads)
    k_intr_points(DevProblem p, const double* __restrict__ points, const double* __restrict__ intr,
                  const double* __restrict__ cw, const double* __restrict__ prec,
                  const double* __restrict__ pt_s, double* __restrict__ pt_vw,
                  double* __restrict__ intr_part) {
  __shared__ double red[kThreads / 64][kIntrPart];
  constexpr int kSum = 18 + 21 + 6;  // W, sum E^T E (upper), sum E^T r~
  const int groups = blockDim.x / G, sub = threadIdx.x % G;
  double acc[kIntrPart];
#pragma unroll
  for (int e = 0; e < kIntrPart; ++e) acc[e] = 0.0;
  // the loop bound is uniform over the workgroup, so every lane of a group shuffles together
  for (int b0 = blockIdx.x * groups; b0 < p.npt; b0 += gridDim.x * groups) {
    const int b = b0 + int(threadIdx.x) / G;
    const bool live = b < p.npt;
    double t[kSum];
#pragma unroll
    for (int e = 0; e < kSum; ++e) t[e] = 0.0;
    double x[3] = {0.0, 0.0, 0.0};
    if (live) {
      x[0] = points[3 * b];
      x[1] = points[3 * b + 1];
      x[2] = points[3 * b + 2];
      const double* in = intr + 6 * int64_t(p.intr_member);
      for (int o = p.pt_ptr[b] + sub; o < p.pt_ptr[b + 1]; o += G) {
        if (p.obs_intr[o] != p.intr_member) continue;
        const double* c = cw + 12 * p.obs_pose[o];
        double xc[3], px[2], D[6], E[12], A[6];
        to_camera(c, x, xc);
        eucm_project(in, xc, px, D, E);
        lift_rotation(c, D, A);
        const double r0 = p.obs_uv[2 * o] - px[0], r1 = p.obs_uv[2 * o + 1] - px[1];
        const double w = robust_weight(p.robust, p.delta, sqrt(r0 * r0 + r1 * r1));
        const double rw0 = p.robust == 1 ? w * r0 : r0, rw1 = p.robust == 1 ? w * r1 : r1;
        int u = 18;
#pragma unroll
        for (int a = 0; a < 6; ++a) {
#pragma unroll
          for (int j = 0; j < 3; ++j) t[3 * a + j] += E[a] * A[j] + E[6 + a] * A[3 + j];
#pragma unroll
          for (int bb = a; bb < 6; ++bb) t[u++] += E[a] * E[bb] + E[6 + a] * E[6 + bb];
          t[39 + a] += E[a] * rw0 + E[6 + a] * rw1;
        }
      }
    }
#pragma unroll
    for (int off = 1; off < G; off <<= 1)
#pragma unroll
      for (int e = 0; e < kSum; ++e) t[e] += __shfl_xor(t[e], off, 64);
    if (live && sub == 0) {
      const double* W = t;
#pragma unroll
      for (int e = 0; e < 21; ++e) acc[e] += t[18 + e];
#pragma unroll
      for (int a = 0; a < 6; ++a) {
        acc[21 + a] += t[39 + a];
        acc[27 + a] += t[39 + a];
      }
      double* out = pt_vw + 36 * int64_t(b);
      if (p.pt_block[b] >= 0) {
        const double* h = prec + kPtRec * int64_t(b) + 3;
        const double Hi[9] = {h[0], h[1], h[2], h[1], h[3], h[4], h[2], h[4], h[5]};
        double V[18];
#pragma unroll
        for (int a = 0; a < 6; ++a)
#pragma unroll
          for (int j = 0; j < 3; ++j)
            V[3 * a + j] = W[3 * a] * Hi[j] + W[3 * a + 1] * Hi[3 + j] + W[3 * a + 2] * Hi[6 + j];
        int u = 0;
#pragma unroll
        for (int a = 0; a < 6; ++a)
#pragma unroll
          for (int bb = a; bb < 6; ++bb)
            acc[u++] -= V[3 * a] * W[3 * bb] + V[3 * a + 1] * W[3 * bb + 1] + V[3 * a + 2] * W[3 * bb + 2];
        const double sb[3] = {pt_s[3 * b], pt_s[3 * b + 1], pt_s[3 * b + 2]};
#pragma unroll
        for (int a = 0; a < 6; ++a)
          acc[21 + a] -= V[3 * a] * sb[0] + V[3 * a + 1] * sb[1] + V[3 * a + 2] * sb[2];
#pragma unroll
        for (int e = 0; e < 18; ++e) {
          out[e] = V[e];
          out[18 + e] = W[e];
        }
      } else {
#pragma unroll
        for (int e = 0; e < 36; ++e) out[e] = 0.0;
      }
    }
  }
  double total = 0.0;
  block_sum_n<kIntrPart>(acc, red, &total);
  if (int(threadIdx.x) < kIntrPart) intr_part[kIntrPart * int64_t(blockIdx.x) + threadIdx.x] = total;
}

// The S block (intrinsics, camera) per chunk of at most kOwnChunk camera-major observations
// (the own-term chunks), one wavefront per chunk, lanes striding over its observations:
// chunk sums into cpart[36 * chunk + e] (a butterfly leaves every lane with the totals).
__global__ void __launch_bounds__(64)
    k_intr_cam_chunks(DevProblem p, const double* __restrict__ points,
                      const double* __restrict__ intr, const double* __restrict__ cw,
                      const double* __restrict__ pt_vw, double* __restrict__ cpart) {
  const int ch = blockIdx.x, lane = threadIdx.x;
  const int cam = p.own_chunks[3 * ch], q0 = p.own_chunks[3 * ch + 1], q1 = p.own_chunks[3 * ch + 2];
  double acc[36];
#pragma unroll
  for (int e = 0; e < 36; ++e) acc[e] = 0.0;
  const bool any = p.intr_slot[cam] >= 0;
  for (int q = q0 + lane; any && q < q1; q += 64) {
    const int o = p.cam_obs[q];
    if (p.obs_intr[o] != p.intr_member) continue;
    const double* c = cw + 12 * p.cam_member[cam];
    const double* in = intr + 6 * int64_t(p.intr_member);
    const int b = p.obs_point[o];
    const double x[3] = {points[3 * b], points[3 * b + 1], points[3 * b + 2]};
    double xc[3], px[2], D[6], E[12], A[6];
    to_camera(c, x, xc);
    eucm_project(in, xc, px, D, E);
    lift_rotation(c, D, A);
    double G[18];
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
      for (int j = 0; j < 3; ++j) G[3 * a + j] = -(E[a] * A[j] + E[6 + a] * A[3 + j]);
    if (p.pt_block[b] >= 0) {
      const double* V = pt_vw + 36 * int64_t(b);
      double M[9];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) M[3 * i + j] = A[i] * A[j] + A[3 + i] * A[3 + j];
#pragma unroll
      for (int a = 0; a < 6; ++a)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          G[3 * a + j] += V[3 * a] * M[j] + V[3 * a + 1] * M[3 + j] + V[3 * a + 2] * M[6 + j];
    }
    // G P = [G hat(-x) | G] = [-(G hat(x)) | G]
#pragma unroll
    for (int a = 0; a < 6; ++a) {
      const double g0 = G[3 * a], g1 = G[3 * a + 1], g2 = G[3 * a + 2];
      acc[6 * a + 0] -= g1 * x[2] - g2 * x[1];
      acc[6 * a + 1] -= g2 * x[0] - g0 * x[2];
      acc[6 * a + 2] -= g0 * x[1] - g1 * x[0];
      acc[6 * a + 3] += g0;
      acc[6 * a + 4] += g1;
      acc[6 * a + 5] += g2;
    }
  }
  double mine = 0.0;
#pragma unroll
  for (int e = 0; e < 36; ++e) {
    double v = acc[e];
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) v += __shfl_xor(v, off, 64);
    if (lane == e) mine = v;
  }
  if (lane < 36) cpart[36 * int64_t(ch) + lane] = mine;
}

// One wavefront per pose block: its chunks' sums in chunk order, stored into the slot
// (intrinsics, camera) (k_schur_blocks leaves it zero).
__global__ void __launch_bounds__(kThreads)
    k_intr_cam_store(DevProblem p, int npb, const double* __restrict__ cpart,
                     double* __restrict__ front, double* __restrict__ bsr) {
  const int cam = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (cam >= npb || lane >= 36) return;
  const int slot = p.intr_slot[cam];
  if (slot < 0) return;
  double v = 0.0;
  for (int ch = p.cam_chunk_ptr[cam]; ch < p.cam_chunk_ptr[cam + 1]; ++ch) v += cpart[36 * int64_t(ch) + lane];
  const int a = lane / 6, bb = lane % 6;
  if (bsr) {
    bsr[36 * int64_t(slot) + lane] = v;
  } else {
    const int ldc = p.s_ld[slot], ld = ldc >> 1;
    front[p.s_off[slot] + ((ldc & 1) ? a * ld + bb : bb * ld + a)] = v;
  }
}

// Sums k_intr_points' partials (fixed order) into the intrinsics' diagonal S block (added to
// the damping k_schur_blocks stored there), rhs_c and neg_grad_c.
__global__ void __launch_bounds__(kThreads)
    k_intr_finish(DevProblem p, const double* __restrict__ intr_part, int nblk,
                  double* __restrict__ front, double* __restrict__ bsr, double* __restrict__ rhs_nat,
                  double* __restrict__ grad_cam) {
  __shared__ double tot[kIntrPart];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int e = wid; e < kIntrPart; e += nw) {
    double v = 0.0;
    for (int i = lane; i < nblk; i += 64) v += intr_part[kIntrPart * int64_t(i) + e];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) tot[e] = v;
  }
  __syncthreads();
  const int t = threadIdx.x, c = p.intr_block;
  if (t < 36) {
    const int a = t / 6, bb = t % 6, lo = min(a, bb), hi = max(a, bb);
    const double v = tot[lo * 6 - lo * (lo - 1) / 2 + (hi - lo)];
    const int64_t k = p.intr_diag_slot;
    if (bsr) {
      bsr[36 * k + t] += v;
    } else {
      const int ldc = p.s_ld[k], ld = ldc >> 1;
      front[p.s_off[k] + ((ldc & 1) ? a * ld + bb : bb * ld + a)] += v;
    }
  } else if (t < 42) {
    rhs_nat[6 * c + t - 36] = tot[21 + t - 36];
  } else if (t < 48) {
    grad_cam[6 * c + t - 42] = tot[27 + t - 42];
  }
}

// ---- pose update T <- exp(delta) T ----------------------------------------------------------
// Pose blocks 0 .. npb - 1 (the free intrinsics' block, if any, follows them).
__global__ void __launch_bounds__(kThreads)
    k_update_poses(DevProblem p, int npb, const double* __restrict__ poses,
                   const double* __restrict__ dx_cam, const double* __restrict__ grad_cam,
                   double* __restrict__ poses_new, double* __restrict__ partial, DevScalars* sc) {
  double dd = 0, dg = 0;
  double mx = 0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < npb; c += gridDim.x * blockDim.x) {
    const int pm = p.cam_member[c];
    double t[6], T0[7], T1[7];
    for (int k = 0; k < 6; ++k) {
      t[k] = dx_cam[6 * c + k];
      dd += t[k] * t[k];
      dg += t[k] * grad_cam[6 * c + k];
      mx = fmax(mx, fabs(t[k]));
    }
    for (int k = 0; k < 7; ++k) T0[k] = poses[7 * pm + k];
    se3_oplus(T0, t, T1);
    for (int k = 0; k < 7; ++k) poses_new[7 * pm + k] = T1[k];
  }
  wave_atomic_max_abs(&sc->max_delta_bits, mx);
  block_reduce2(dd, dg, partial);
}

// Free intrinsics: additive update (perspective_camera.rs:85-88) into intr_new, which holds a
// copy of intr; the block's share of delta.delta, delta.neg_grad and max |delta| goes to
// partial[0..1] (one wavefront).
__global__ void k_update_intrinsics(DevProblem p, const double* __restrict__ dx_cam,
                                    const double* __restrict__ grad_cam,
                                    const double* __restrict__ intr, double* __restrict__ intr_new,
                                    double* __restrict__ partial, DevScalars* sc) {
  const int k = threadIdx.x, c = p.intr_block;
  double dd = 0, dg = 0, mx = 0;
  if (k < 6) {
    const double t = dx_cam[6 * c + k];
    const int64_t i = int64_t(6) * p.intr_member + k;
    intr_new[i] = intr[i] + t;
    dd = t * t;
    dg = t * grad_cam[6 * c + k];
    mx = fabs(t);
  }
  wave_atomic_max_abs(&sc->max_delta_bits, mx);
  for (int off = 32; off > 0; off >>= 1) {
    dd += __shfl_down(dd, off, 64);
    dg += __shfl_down(dg, off, 64);
  }
  if (k == 0) {
    partial[0] = dd;
    partial[1] = dg;
  }
}

// ---- equality constraints: range-space KKT -------------------------------------------------
// TranslationNormConstraint (ba_scale_constraint.rs:44-99): c = |t| - r; with T <- exp(x) T,
// dt/dx = [-hat(t) | I] at x = 0, so dc/dx = [t^T (-hat(t)) | t^T] / |t| = [0 | t^T / |t|].
// One workgroup: each term's G and c, each set's c summed over its terms in order, and (with
// add_rhs) G^T lambda added to the free rows, one thread in term order.
__global__ void k_eq_eval(DevProblem p, DevEq q, const double* __restrict__ poses,
                          const double* __restrict__ lambda, double* __restrict__ rhs_nat,
                          double* __restrict__ grad_cam, int add_rhs) {
  for (int t = threadIdx.x; t < q.nterm; t += blockDim.x) {
    const double* T = poses + 7 * int64_t(q.term_member[t]);
    const double n = sqrt(T[4] * T[4] + T[5] * T[5] + T[6] * T[6]);
    double* g = q.G + 7 * int64_t(t);
    g[0] = g[1] = g[2] = 0.0;
    g[3] = T[4] / n;
    g[4] = T[5] / n;
    g[5] = T[6] / n;
    g[6] = n - q.radius[t];
  }
  __syncthreads();
  for (int k = threadIdx.x; k < q.nset; k += blockDim.x) {
    double c = 0.0;
    for (int t = q.set_ptr[k]; t < q.set_ptr[k + 1]; ++t) c += q.G[7 * int64_t(t) + 6];
    q.c[k] = c;
  }
  if (add_rhs && threadIdx.x == 0)
    for (int k = 0; k < q.nset; ++k)
      for (int t = q.set_ptr[k]; t < q.set_ptr[k + 1]; ++t) {
        const int blk = p.pose_block[q.term_member[t]];
        if (blk < 0) continue;
        for (int i = 0; i < 6; ++i) {
          const double v = q.G[7 * int64_t(t) + i] * lambda[k];
          rhs_nat[6 * int64_t(blk) + i] += v;
          grad_cam[6 * int64_t(blk) + i] += v;
        }
      }
}

// b = G_k^T in camera order (zeros but the set's pose blocks), one workgroup.
__global__ void k_eq_column(DevProblem p, DevEq q, int k) {
  for (int64_t e = threadIdx.x; e < 6 * int64_t(p.nb); e += blockDim.x) q.b[e] = 0.0;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int t = q.set_ptr[k]; t < q.set_ptr[k + 1]; ++t) {
      const int blk = p.pose_block[q.term_member[t]];
      if (blk < 0) continue;
      for (int i = 0; i < 6; ++i) q.b[6 * int64_t(blk) + i] += q.G[7 * int64_t(t) + i];
    }
}

// G_k . v over the set's pose blocks (v in camera order)
__device__ double eq_dot(const DevProblem& p, const DevEq& q, int k, const double* v) {
  double acc = 0.0;
  for (int t = q.set_ptr[k]; t < q.set_ptr[k + 1]; ++t) {
    const int blk = p.pose_block[q.term_member[t]];
    if (blk < 0) continue;
    for (int i = 0; i < 6; ++i) acc += q.G[7 * int64_t(t) + i] * v[6 * int64_t(blk) + i];
  }
  return acc;
}

// One workgroup: M = G Y, its inverse (Gauss-Jordan with partial pivoting; a zero or
// non-finite pivot is the reference's SingularKktConstraint), dlambda = M^-1 (G u + c),
// dx = u - Y dlambda, lambda_next, and the multiplier rows of the step sums into partial[0..1].
__global__ void __launch_bounds__(kThreads)
    k_eq_kkt(DevProblem p, DevEq q, const double* __restrict__ lambda,
             double* __restrict__ lambda_next, double* __restrict__ dx_cam,
             double* __restrict__ partial, DevScalars* sc) {
  __shared__ double M[kMaxEqSets][2 * kMaxEqSets], w[kMaxEqSets], dl[kMaxEqSets];
  __shared__ int singular;
  const int n = q.nset, tid = threadIdx.x;
  const int64_t nf = 6 * int64_t(p.nb);
  for (int e = tid; e < n * n; e += blockDim.x) {
    const int a = e / n, b = e % n;
    M[a][b] = eq_dot(p, q, a, q.Y + nf * b);
    M[a][n + b] = a == b ? 1.0 : 0.0;
  }
  for (int a = tid; a < n; a += blockDim.x) w[a] = eq_dot(p, q, a, dx_cam) + q.c[a];
  if (tid == 0) singular = 0;
  __syncthreads();
  if (tid == 0) {
    for (int c = 0; c < n && !singular; ++c) {
      int piv = c;
      for (int r = c + 1; r < n; ++r)
        if (fabs(M[r][c]) > fabs(M[piv][c])) piv = r;
      const double pv = M[piv][c];
      if (pv == 0.0 || !isfinite(pv)) {
        singular = 1;
        break;
      }
      for (int k = 0; k < 2 * n; ++k) {
        const double tmp = M[c][k];
        M[c][k] = M[piv][k];
        M[piv][k] = tmp;
      }
      for (int k = 0; k < 2 * n; ++k) M[c][k] /= pv;
      for (int r = 0; r < n; ++r) {
        if (r == c) continue;
        const double fct = M[r][c];
        for (int k = 0; k < 2 * n; ++k) M[r][k] -= fct * M[c][k];
      }
    }
    double dd = 0.0, dg = 0.0, mx = 0.0;
    for (int a = 0; a < n; ++a) {
      double v = 0.0;
      if (!singular)
        for (int b = 0; b < n; ++b) v += M[a][n + b] * w[b];
      dl[a] = v;
      dd += v * v;
      dg += v * -q.c[a];
      mx = fmax(mx, fabs(v));
    }
    if (singular) sc->bad_pivot = 1;
    partial[0] = dd;
    partial[1] = dg;
    if (mx > 0.0) atomicMax(&sc->max_delta_bits, (unsigned long long)__double_as_longlong(mx));
  }
  __syncthreads();
  for (int64_t e = tid; e < nf; e += blockDim.x) {
    double v = dx_cam[e];
    for (int b = 0; b < n; ++b) v -= q.Y[nf * b + e] * dl[b];
    dx_cam[e] = v;
  }
  for (int a = tid; a < n; a += blockDim.x) lambda_next[a] = lambda[a] - dl[a];
}

// ---- merit: sum over terms of r~^T r~ ------------------------------------------------------
template <int kModel>
__global__ void __launch_bounds__(kThreads)
    k_cost(DevProblem p, const int32_t* __restrict__ obs_point, const double* __restrict__ poses,
           const double* __restrict__ points, const double* __restrict__ intr,
           double* __restrict__ partial) {
  double acc = 0;
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < p.nobs; o += gridDim.x * blockDim.x) {
    const double* T = poses + 7 * p.obs_pose[o];
    const double* in = intr + intr_stride<kModel>() * p.obs_intr[o];
    const int b = obs_point[o];
    double inv[7], R[9], xc[3];
    se3_inverse(T, inv);
    so3_matrix(inv, R);
    const double x[3] = {points[3 * b], points[3 * b + 1], points[3 * b + 2]};
    mat3_vec(R, x, xc);
    xc[0] += inv[4];
    xc[1] += inv[5];
    xc[2] += inv[6];
    double pxl[2];
    model_pixel<kModel>(in, xc, pxl);
    const double r0 = p.obs_uv[2 * o] - pxl[0];
    const double r1 = p.obs_uv[2 * o + 1] - pxl[1];
    if (p.robust == 1) {
      const double w = robust_weight(1, p.delta, sqrt(r0 * r0 + r1 * r1));
      const double a = w * r0, b2 = w * r1;
      acc += a * a + b2 * b2;
    } else {
      acc += r0 * r0 + r1 * r1;
    }
  }
  block_reduce2(acc, 0.0, partial);
}

// Sum partials in a fixed order (strided per thread, then a fixed tree); finish the
// step scalars.
__device__ double block_sum_fixed(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (int(threadIdx.x) < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(256)
    k_finalize(const double* __restrict__ partial, int n_pt_blocks, int n_pose_blocks,
               int n_cost_blocks, const double* __restrict__ grad_cam, int nb, DevScalars* sc,
               int what) {
  __shared__ double red[256];
  if (what == 0) {  // dd, dg from back-substitution + pose update partials
    double dd = 0, dg = 0;
    for (int i = threadIdx.x; i < n_pt_blocks; i += blockDim.x) {
      dd += partial[2 * i];
      dg += partial[2 * i + 1];
    }
    const double* q = partial + 2 * n_pt_blocks;
    for (int i = threadIdx.x; i < n_pose_blocks; i += blockDim.x) {
      dd += q[2 * i];
      dg += q[2 * i + 1];
    }
    dd = block_sum_fixed(dd, red);
    dg = block_sum_fixed(dg, red);
    if (threadIdx.x == 0) {
      sc->dd = dd;
      sc->dg = dg;
    }
  } else if (what == 1) {  // cost
    double c = 0;
    for (int i = threadIdx.x; i < n_cost_blocks; i += blockDim.x) c += partial[2 * i];
    c = block_sum_fixed(c, red);
    if (threadIdx.x == 0) sc->cost = c;
  } else {  // max |neg_grad| over the cameras
    double m = 0;
    for (int e = threadIdx.x; e < 6 * nb; e += blockDim.x) m = fmax(m, fabs(grad_cam[e]));
    red[threadIdx.x] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      double mm = 0;
      for (int i = 0; i < int(blockDim.x); ++i) mm = fmax(mm, red[i]);
      atomicMax(&sc->max_grad_bits, (unsigned long long)__double_as_longlong(mm));
    }
  }
}

// ---- pose graph (SE(3) between factor) --------------------------------------------------------
// PoseGraph3CostTerm (pose_graph3.rs:33-79): r = log(T_m^-1 T_n T_mn^-1),
// D = J_l^-1(log(a b)) Adj(a) with a = T_m^-1, b = T_n T_mn^-1; J_m = -D, J_n = D.
__device__ void pose_graph_eval(const double* m, const double* n, const double* mn, double r[6],
                                double D[36]) {
  double a[7], mni[7], b[7], ab[7], Jl[36], Ad[36];
  se3_inverse(m, a);
  se3_inverse(mn, mni);
  se3_mul(n, mni, b);
  double t1[7];
  se3_mul(a, n, t1);
  double t2[7];
  se3_mul(t1, mni, t2);
  se3_log(t2, r);
  se3_mul(a, b, ab);
  double t[6];
  se3_log(ab, t);
  se3_inv_left_jacobian(t, Jl);
  se3_adj(a, Ad);
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      double s = 0;
      for (int k = 0; k < 6; ++k) s += Jl[6 * i + k] * Ad[6 * k + j];
      D[6 * i + j] = s;
    }
}

// Helpers of the pose terms' linearization.  Right-multiplication of a 6x6 row-major block by
// the centring matrix G(c) = [[I, 0], [hat(c), I]]: (X G)[:, 0:3] = X[:, 0:3] + X[:, 3:6] hat(c),
// (X G)[:, 3:6] = X[:, 3:6].
__device__ __forceinline__ void pose_right_g(const double* X, const double c[3], double* Y) {
  for (int r = 0; r < 6; ++r) {
    const double* x = X + 6 * r;
    Y[6 * r + 0] = x[0] + (x[4] * c[2] - x[5] * c[1]);
    Y[6 * r + 1] = x[1] + (-x[3] * c[2] + x[5] * c[0]);
    Y[6 * r + 2] = x[2] + (x[3] * c[1] - x[4] * c[0]);
    for (int k = 3; k < 6; ++k) Y[6 * r + k] = x[k];
  }
}
// G(ca)^T X G(cb) (X row-major 6x6)
__device__ __forceinline__ void pose_sandwich(const double* X, const double ca[3], const double cb[3],
                                              double* Y) {
  double XG[36], Xt[36], T1[36];
  pose_right_g(X, cb, XG);
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) Xt[6 * j + i] = XG[6 * i + j];
  pose_right_g(Xt, ca, T1);  // (X G_b)^T G_a = (G_a^T X G_b)^T
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) Y[6 * i + j] = T1[6 * j + i];
}
// Adds element (a, b) of a row-m / column-n block into S slot `slot` (its front position, or
// the row-major slot-order blocks of the iterative solvers); `flip` when the slot holds the
// transpose.
__device__ __forceinline__ void pose_add_block(const DevProblem& p, double* C, double* bsr, int slot,
                                               const double* X, int flip) {
  if (slot < 0) return;
  if (bsr) {
    for (int e = 0; e < 36; ++e) {
      const int a = e / 6, b = e % 6;
      const double v = flip ? X[6 * b + a] : X[e];
      if (v != 0.0) atomic_add_d(bsr + 36 * int64_t(slot) + e, v);
    }
    return;
  }
  double* B = C + p.s_off[slot];
  const int ldc = p.s_ld[slot], ld = ldc >> 1;
  for (int e = 0; e < 36; ++e) {
    const int a = e / 6, b = e % 6;
    const double v = flip ? X[6 * b + a] : X[e];
    if (v != 0.0) atomic_add_d(B + ((ldc & 1) ? a * ld + b : b * ld + a), v);
  }
}
// rhs += sign G(c)^T v = sign [v_w + v_v x c; v_v], grad_cam += sign v (block cb)
__device__ __forceinline__ void pose_add_rhs(double* rhs_nat, double* grad_cam, int cb, double sign,
                                             const double* v, const double c[3]) {
  const double r[6] = {v[0] + (v[4] * c[2] - v[5] * c[1]), v[1] + (v[5] * c[0] - v[3] * c[2]),
                       v[2] + (v[3] * c[1] - v[4] * c[0]), v[3], v[4], v[5]};
  for (int k = 0; k < 6; ++k) {
    atomic_add_d(rhs_nat + 6 * cb + k, sign * r[k]);
    atomic_add_d(grad_cam + 6 * cb + k, sign * v[k]);
  }
}

// With `centered` (cw != nullptr), the blocks and rhs are those of the centred camera
// tangents: G_m^T H_mn G_n and G_m^T b_m (see k_schur_slots).
__global__ void k_pose_graph_linearize(DevPoseGraph g, DevProblem p, const double* __restrict__ poses,
                                       double nu, double* __restrict__ C, double* __restrict__ rhs_nat,
                                       double* __restrict__ grad_cam, const double* __restrict__ cw,
                                       double* __restrict__ bsr) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.n) return;
  double r[6], D[36];
  pose_graph_eval(poses + 7 * g.m[t], poses + 7 * g.nn[t], g.meas + 7 * t, r, D);
  double nrm = 0;
  for (int k = 0; k < 6; ++k) nrm += r[k] * r[k];
  const double w = robust_weight(g.robust, g.delta, sqrt(nrm));
  double rw[6];
  for (int k = 0; k < 6; ++k) rw[k] = g.robust == 1 ? w * r[k] : r[k];
  double DtD[36], Dtr[6];
  for (int i = 0; i < 6; ++i) {
    double s = 0;
    for (int k = 0; k < 6; ++k) s += D[6 * k + i] * rw[k];
    Dtr[i] = s;
    for (int j = 0; j < 6; ++j) {
      double q = 0;
      for (int k = 0; k < 6; ++k) q += D[6 * k + i] * D[6 * k + j];
      DtD[6 * i + j] = q;
    }
  }
  const int cm = p.pose_block[g.m[t]], cn = p.pose_block[g.nn[t]];
  const int* dst = g.dest + 4 * t;
  const double damp = g.first[t] ? nu : 0.0;
  // camera centres (zero when not centred: G = I)
  double cmc[3] = {0, 0, 0}, cnc[3] = {0, 0, 0};
  if (cw) {
    cam_center(cw + 12 * g.m[t], cmc);
    cam_center(cw + 12 * g.nn[t], cnc);
  }
  // H_mm = H_nn = D^T D (+ nu I), H_mn = J_m^T J_n = -D^T D (symmetric, so the stored
  // orientation of the (max, min) block does not matter)
  double Hmm[36], Hmn[36], X[36];
  for (int e = 0; e < 36; ++e) {
    Hmm[e] = DtD[e] + ((e / 6 == e % 6) ? damp : 0.0);
    Hmn[e] = -DtD[e];
  }
  pose_sandwich(Hmm, cmc, cmc, X);
  pose_add_block(p, C, bsr, dst[0], X, 0);
  pose_sandwich(Hmm, cnc, cnc, X);
  pose_add_block(p, C, bsr, dst[1], X, 0);
  pose_sandwich(Hmn, cmc, cnc, X);
  pose_add_block(p, C, bsr, dst[2], X, dst[3]);
  // neg_grad_m -= J_m^T r~ = +D^T r~ ; neg_grad_n -= D^T r~
  if (cm >= 0) pose_add_rhs(rhs_nat, grad_cam, cm, 1.0, Dtr, cmc);
  if (cn >= 0) pose_add_rhs(rhs_nat, grad_cam, cn, -1.0, Dtr, cnc);
}

// Isometry3PriorCostTerm (isometry3_prior.rs:46-88): r = log(T E^-1) and its Jacobian by
// forward-mode duals of log(exp(x) T E^-1) at x = 0 (the reference's own derivative, which
// also fixes the small-angle behaviour).
__device__ void prior_eval(const double* pose, const double* mean, double r[6], double J[36]) {
  using D6 = dual::Dual<6>;
  D6 x[6], ex[7], T[7], Em[7], Ei[7], Tx[7], TE[7], gv[6];
  for (int k = 0; k < 6; ++k) x[k] = D6::var(0.0, k);
  for (int k = 0; k < 7; ++k) {
    T[k] = D6(pose[k]);
    Em[k] = D6(mean[k]);
  }
  se3_inverse(Em, Ei);
  se3_exp(x, ex);
  se3_mul(ex, T, Tx);
  se3_mul(Tx, Ei, TE);
  se3_log(TE, gv);
  for (int i = 0; i < 6; ++i) {
    r[i] = gv[i].re;
    for (int k = 0; k < 6; ++k) J[6 * i + k] = gv[i].d[k];
  }
}

// r~ = w(|r|) r and, with the precision W, lambda = W r~ and the cost r~^T lambda (make,
// evaluated_term.rs:286-327)
__device__ __forceinline__ double prior_weighted(const DevPoseGraph& g, int t, const double r[6],
                                                 double rw[6], double lam[6]) {
  double nrm = 0;
  for (int k = 0; k < 6; ++k) nrm += r[k] * r[k];
  const double w = robust_weight(g.robust, g.delta, sqrt(nrm));
  for (int k = 0; k < 6; ++k) rw[k] = g.robust == 1 ? w * r[k] : r[k];
  const double* W = g.prec + 36 * int64_t(t);
  double c = 0;
  for (int i = 0; i < 6; ++i) {
    double s = 0;
    for (int k = 0; k < 6; ++k) s += W[6 * i + k] * rw[k];
    lam[i] = s;
    c += rw[i] * s;
  }
  return c;
}

// One thread per prior term: H = J^T W J (+ nu I on the term's first of a reduction range)
// into the pose's diagonal S slot, neg_grad -= J^T W r~.
__global__ void k_prior_linearize(DevPoseGraph g, DevProblem p, const double* __restrict__ poses,
                                  double nu, double* __restrict__ C, double* __restrict__ rhs_nat,
                                  double* __restrict__ grad_cam, const double* __restrict__ cw,
                                  double* __restrict__ bsr) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.n) return;
  const int cm = p.pose_block[g.m[t]];
  if (cm < 0) return;  // a constant pose: its prior adds cost only
  double r[6], J[36], rw[6], lam[6];
  prior_eval(poses + 7 * g.m[t], g.meas + 7 * t, r, J);
  prior_weighted(g, t, r, rw, lam);
  const double* W = g.prec + 36 * int64_t(t);
  double WJ[36], H[36], Jtl[6];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      double s = 0;
      for (int k = 0; k < 6; ++k) s += W[6 * i + k] * J[6 * k + j];
      WJ[6 * i + j] = s;
    }
  const double damp = g.first[t] ? nu : 0.0;
  for (int i = 0; i < 6; ++i) {
    double s = 0;
    for (int k = 0; k < 6; ++k) s += J[6 * k + i] * lam[k];
    Jtl[i] = s;
    for (int j = 0; j < 6; ++j) {
      double q = 0;
      for (int k = 0; k < 6; ++k) q += J[6 * k + i] * WJ[6 * k + j];
      H[6 * i + j] = q + (i == j ? damp : 0.0);
    }
  }
  double c[3] = {0, 0, 0}, X[36];
  if (cw) cam_center(cw + 12 * g.m[t], c);
  pose_sandwich(H, c, c, X);
  pose_add_block(p, C, bsr, g.dest[4 * t], X, 0);
  pose_add_rhs(rhs_nat, grad_cam, cm, -1.0, Jtl, c);
}

__global__ void k_prior_cost(DevPoseGraph g, const double* __restrict__ poses,
                             double* __restrict__ partial) {
  double acc = 0;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < g.n; t += gridDim.x * blockDim.x) {
    double ei[7], te[7], r[6], rw[6], lam[6];
    se3_inverse(g.meas + 7 * t, ei);
    se3_mul(poses + 7 * g.m[t], ei, te);
    se3_log(te, r);
    acc += prior_weighted(g, t, r, rw, lam);
  }
  block_reduce2(acc, 0.0, partial);
}

// delta = G delta': nu += c x omega for every free camera (centred solves).
__global__ void k_uncenter(int nb, const int32_t* __restrict__ cam_member,
                           const double* __restrict__ cw, double* __restrict__ dx) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nb) return;
  double ctr[3];
  cam_center(cw + 12 * cam_member[c], ctr);
  double* d = dx + 6 * c;
  const double w0 = d[0], w1 = d[1], w2 = d[2];
  d[3] += ctr[1] * w2 - ctr[2] * w1;
  d[4] += ctr[2] * w0 - ctr[0] * w2;
  d[5] += ctr[0] * w1 - ctr[1] * w0;
}

__global__ void k_pose_graph_cost(DevPoseGraph g, const double* __restrict__ poses,
                                  double* __restrict__ partial) {
  double acc = 0;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < g.n; t += gridDim.x * blockDim.x) {
    double mi[7], mni[7], t1[7], t2[7], r[6];
    se3_inverse(poses + 7 * g.m[t], mi);
    se3_inverse(g.meas + 7 * t, mni);
    se3_mul(mi, poses + 7 * g.nn[t], t1);
    se3_mul(t1, mni, t2);
    se3_log(t2, r);
    double nrm = 0;
    for (int k = 0; k < 6; ++k) nrm += r[k] * r[k];
    if (g.robust == 1) {
      const double w = robust_weight(1, g.delta, sqrt(nrm));
      double c = 0;
      for (int k = 0; k < 6; ++k) c += (w * r[k]) * (w * r[k]);
      acc += c;
    } else {
      acc += nrm;
    }
  }
  block_reduce2(acc, 0.0, partial);
}

// ---- stage-level evaluation ----------------------------------------------------------------
template <typename T>
__global__ void k_eval_pinhole(int64_t n, const double* __restrict__ intr,
                               const double* __restrict__ pose, const double* __restrict__ pt,
                               const double* __restrict__ uv, int robust, double delta,
                               double* res, double* jpose, double* jpt, double* cost,
                               double* weight) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= n) return;
  double inv[7], R[9], cw[12];
  se3_inverse(pose + 7 * t, inv);
  so3_matrix(inv, R);
  for (int k = 0; k < 9; ++k) cw[k] = R[k];
  cw[9] = inv[4];
  cw[10] = inv[5];
  cw[11] = inv[6];
  const double x[3] = {pt[3 * t], pt[3 * t + 1], pt[3 * t + 2]};
  double xc[3];
  to_camera(cw, x, xc);
  const double* in = intr + 4 * t;
  T A[6];
  obs_jacobian<T>(cw, in, xc, A);
  const double px = xc[0] / xc[2], py = xc[1] / xc[2];
  const double r0 = uv[2 * t] - (px * in[0] + in[2]);
  const double r1 = uv[2 * t + 1] - (py * in[1] + in[3]);
  const double w = robust_weight(robust, delta, sqrt(r0 * r0 + r1 * r1));
  if (res) {
    res[2 * t] = r0;
    res[2 * t + 1] = r1;
  }
  if (jpose) {
    const T nx[3] = {-T(x[0]), -T(x[1]), -T(x[2])};
    for (int i = 0; i < 2; ++i) {
      const T* a = A + 3 * i;
      // A * hat(-x): hat(nx) rows [0,-nx2,nx1],[nx2,0,-nx0],[-nx1,nx0,0]
      jpose[12 * t + 6 * i + 0] = double(a[1] * nx[2] - a[2] * nx[1]);
      jpose[12 * t + 6 * i + 1] = double(-a[0] * nx[2] + a[2] * nx[0]);
      jpose[12 * t + 6 * i + 2] = double(a[0] * nx[1] - a[1] * nx[0]);
      for (int j = 0; j < 3; ++j) jpose[12 * t + 6 * i + 3 + j] = double(a[j]);
    }
  }
  if (jpt)
    for (int k = 0; k < 6; ++k) jpt[6 * t + k] = -double(A[k]);
  const double a0 = robust == 1 ? w * r0 : r0, a1 = robust == 1 ? w * r1 : r1;
  if (cost) cost[t] = a0 * a0 + a1 * a1;
  if (weight) weight[t] = w;
}

__global__ void k_eval_eucm(int64_t n, const double* __restrict__ intr,
                            const double* __restrict__ pose, const double* __restrict__ pt,
                            const double* __restrict__ uv, int robust, double delta, double* res,
                            double* jintr, double* jpose, double* jpt, double* cost,
                            double* weight) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= n) return;
  double inv[7], R[9], cw[12];
  se3_inverse(pose + 7 * t, inv);
  so3_matrix(inv, R);
  for (int k = 0; k < 9; ++k) cw[k] = R[k];
  cw[9] = inv[4];
  cw[10] = inv[5];
  cw[11] = inv[6];
  const double x[3] = {pt[3 * t], pt[3 * t + 1], pt[3 * t + 2]};
  double xc[3], px[2], D[6], E[12], A[6];
  to_camera(cw, x, xc);
  eucm_project(intr + 6 * t, xc, px, D, E);
  lift_rotation(cw, D, A);
  const double r0 = uv[2 * t] - px[0], r1 = uv[2 * t + 1] - px[1];
  const double w = robust_weight(robust, delta, sqrt(r0 * r0 + r1 * r1));
  if (res) {
    res[2 * t] = r0;
    res[2 * t + 1] = r1;
  }
  if (jintr)
    for (int k = 0; k < 12; ++k) jintr[12 * t + k] = -E[k];
  if (jpose) {
    for (int i = 0; i < 2; ++i) {
      const double* a = A + 3 * i;
      jpose[12 * t + 6 * i + 0] = a[1] * (-x[2]) - a[2] * (-x[1]);
      jpose[12 * t + 6 * i + 1] = -a[0] * (-x[2]) + a[2] * (-x[0]);
      jpose[12 * t + 6 * i + 2] = a[0] * (-x[1]) - a[1] * (-x[0]);
      for (int j = 0; j < 3; ++j) jpose[12 * t + 6 * i + 3 + j] = a[j];
    }
  }
  if (jpt)
    for (int k = 0; k < 6; ++k) jpt[6 * t + k] = -A[k];
  const double a0 = robust == 1 ? w * r0 : r0, a1 = robust == 1 ? w * r1 : r1;
  if (cost) cost[t] = a0 * a0 + a1 * a1;
  if (weight) weight[t] = w;
}

__global__ void k_eval_pose_graph(int64_t n, const double* __restrict__ m,
                                  const double* __restrict__ nn, const double* __restrict__ mn,
                                  double* res, double* jm, double* jn) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= n) return;
  double r[6], D[36];
  pose_graph_eval(m + 7 * t, nn + 7 * t, mn + 7 * t, r, D);
  for (int k = 0; k < 6; ++k) res[6 * t + k] = r[k];
  for (int k = 0; k < 36; ++k) {
    jm[36 * t + k] = -D[k];
    jn[36 * t + k] = D[k];
  }
}

int blocks_for(int64_t n, int threads) { return int((n + threads - 1) / threads); }

}  // namespace

// ---- launchers ---------------------------------------------------------------------------------
void launch_cam_cache(hipStream_t s, int npose, const double* poses, double* cw) {
  if (npose == 0) return;
  hipLaunchKernelGGL(k_cam_cache, dim3(blocks_for(npose, kThreads)), dim3(kThreads), 0, s, npose,
                     poses, cw);
}

void launch_point_prep(hipStream_t s, int precision, const DevProblem& p, const double* poses,
                       const double* points, const double* intr, const double* cw, double nu,
                       DevLin& w, DevScalars* sc) {
  (void)poses;
  if (p.npt == 0) return;
  const dim3 g(blocks_for(p.npt, kThreads)), b(kThreads);
  const bool eucm = p.cam_model == kEucm;
  if (precision == 0)
    hipLaunchKernelGGL((eucm ? k_point_prep<double, kEucm> : k_point_prep<double, kPinhole>), g, b,
                       0, s, p, points, intr, cw, nu, w.pt_s, w.prec, sc);
  else
    hipLaunchKernelGGL((eucm ? k_point_prep<float, kEucm> : k_point_prep<float, kPinhole>), g, b, 0,
                       s, p, points, intr, cw, nu, w.pt_s, w.prec, sc);
}

void launch_camera_rows(hipStream_t s, int precision, const DevProblem& p, const double* points,
                        const double* intr, const double* cw, double nu, DevLin& w, DevFactor& f,
                        DevScalars* sc) {
  (void)sc;
  if (p.nslot == 0) return;
  const int wpb = kThreads / 64;
  const dim3 go(unsigned(p.n_own_chunks)), gs(unsigned((p.nslot + wpb - 1) / wpb));
  const bool eucm = p.cam_model == kEucm;
  if (precision == 0) {
    double* v = static_cast<double*>(w.vobs);
    if (p.n_own_chunks > 0)
      hipLaunchKernelGGL((eucm ? k_cam_obs<double, kEucm> : k_cam_obs<double, kPinhole>), go,
                         dim3(kThreads), 0, s, p, points, intr, cw, w.prec, 0, v, w.own_part);
    hipLaunchKernelGGL((k_schur_blocks<double, false>), gs, dim3(kThreads), 0, s, p, v, w.own_part, cw, nu,
                       f.front, w.rhs_nat, w.grad_cam, w.bsr);
  } else {
    float* v = static_cast<float*>(w.vobs);
    if (p.n_own_chunks > 0)
      hipLaunchKernelGGL((eucm ? k_cam_obs<float, kEucm> : k_cam_obs<float, kPinhole>), go,
                         dim3(kThreads), 0, s, p, points, intr, cw, w.prec, 1, v, w.own_part);
    hipLaunchKernelGGL((k_schur_blocks<float, true>), gs, dim3(kThreads), 0, s, p, v, w.own_part, cw, nu,
                       f.front, w.rhs_nat, w.grad_cam, w.bsr);
  }
}

// Zero the lower triangle (diagonal included) of every front: the assembly, extend-add and
// factorization read and write only that part.  Workgroup (node, g) takes columns 4 g ..
// 4 g + 3 of the node's front, one wavefront per column.
__global__ void __launch_bounds__(kThreads) k_zero_fronts(DevFactor f) {
  const int X = blockIdx.x, F = 6 * f.sn_nf[X];
  const int c = 4 * blockIdx.y + (threadIdx.x >> 6);
  if (c >= F) return;
  double* col = f.front + f.sn_off[X] + int64_t(c) * F;
  for (int r = c + (threadIdx.x & 63); r < F; r += 64) col[r] = 0.0;
}

void launch_zero_fronts(hipStream_t s, const DevFactor& f) {
  const Symbolic& S = *f.host;
  hipLaunchKernelGGL(k_zero_fronts, dim3(S.nsn, (6 * S.max_front + 3) / 4), dim3(kThreads), 0, s, f);
}

void launch_fill(hipStream_t s, double* ptr, int64_t n, double v) {
  if (n == 0) return;
  const int g = int(std::min<int64_t>((n + kThreads - 1) / kThreads, 4096));
  hipLaunchKernelGGL(k_fill, dim3(g), dim3(kThreads), 0, s, ptr, n, v);
}

void launch_zero_scalars(hipStream_t s, DevScalars* sc) {
  hipLaunchKernelGGL(k_zero_scalars, dim3(1), dim3(64), 0, s, sc);
}

void launch_rcs_factor(hipStream_t s, const DevFactor& f, DevScalars* sc, const double* rhs_nat,
                       double* y) {
  const Symbolic& S = *f.host;
  // the right-hand side in elimination order, then into each node's front rows
  const int g = blocks_for(6 * int64_t(f.nb), kThreads);
  hipLaunchKernelGGL(k_permute_in, dim3(g), dim3(kThreads), 0, s, f.nb, f.perm, rhs_nat, y);
  hipLaunchKernelGGL(k_mf_rhs_init, dim3(S.nsn), dim3(kThreads), 0, s, f, y);
  // the leaves' first diagonal blocks; every other node's is factored in a round
  if (!S.leaves.empty())
    hipLaunchKernelGGL(k_mf_diag0, dim3(S.leaves.size()), dim3(kThreads), 0, s, f, f.leaves, sc, y);
  for (size_t li = 0; li < S.launches.size(); ++li) {
    const auto& L = S.launches[li];
    if (L.tile_count == 0) continue;
    const int32_t* items = f.launch_tiles + 4 * int64_t(L.tile_begin);
    if (f.step_events) SGPU_HIP(hipEventRecord(f.step_events[2 * li], s));
    if (L.fused) {
      hipLaunchKernelGGL(k_mf_step, dim3(L.tile_count), dim3(kThreads), 0, s, f, items, sc, y);
    } else {
      hipLaunchKernelGGL(k_mf_panel, dim3(L.item_count), dim3(kThreads), 0, s, f,
                         f.panel_items + 3 * int64_t(L.item_begin), y);
      hipLaunchKernelGGL(k_mf_update, dim3(L.tile_count), dim3(kThreads), 0, s, f, items, sc, y);
    }
    if (f.step_events) SGPU_HIP(hipEventRecord(f.step_events[2 * li + 1], s));
  }
}

// whole wavefronts covering the level's largest node (one thread per node row)
static int tri_threads(const Symbolic& S, int level) {
  const int K = 6 * S.lvl_max_nc[level];
  if (K > kTriMax) throw std::runtime_error("supernode larger than the triangular-solve block");
  return std::max(64, (K + 63) / 64 * 64);
}

void launch_rcs_solve(hipStream_t s, const DevFactor& f, double* y, double* dx_cam) {
  const Symbolic& S = *f.host;
  const int g = blocks_for(6 * int64_t(f.nb), kThreads);
  const int nlev = int(S.lvl_ptr.size()) - 1;
  for (int l = nlev - 1; l >= 0; --l) {
    const int n = S.lvl_ptr[l + 1] - S.lvl_ptr[l];
    const int nc = S.solve_cols_ptr[l + 1] - S.solve_cols_ptr[l];
    if (nc > 0)
      hipLaunchKernelGGL(k_mf_bwd_rect, dim3(nc), dim3(kThreads), 0, s, f,
                         f.solve_cols + 2 * int64_t(S.solve_cols_ptr[l]), y);
    hipLaunchKernelGGL(k_mf_bwd_tri, dim3(n), dim3(tri_threads(S, l)), 0, s, f,
                       f.lvl_sn + S.lvl_ptr[l], y);
  }
  if (f.scalar_pivots)
    hipLaunchKernelGGL(k_permute_out_scalar_rule, dim3(g), dim3(kThreads), 0, s, f.nb, f.perm, f.piv_d,
                       y, dx_cam);
  else
    hipLaunchKernelGGL(k_permute_out, dim3(g), dim3(kThreads), 0, s, f.nb, f.perm, y, dx_cam);
}

void launch_rcs_forward(hipStream_t s, const DevFactor& f, const double* b_nat, double* b_perm,
                        double* z) {
  const Symbolic& S = *f.host;
  if (!f.sn_child_ptr) throw Error(SGPU_ERR_UNSUPPORTED, "forward solve without child lists");
  const int g = blocks_for(6 * int64_t(f.nb), kThreads);
  hipLaunchKernelGGL(k_permute_in, dim3(g), dim3(kThreads), 0, s, f.nb, f.perm, b_nat, b_perm);
  hipLaunchKernelGGL(k_mf_rhs_init, dim3(S.nsn), dim3(kThreads), 0, s, f, b_perm);
  const int nlev = int(S.lvl_ptr.size()) - 1;
  for (int l = 0; l < nlev; ++l) {
    const int n = S.lvl_ptr[l + 1] - S.lvl_ptr[l];
    hipLaunchKernelGGL(k_mf_fwd, dim3(n), dim3(kThreads), 0, s, f, f.lvl_sn + S.lvl_ptr[l], z);
  }
}

void launch_point_coupling(hipStream_t s, const DevProblem& p, const double* points,
                           const double* intr, const double* cw, const double* prec, int b,
                           double* out, int32_t* cam, double* hinv) {
  hipLaunchKernelGGL((p.cam_model == kEucm ? k_point_coupling<kEucm> : k_point_coupling<kPinhole>),
                     dim3(1), dim3(64), 0, s, p, points, intr, cw, prec, b, out, cam, hinv);
}

// Lanes per point of k_intr_points from the mean track length.
static int intr_group(const DevProblem& p) {
  const int64_t mean = p.npt > 0 ? int64_t(p.nobs) / p.npt : 1;
  return mean >= 24 ? 16 : mean >= 6 ? 4 : 1;
}

// at most 256 workgroups, each striding over groups of points (fewer partials to sum)
static int intr_point_blocks(const DevProblem& p) {
  return std::max(1, std::min(blocks_for(int64_t(p.npt) * intr_group(p), kThreads), 256));
}

void launch_intrinsics_rows(hipStream_t s, const DevProblem& p, const double* points,
                            const double* intr, const double* cw, DevLin& w, DevFactor& f) {
  if (p.intr_block < 0) return;
  const int gpt = intr_point_blocks(p);
  switch (intr_group(p)) {
#define SGPU_INTR_POINTS(G)                                                                   \
  case G:                                                                                     \
    hipLaunchKernelGGL(k_intr_points<G>, dim3(gpt), dim3(kThreads), 0, s, p, points, intr, cw, \
                       w.prec, w.pt_s, w.pt_vw, w.intr_part);                                 \
    break;
    SGPU_INTR_POINTS(1) SGPU_INTR_POINTS(4) SGPU_INTR_POINTS(16)
#undef SGPU_INTR_POINTS
  }
  if (p.n_own_chunks > 0) {
    hipLaunchKernelGGL(k_intr_cam_chunks, dim3(p.n_own_chunks), dim3(64), 0, s, p, points, intr,
                       cw, w.pt_vw, w.intr_cpart);
    hipLaunchKernelGGL(k_intr_cam_store, dim3(blocks_for(int64_t(p.intr_block) * 64, kThreads)),
                       dim3(kThreads), 0, s, p, p.intr_block, w.intr_cpart, f.front, w.bsr);
  }
  hipLaunchKernelGGL(k_intr_finish, dim3(1), dim3(kThreads), 0, s, p, w.intr_part, gpt, f.front,
                     w.bsr, w.rhs_nat, w.grad_cam);
}

int intrinsics_partials(const DevProblem& p) { return kIntrPart * intr_point_blocks(p); }

// Diagonal of the assembled S (camera order) from the diagonal slots of the fronts: the
// diagonal block of row c is its last slot (columns ascending, c the largest).
__global__ void k_gather_s_diag(DevProblem p, const double* __restrict__ front, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 6 * p.nb) return;
  const int c = t / 6, a = t % 6;
  const int k = p.s_ptr[c + 1] - 1;
  const int ld = p.s_ld[k] >> 1;
  out[t] = front[p.s_off[k] + int64_t(a) * ld + a];
}

void launch_gather_s_diag(hipStream_t s, const DevProblem& p, const double* front, double* out) {
  if (p.nb == 0) return;
  hipLaunchKernelGGL(k_gather_s_diag, dim3(blocks_for(6 * int64_t(p.nb), kThreads)), dim3(kThreads), 0, s,
                     p, front, out);
}

void launch_scatter_s(hipStream_t s, const DevProblem& p, int64_t nslots, const double* blocks,
                      DevFactor& f) {
  if (nslots == 0) return;
  hipLaunchKernelGGL(k_scatter_s, dim3(blocks_for(nslots * 36, kThreads)), dim3(kThreads), 0, s, p,
                     nslots, blocks, f.front);
}

void launch_backsub_points(hipStream_t s, int precision, const DevProblem& p, const double* points,
                           const double* intr, const double* cw, const DevLin& w,
                           double* points_new, DevScalars* sc) {
  const int g = std::max(1, std::min(blocks_for(p.npt, kThreads), kCostBlocks));
  auto kernel = k_backsub<double, kPinhole, false>;
  if (precision == 0 && p.intr_block >= 0)
    kernel = k_backsub<double, kEucm, true>;  // free intrinsics: FP64 EUCM only
  else if (precision == 0)
    kernel = p.cam_model == kEucm ? k_backsub<double, kEucm, false> : k_backsub<double, kPinhole, false>;
  if (precision == 0) {
    hipLaunchKernelGGL(kernel, dim3(g), dim3(kThreads), 0, s, p, points, intr, cw, w.prec, w.pt_s,
                       w.dx_cam, w.pt_vw, points_new, w.partial, sc);
  } else {
    auto kf = p.cam_model == kEucm ? k_backsub<float, kEucm, false> : k_backsub<float, kPinhole, false>;
    hipLaunchKernelGGL(kf, dim3(g), dim3(kThreads), 0, s, p, points, intr, cw, w.prec, w.pt_s,
                       w.dx_cam, w.pt_vw, points_new, w.partial, sc);
  }
}

// Partials of the step sums: 2 per back-substitution workgroup, 2 per pose-update workgroup,
// then 2 for the free intrinsics' block and 2 for the equality constraints' multipliers.
static int pose_update_blocks(const DevProblem& p) {
  const int npb = p.intr_block >= 0 ? p.intr_block : p.nb;
  return std::max(1, std::min(blocks_for(npb, kThreads), kCostBlocks));
}

void launch_update_poses(hipStream_t s, const DevProblem& p, const double* poses,
                         const double* intr, const DevLin& w, double* poses_new, double* intr_new,
                         DevScalars* sc) {
  const int gpt = std::max(1, std::min(blocks_for(p.npt, kThreads), kCostBlocks));
  const int g = pose_update_blocks(p);
  const int npb = p.intr_block >= 0 ? p.intr_block : p.nb;
  hipLaunchKernelGGL(k_update_poses, dim3(g), dim3(kThreads), 0, s, p, npb, poses, w.dx_cam,
                     w.grad_cam, poses_new, w.partial + 2 * gpt, sc);
  if (p.intr_block >= 0)
    hipLaunchKernelGGL(k_update_intrinsics, dim3(1), dim3(64), 0, s, p, w.dx_cam, w.grad_cam, intr,
                       intr_new, w.partial + 2 * gpt + 2 * g, sc);
}

void launch_eq_eval(hipStream_t s, const DevProblem& p, const DevEq& q, const double* poses,
                    const double* lambda, DevLin& w, bool add_rhs) {
  if (q.nset == 0) return;
  hipLaunchKernelGGL(k_eq_eval, dim3(1), dim3(64), 0, s, p, q, poses, lambda, w.rhs_nat, w.grad_cam,
                     add_rhs ? 1 : 0);
}

void launch_eq_solve(hipStream_t s, const DevProblem& p, const DevEq& q, const DevFactor& f,
                     const double* lambda, double* lambda_next, DevLin& w, DevScalars* sc) {
  if (q.nset == 0) return;
  const int64_t nf = 6 * int64_t(p.nb);
  for (int k = 0; k < q.nset; ++k) {
    hipLaunchKernelGGL(k_eq_column, dim3(1), dim3(kThreads), 0, s, p, q, k);
    launch_rcs_forward(s, f, q.b, q.bperm, q.z);
    launch_rcs_solve(s, f, q.z, q.Y + nf * k);
  }
  const int gpt = std::max(1, std::min(blocks_for(p.npt, kThreads), kCostBlocks));
  double* part = w.partial + 2 * gpt + 2 * (pose_update_blocks(p) + (p.intr_block >= 0 ? 1 : 0));
  hipLaunchKernelGGL(k_eq_kkt, dim3(1), dim3(kThreads), 0, s, p, q, lambda, lambda_next, w.dx_cam,
                     part, sc);
}

void launch_finalize_step(hipStream_t s, const DevProblem& p, const DevLin& w, DevScalars* sc) {
  const int gpt = std::max(1, std::min(blocks_for(p.npt, kThreads), kCostBlocks));
  const int gpo = pose_update_blocks(p) + (p.intr_block >= 0 ? 1 : 0) + (p.n_eq > 0 ? 1 : 0);
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, w.partial, gpt, gpo, 0, nullptr, 0, sc,
                     0);
}

void launch_cost(hipStream_t s, const DevProblem& p, const double* poses, const double* points,
                 const double* intr, double* partial, int32_t partial_cap, DevScalars* sc) {
  (void)partial_cap;
  hipLaunchKernelGGL((p.cam_model == kEucm ? k_cost<kEucm> : k_cost<kPinhole>), dim3(kCostBlocks),
                     dim3(kThreads), 0, s, p, p.obs_point, poses, points, intr, partial);
  (void)sc;
}

void launch_finalize_cost(hipStream_t s, const double* partial, int nblocks, DevScalars* sc) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, partial, 0, 0, nblocks, nullptr, 0, sc,
                     1);
}

void launch_max_grad(hipStream_t s, const double* grad_cam, int nb, DevScalars* sc) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, nullptr, 0, 0, 0, grad_cam, nb, sc, 2);
}

void launch_pose_graph_linearize(hipStream_t s, const DevPoseGraph& g, const DevProblem& p,
                                 const double* poses, double nu, DevLin& w, DevFactor& f,
                                 const double* centered_cw) {
  if (g.n == 0) return;
  if (g.prior)
    hipLaunchKernelGGL(k_prior_linearize, dim3(blocks_for(g.n, 64)), dim3(64), 0, s, g, p, poses, nu,
                       f.front, w.rhs_nat, w.grad_cam, centered_cw, w.bsr);
  else
    hipLaunchKernelGGL(k_pose_graph_linearize, dim3(blocks_for(g.n, 128)), dim3(128), 0, s, g, p,
                       poses, nu, f.front, w.rhs_nat, w.grad_cam, centered_cw, w.bsr);
}

void launch_uncenter(hipStream_t s, const DevProblem& p, const double* cw, double* dx_cam) {
  if (p.nb == 0) return;
  hipLaunchKernelGGL(k_uncenter, dim3(blocks_for(p.nb, kThreads)), dim3(kThreads), 0, s, p.nb,
                     p.cam_member, cw, dx_cam);
}

void launch_pose_graph_cost(hipStream_t s, const DevPoseGraph& g, const double* poses,
                            double* partial, int32_t partial_cap, DevScalars* sc) {
  (void)partial_cap;
  (void)sc;
  if (g.prior)
    hipLaunchKernelGGL(k_prior_cost, dim3(kCostBlocks), dim3(kThreads), 0, s, g, poses, partial);
  else
    hipLaunchKernelGGL(k_pose_graph_cost, dim3(kCostBlocks), dim3(kThreads), 0, s, g, poses, partial);
}

void launch_eval_pinhole(hipStream_t s, int precision, int64_t n, const double* intr,
                         const double* pose, const double* pt, const double* uv, int robust,
                         double delta, double* res, double* jpose, double* jpt, double* cost,
                         double* weight) {
  if (n == 0) return;
  const int g = blocks_for(n, kThreads);
  if (precision == 0)
    hipLaunchKernelGGL(k_eval_pinhole<double>, dim3(g), dim3(kThreads), 0, s, n, intr, pose, pt, uv,
                       robust, delta, res, jpose, jpt, cost, weight);
  else
    hipLaunchKernelGGL(k_eval_pinhole<float>, dim3(g), dim3(kThreads), 0, s, n, intr, pose, pt, uv,
                       robust, delta, res, jpose, jpt, cost, weight);
}

void launch_eval_eucm(hipStream_t s, int64_t n, const double* intr, const double* pose,
                      const double* pt, const double* uv, int robust, double delta, double* res,
                      double* jintr, double* jpose, double* jpt, double* cost, double* weight) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_eval_eucm, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, s, n, intr,
                     pose, pt, uv, robust, delta, res, jintr, jpose, jpt, cost, weight);
}

// Isometry3 prior terms on host-given inputs: residual, dual Jacobian and cost per term.
__global__ void k_eval_prior(int64_t n, const double* __restrict__ poses, const double* __restrict__ means,
                             const double* __restrict__ prec, int robust, double delta, double* res,
                             double* jac, double* cost) {
  const int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= n) return;
  double r[6], J[36], rw[6], lam[6];
  prior_eval(poses + 7 * t, means + 7 * t, r, J);
  DevPoseGraph g{};
  g.robust = robust;
  g.delta = delta;
  g.prec = prec;
  cost[t] = prior_weighted(g, int(t), r, rw, lam);
  for (int k = 0; k < 6; ++k) res[6 * t + k] = r[k];
  for (int k = 0; k < 36; ++k) jac[36 * t + k] = J[k];
}

void launch_eval_prior(hipStream_t s, int64_t n, const double* poses, const double* means,
                       const double* prec, int robust, double delta, double* res, double* jac,
                       double* cost) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_eval_prior, dim3(blocks_for(n, 64)), dim3(64), 0, s, n, poses, means, prec,
                     robust, delta, res, jac, cost);
}

void launch_eval_pose_graph(hipStream_t s, int64_t n, const double* m, const double* nn,
                            const double* mn, double* res, double* jm, double* jn) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_eval_pose_graph, dim3(blocks_for(n, 128)), dim3(128), 0, s, n, m, nn, mn,
                     res, jm, jn);
}

}  // namespace sgpu
