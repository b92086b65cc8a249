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

#include "kernels.hpp"
#include "lie.hpp"

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

__device__ __forceinline__ void atomic_max_abs(unsigned long long* p, double v) {
  atomicMax(p, (unsigned long long)__double_as_longlong(fabs(v)));
}

// Huber weight sqrt(rho(|r|)) / |r| (robust_kernel.rs:13-35); 1 without a kernel.
__device__ __forceinline__ double robust_weight(int robust, double delta, double nrm) {
  if (robust != 1) return 1.0;
  if (nrm == 0.0) return 0.0;
  const double a = fabs(nrm) <= delta ? nrm * nrm : 2.0 * delta * fabs(nrm) - delta * delta;
  return sqrt(a) / nrm;
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
  xc[0] = cw[0] * x[0] + cw[1] * x[1] + cw[2] * x[2] + cw[9];
  xc[1] = cw[3] * x[0] + cw[4] * x[1] + cw[5] * x[2] + cw[10];
  xc[2] = cw[6] * x[0] + cw[7] * x[1] + cw[8] * x[2] + cw[11];
}

// A = diag(fx, fy) * dx_proj_x(x_c) * R_cw  (perspective.rs:54-72, affine.rs:77-85).
template <typename T>
__device__ __forceinline__ void obs_jacobian(const double* __restrict__ cw,
                                             const double* __restrict__ in, const double xc[3],
                                             T A[6]) {
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

// ---- kernel: per-point H_ll, its inverse and the point gradient --------------------------
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_point_prep(DevProblem p, const double* __restrict__ points, const double* __restrict__ intr,
                 const double* __restrict__ cw, double nu, double* __restrict__ hinv_out,
                 double* __restrict__ s_out, double* __restrict__ y_out, DevScalars* sc) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= p.npt) return;
  if (p.pt_block[b] < 0) return;
  const double x[3] = {points[3 * b], points[3 * b + 1], points[3 * b + 2]};
  T H[9] = {T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0)};
  T s[3] = {T(0), T(0), T(0)};
  for (int o = p.pt_ptr[b]; o < p.pt_ptr[b + 1]; ++o) {
    const double* c = cw + 12 * p.obs_pose[o];
    const double* in = intr + 4 * p.obs_intr[o];
    double xc[3];
    to_camera(c, x, xc);
    T A[6], v[3], M[9];
    obs_jacobian<T>(c, in, xc, A);
    obs_residual<T>(in, p.obs_uv + 2 * o, xc, p.robust, p.delta, A, v);
    gram(A, M);
#pragma unroll
    for (int k = 0; k < 9; ++k) H[k] += M[k];
#pragma unroll
    for (int k = 0; k < 3; ++k) s[k] += v[k];
  }
  const T damp = T(nu * double(p.pt_nu[b]));
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
    inv[0] = a / det;
    inv[1] = (m13 * m32 - m33 * m12) / det;
    inv[2] = (m12 * m23 - m22 * m13) / det;
    inv[3] = -bb / det;
    inv[4] = (m11 * m33 - m31 * m13) / det;
    inv[5] = (m13 * m21 - m23 * m11) / det;
    inv[6] = c / det;
    inv[7] = (m12 * m31 - m32 * m11) / det;
    inv[8] = (m11 * m22 - m21 * m12) / det;
  }
  for (int k = 0; k < 9; ++k) hinv_out[9 * b + k] = inv[k];
  for (int k = 0; k < 3; ++k) {
    s_out[3 * b + k] = double(s[k]);
    y_out[3 * b + k] = inv[3 * k] * double(s[0]) + inv[3 * k + 1] * double(s[1]) +
                       inv[3 * k + 2] * double(s[2]);
    atomic_max_abs(&sc->max_grad_bits, double(s[k]));
  }
}

// P^T N P for P = [hat(-x) | I] and a 3x3 N: [[H N H^T, H N], [N H^T, N]] with H = hat(x).
// Adds sign * (P^T N P) to a 6x6 row-major accumulator via `add(idx, value)`.
template <typename T, typename Add>
__device__ __forceinline__ void add_lift(const T x[3], const T N[9], T sign, Add add) {
  // HN = hat(x) N ; NHt = N hat(x)^T = -N hat(x)
  T HN[9], NHt[9], HNHt[9];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    // hat(x) rows: [0,-x2,x1], [x2,0,-x0], [-x1,x0,0]
    HN[j] = -x[2] * N[3 + j] + x[1] * N[6 + j];
    HN[3 + j] = x[2] * N[j] - x[0] * N[6 + j];
    HN[6 + j] = -x[1] * N[j] + x[0] * N[3 + j];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    // (N hat(x)^T)[i][j] = sum_t N[i][t] hat(x)[j][t]
    NHt[3 * i + 0] = -N[3 * i + 1] * x[2] + N[3 * i + 2] * x[1];
    NHt[3 * i + 1] = N[3 * i + 0] * x[2] - N[3 * i + 2] * x[0];
    NHt[3 * i + 2] = -N[3 * i + 0] * x[1] + N[3 * i + 1] * x[0];
    HNHt[3 * i + 0] = -HN[3 * i + 1] * x[2] + HN[3 * i + 2] * x[1];
    HNHt[3 * i + 1] = HN[3 * i + 0] * x[2] - HN[3 * i + 2] * x[0];
    HNHt[3 * i + 2] = -HN[3 * i + 0] * x[1] + HN[3 * i + 1] * x[0];
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

template <typename T>
__device__ __forceinline__ void mat3_mul_t(const T A[9], const T B[9], T C[9]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}

// ---- kernel: one S row (camera c1) with its reduced rhs -----------------------------------
// One workgroup per free camera: visits the camera's observations and, for each, the
// co-observations of the same point with camera block <= c1.  Row blocks accumulate in
// LDS (row slots found by binary search) and are written once to the factor's
// unified storage (permuted, transposed where the permutation flips the block).
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_camera_rows(DevProblem p, const int32_t* __restrict__ obs_point,
                  const double* __restrict__ points, const double* __restrict__ intr,
                  const double* __restrict__ cw, double nu, const double* __restrict__ hinv_g,
                  const double* __restrict__ y_g, double* __restrict__ C,
                  double* __restrict__ rhs_nat, double* __restrict__ grad_cam, DevScalars* sc) {
  __shared__ int32_t cols[kMaxRowSlots];
  __shared__ T acc[kMaxRowSlots * 36];
  __shared__ T racc[6], gacc[6];
  const int c1 = blockIdx.x;
  const int row0 = p.s_ptr[c1], nrow = p.s_ptr[c1 + 1] - row0;
  const bool use_lds = nrow <= kMaxRowSlots;
  const int tid = threadIdx.x;
  if (use_lds) {
    for (int k = tid; k < nrow; k += blockDim.x) cols[k] = p.s_col[row0 + k];
    for (int k = tid; k < nrow * 36; k += blockDim.x) acc[k] = T(0);
  }
  if (tid < 6) {
    racc[tid] = T(0);
    gacc[tid] = T(0);
  }
  __syncthreads();
  const int32_t* gcols = p.s_col + row0;
  auto slot_of = [&](int c2) -> int {
    int lo = 0, hi = nrow - 1;
    const int32_t* cc = use_lds ? cols : gcols;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (cc[mid] < c2)
        lo = mid + 1;
      else
        hi = mid;
    }
    return lo;
  };
  auto add_block = [&](int slot, const T x[3], const T N[9], T sign) {
    if (use_lds) {
      T* dst = acc + 36 * slot;
      add_lift(x, N, sign, [&](int e, T v) { lds_add(dst + e, v); });
    } else {
      const int code = p.s_dest[row0 + slot];
      double* dst = C + 36 * (code >> 1);
      const bool tr = code & 1;
      add_lift(x, N, sign, [&](int e, T v) {
        const int idx = tr ? (e % 6) * 6 + e / 6 : e;
        atomic_add_d(dst + idx, double(v));
      });
    }
  };
  const double* cw1 = cw + 12 * p.cam_member[c1];
  for (int idx = p.cam_ptr[c1] + tid; idx < p.cam_ptr[c1 + 1]; idx += blockDim.x) {
    const int o = p.cam_obs[idx];
    const int b = obs_point[o];
    const double xd[3] = {points[3 * b], points[3 * b + 1], points[3 * b + 2]};
    const T x[3] = {T(xd[0]), T(xd[1]), T(xd[2])};
    const double* in = intr + 4 * p.obs_intr[o];
    double xc[3];
    to_camera(cw1, xd, xc);
    T A[6], v[3], Mo[9];
    obs_jacobian<T>(cw1, in, xc, A);
    obs_residual<T>(in, p.obs_uv + 2 * o, xc, p.robust, p.delta, A, v);
    gram(A, Mo);
    // own gradient: neg_grad_p -= P^T v  (P^T t = [x cross t; t])
    T g[6];
    g[0] = -(x[1] * v[2] - x[2] * v[1]);
    g[1] = -(x[2] * v[0] - x[0] * v[2]);
    g[2] = -(x[0] * v[1] - x[1] * v[0]);
    g[3] = -v[0];
    g[4] = -v[1];
    g[5] = -v[2];
#pragma unroll
    for (int k = 0; k < 6; ++k) lds_add(&gacc[k], g[k]);
    if (p.pt_block[b] < 0) {  // constant point: H_pp only, no Schur term
#pragma unroll
      for (int k = 0; k < 6; ++k) lds_add(&racc[k], g[k]);
      add_block(nrow - 1, x, Mo, T(1));
      continue;
    }
    T Hi[9], y[3];
#pragma unroll
    for (int k = 0; k < 9; ++k) Hi[k] = T(hinv_g[9 * b + k]);
#pragma unroll
    for (int k = 0; k < 3; ++k) y[k] = T(y_g[3 * b + k]);
    // rhs += P^T (M_o y - v)
    T t[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) t[k] = Mo[3 * k] * y[0] + Mo[3 * k + 1] * y[1] + Mo[3 * k + 2] * y[2] - v[k];
    lds_add(&racc[0], x[1] * t[2] - x[2] * t[1]);
    lds_add(&racc[1], x[2] * t[0] - x[0] * t[2]);
    lds_add(&racc[2], x[0] * t[1] - x[1] * t[0]);
    lds_add(&racc[3], t[0]);
    lds_add(&racc[4], t[1]);
    lds_add(&racc[5], t[2]);
    T Z[9];
    mat3_mul_t(Mo, Hi, Z);
    for (int j = p.pt_ptr[b]; j < p.pt_ptr[b + 1]; ++j) {
      const int pj = p.obs_pose[j];
      const int c2 = p.pose_block[pj];
      if (c2 < 0) continue;
      if (c2 > c1) break;
      T Mj[9];
      if (j == o) {
#pragma unroll
        for (int k = 0; k < 9; ++k) Mj[k] = Mo[k];
      } else {
        const double* cwj = cw + 12 * pj;
        double xcj[3];
        to_camera(cwj, xd, xcj);
        T Aj[6];
        obs_jacobian<T>(cwj, intr + 4 * p.obs_intr[j], xcj, Aj);
        gram(Aj, Mj);
      }
      T N[9];
      mat3_mul_t(Z, Mj, N);
      if (j == o) {
#pragma unroll
        for (int k = 0; k < 9; ++k) N[k] -= Mo[k];
      }
      add_block(slot_of(c2), x, N, T(-1));
    }
  }
  __syncthreads();
  const T damp = T(nu * double(p.cam_nu[c1]));
  if (use_lds) {
    if (tid < 6) acc[36 * (nrow - 1) + 7 * tid] += damp;
    __syncthreads();
    for (int k = tid; k < nrow * 36; k += blockDim.x) {
      const int slot = k / 36, e = k % 36;
      const int code = p.s_dest[row0 + slot];
      const int idx = (code & 1) ? (e % 6) * 6 + e / 6 : e;
      C[36 * (code >> 1) + idx] = double(acc[k]);
    }
  } else if (tid < 6) {
    const int code = p.s_dest[row0 + nrow - 1];
    atomic_add_d(C + 36 * (code >> 1) + 7 * tid, double(damp));
  }
  if (tid < 6) {
    rhs_nat[6 * c1 + tid] = double(racc[tid]);
    grad_cam[6 * c1 + tid] = double(gacc[tid]);
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

// ---- RCS block LDLt: one workgroup per column of an etree level ---------------------------
// ldlt_decompose_inplace (dense_ldlt.rs:164-256) on a 6x6 block with the PSD pivot rule.
__device__ bool ldlt6(const double* A, double* L, double* d) {
  double a[36];
  for (int k = 0; k < 36; ++k) a[k] = A[k];
  double max_abs_pivot = 1.0;
  const double tol = 1e-12;
  for (int j = 0; j < 6; ++j) {
    for (int k = 0; k < j; ++k) {
      double s = a[6 * j + k];
      for (int q = 0; q < k; ++q) s -= a[6 * j + q] * d[q] * a[6 * k + q];
      const double zt = fmax(max_abs_pivot, 1.0) * tol;
      a[6 * j + k] = fabs(d[k]) <= zt ? 0.0 : s / d[k];
    }
    double djj = a[6 * j + j];
    for (int q = 0; q < j; ++q) {
      const double l = a[6 * j + q];
      djj -= d[q] * l * l;
    }
    if (!isfinite(djj)) return false;
    const double zt = fmax(max_abs_pivot, 1.0) * tol;
    if (fabs(djj) <= zt) {
      d[j] = 0.0;
    } else {
      d[j] = djj;
      max_abs_pivot = fmax(max_abs_pivot, fabs(djj));
    }
    a[6 * j + j] = 1.0;
  }
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c) L[6 * r + c] = c > r ? 0.0 : a[6 * r + c];
  return true;
}

__global__ void __launch_bounds__(kThreads) k_factor_level(DevFactor f, int level, DevScalars* sc) {
  const int j = f.lvl_col[f.lvl_ptr[level] + blockIdx.x];
  const int nb = f.nb;
  const int tid = threadIdx.x;
  const int r0 = f.r_ptr[j], r1 = f.r_ptr[j + 1];
  // phase 1: M_jk = D_k L_jk^T for k in reach(j)
  for (int e = tid; e < (r1 - r0) * 36; e += blockDim.x) {
    const int r = r0 + e / 36, q = e % 36, a = q / 6, bc = q % 6;
    const double* Dk = f.Dfull + 36 * f.r_col[r];
    const double* Ljk = f.C + 36 * (nb + f.r_lpos[r]);
    double s = 0;
#pragma unroll
    for (int t = 0; t < 6; ++t) s += Dk[6 * a + t] * Ljk[6 * bc + t];
    f.M[36 * r + q] = s;
  }
  __syncthreads();
  // phase 2: C(:, j) -= sum_k L(:, k) M_jk over the gather lists of column j's slots
  const int nslot = 1 + f.l_ptr[j + 1] - f.l_ptr[j];
  for (int e = tid; e < nslot * 36; e += blockDim.x) {
    const int s = e / 36, q = e % 36, a = q / 6, bc = q % 6;
    const int u = s == 0 ? j : nb + f.l_ptr[j] + s - 1;
    double acc = f.C[36 * u + q];
    for (int w = f.u_ptr[u]; w < f.u_ptr[u + 1]; ++w) {
      const double* Lik = f.C + 36 * (nb + f.u_src[w]) + 6 * a;
      const double* Mjk = f.M + 36 * f.u_m[w] + bc;
      double t = 0;
#pragma unroll
      for (int k = 0; k < 6; ++k) t += Lik[k] * Mjk[6 * k];
      acc -= t;
    }
    f.C[36 * u + q] = acc;
  }
  __syncthreads();
  // phase 3: pivot block
  if (tid == 0) {
    double L[36], d[6];
    if (!ldlt6(f.C + 36 * j, L, d)) {
      atomicOr(&sc->bad_pivot, 1);
      for (int k = 0; k < 36; ++k) L[k] = (k % 7 == 0) ? 1.0 : 0.0;
      for (int k = 0; k < 6; ++k) d[k] = 0.0;
    }
    for (int k = 0; k < 36; ++k) f.Ljj[36 * j + k] = L[k];
    for (int k = 0; k < 6; ++k) f.dvec[6 * j + k] = d[k];
    for (int a = 0; a < 6; ++a)
      for (int b = 0; b < 6; ++b) {
        double s = 0;
        for (int t = 0; t < 6; ++t) s += L[6 * a + t] * d[t] * L[6 * b + t];
        f.Dfull[36 * j + 6 * a + b] = s;
      }
  }
  __syncthreads();
  // phase 4: L_ij = C_ij L_jj^-T D_j^-1 L_jj^-1, one row of one block per thread
  // (ldlt_right_matsolve_inplace, dense_ldlt.rs:300-308)
  const double* L = f.Ljj + 36 * j;
  const double* d = f.dvec + 6 * j;
  for (int e = tid; e < (nslot - 1) * 6; e += blockDim.x) {
    const int s = 1 + e / 6, a = e % 6;
    double* X = f.C + 36 * (nb + f.l_ptr[j] + s - 1) + 6 * a;
    double x[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) x[k] = X[k];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double v = x[c];
      for (int t = 0; t < c; ++t) v -= x[t] * L[6 * c + t];
      x[c] = v;
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) x[c] *= d[c] == 0.0 ? 0.0 : 1.0 / d[c];
#pragma unroll
    for (int c = 5; c >= 0; --c) {
      double v = x[c];
      for (int t = c + 1; t < 6; ++t) v -= x[t] * L[6 * t + c];
      x[c] = v;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) X[k] = x[k];
  }
}

// ---- triangular solves (single workgroup walking the etree levels) ---------------------
__global__ void __launch_bounds__(1024)
    k_triangular_solve(DevFactor f, const double* __restrict__ rhs_nat, double* __restrict__ y,
                       double* __restrict__ dx_cam) {
  const int nb = f.nb, tid = threadIdx.x, nt = blockDim.x;
  for (int e = tid; e < 6 * nb; e += nt) y[e] = rhs_nat[6 * f.perm[e / 6] + e % 6];
  __syncthreads();
  // L y = b (gather over the row structure of L)
  for (int lev = 0; lev < f.nlev; ++lev) {
    const int c0 = f.lvl_ptr[lev], cn = f.lvl_ptr[lev + 1] - c0;
    for (int e = tid; e < cn * 6; e += nt) {
      const int j = f.lvl_col[c0 + e / 6], c = e % 6;
      double acc = y[6 * j + c];
      for (int r = f.r_ptr[j]; r < f.r_ptr[j + 1]; ++r) {
        const double* Ljk = f.C + 36 * (nb + f.r_lpos[r]) + 6 * c;
        const double* yk = y + 6 * f.r_col[r];
        double t = 0;
#pragma unroll
        for (int k = 0; k < 6; ++k) t += Ljk[k] * yk[k];
        acc -= t;
      }
      y[6 * j + c] = acc;
    }
    __syncthreads();
  }
  // D z = y per block (lower, diag with d == 0 -> 0, lower transpose)
  for (int j = tid; j < nb; j += nt) {
    const double* L = f.Ljj + 36 * j;
    const double* d = f.dvec + 6 * j;
    double v[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] = y[6 * j + k];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      double s = v[i];
      for (int k = 0; k < i; ++k) s -= L[6 * i + k] * v[k];
      v[i] = s;
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) v[i] = d[i] == 0.0 ? 0.0 : v[i] / d[i];
#pragma unroll
    for (int i = 5; i >= 0; --i) {
      double s = v[i];
      for (int k = i + 1; k < 6; ++k) s -= L[6 * k + i] * v[k];
      v[i] = s;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) y[6 * j + k] = v[k];
  }
  __syncthreads();
  // L^T x = z (gather over the column structure of L), levels top-down
  for (int lev = f.nlev - 1; lev >= 0; --lev) {
    const int c0 = f.lvl_ptr[lev], cn = f.lvl_ptr[lev + 1] - c0;
    for (int e = tid; e < cn * 6; e += nt) {
      const int j = f.lvl_col[c0 + e / 6], c = e % 6;
      double acc = y[6 * j + c];
      for (int q = f.l_ptr[j]; q < f.l_ptr[j + 1]; ++q) {
        const double* Lij = f.C + 36 * (nb + q) + c;
        const double* xi = y + 6 * f.l_row[q];
        double t = 0;
#pragma unroll
        for (int k = 0; k < 6; ++k) t += Lij[6 * k] * xi[k];
        acc -= t;
      }
      y[6 * j + c] = acc;
    }
    __syncthreads();
  }
  for (int e = tid; e < 6 * nb; e += nt) dx_cam[6 * f.perm[e / 6] + e % 6] = y[e];
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
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_backsub(DevProblem p, const double* __restrict__ points, const double* __restrict__ intr,
              const double* __restrict__ cw, const double* __restrict__ hinv,
              const double* __restrict__ pt_s, const double* __restrict__ dx_cam,
              double* __restrict__ points_new, double* __restrict__ partial, DevScalars* sc) {
  double dd = 0, dg = 0;
  for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < p.npt; b += gridDim.x * blockDim.x) {
    const double xd[3] = {points[3 * b], points[3 * b + 1], points[3 * b + 2]};
    if (p.pt_block[b] < 0) {
      for (int k = 0; k < 3; ++k) points_new[3 * b + k] = xd[k];
      continue;
    }
    const T x[3] = {T(xd[0]), T(xd[1]), T(xd[2])};
    T acc[3] = {T(pt_s[3 * b]), T(pt_s[3 * b + 1]), T(pt_s[3 * b + 2])};
    for (int o = p.pt_ptr[b]; o < p.pt_ptr[b + 1]; ++o) {
      const int pm = p.obs_pose[o];
      const int c = p.pose_block[pm];
      if (c < 0) continue;
      const double* cwo = cw + 12 * pm;
      double xc[3];
      to_camera(cwo, xd, xc);
      T A[6], M[9];
      obs_jacobian<T>(cwo, intr + 4 * p.obs_intr[o], xc, A);
      gram(A, M);
      const double* dx = dx_cam + 6 * c;
      const T w0 = T(dx[0]), w1 = T(dx[1]), w2 = T(dx[2]);
      // P dx = omega x point + nu
      const T q0 = w1 * x[2] - w2 * x[1] + T(dx[3]);
      const T q1 = w2 * x[0] - w0 * x[2] + T(dx[4]);
      const T q2 = w0 * x[1] - w1 * x[0] + T(dx[5]);
#pragma unroll
      for (int k = 0; k < 3; ++k) acc[k] += M[3 * k] * q0 + M[3 * k + 1] * q1 + M[3 * k + 2] * q2;
    }
    const double* Hi = hinv + 9 * b;
    double mx = 0;
    for (int k = 0; k < 3; ++k) {
      const double dk = Hi[3 * k] * double(acc[0]) + Hi[3 * k + 1] * double(acc[1]) +
                        Hi[3 * k + 2] * double(acc[2]);
      points_new[3 * b + k] = xd[k] + dk;
      dd += dk * dk;
      dg += dk * pt_s[3 * b + k];
      mx = fmax(mx, fabs(dk));
    }
    atomic_max_abs(&sc->max_delta_bits, mx);
  }
  block_reduce2(dd, dg, partial);
}

// ---- pose update T <- exp(delta) T ----------------------------------------------------------
__global__ void __launch_bounds__(kThreads)
    k_update_poses(DevProblem p, const double* __restrict__ poses, const double* __restrict__ dx_cam,
                   const double* __restrict__ grad_cam, double* __restrict__ poses_new,
                   double* __restrict__ partial, DevScalars* sc) {
  double dd = 0, dg = 0;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < p.nb; c += gridDim.x * blockDim.x) {
    const int pm = p.cam_member[c];
    double t[6], T0[7], T1[7];
    double mx = 0;
    for (int k = 0; k < 6; ++k) {
      t[k] = dx_cam[6 * c + k];
      dd += t[k] * t[k];
      dg += t[k] * grad_cam[6 * c + k];
      mx = fmax(mx, fabs(t[k]));
    }
    for (int k = 0; k < 7; ++k) T0[k] = poses[7 * pm + k];
    se3_oplus(T0, t, T1);
    for (int k = 0; k < 7; ++k) poses_new[7 * pm + k] = T1[k];
    atomic_max_abs(&sc->max_delta_bits, mx);
  }
  block_reduce2(dd, dg, partial);
}

// ---- merit: sum over terms of r~^T r~ ------------------------------------------------------
__global__ void __launch_bounds__(kThreads)
    k_cost(DevProblem p, const int32_t* __restrict__ obs_point, const double* __restrict__ poses,
           const double* __restrict__ points, const double* __restrict__ intr,
           double* __restrict__ partial) {
  double acc = 0;
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < p.nobs; o += gridDim.x * blockDim.x) {
    const double* T = poses + 7 * p.obs_pose[o];
    const double* in = intr + 4 * p.obs_intr[o];
    const int b = obs_point[o];
    double inv[7], R[9], xc[3];
    se3_inverse(T, inv);
    so3_matrix(inv, R);
    const double x[3] = {points[3 * b], points[3 * b + 1], points[3 * b + 2]};
    mat3_vec(R, x, xc);
    xc[0] += inv[4];
    xc[1] += inv[5];
    xc[2] += inv[6];
    const double px = xc[0] / xc[2], py = xc[1] / xc[2];
    const double r0 = p.obs_uv[2 * o] - (px * in[0] + in[2]);
    const double r1 = p.obs_uv[2 * o + 1] - (py * in[1] + in[3]);
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

// Sum partials in a fixed order; finish the step scalars.
__global__ void k_finalize(const double* __restrict__ partial, int n_pt_blocks, int n_pose_blocks,
                           int n_cost_blocks, const double* __restrict__ grad_cam, int nb,
                           DevScalars* sc, int what) {
  __shared__ double red[256];
  if (what == 0) {  // dd, dg from back-substitution + pose update partials
    if (threadIdx.x == 0) {
      double dd = 0, dg = 0;
      for (int i = 0; i < n_pt_blocks; ++i) {
        dd += partial[2 * i];
        dg += partial[2 * i + 1];
      }
      const double* q = partial + 2 * n_pt_blocks;
      for (int i = 0; i < n_pose_blocks; ++i) {
        dd += q[2 * i];
        dg += q[2 * i + 1];
      }
      sc->dd = dd;
      sc->dg = dg;
    }
  } else if (what == 1) {  // cost
    if (threadIdx.x == 0) {
      double c = 0;
      for (int i = 0; i < n_cost_blocks; ++i) c += partial[2 * i];
      sc->cost = c;
    }
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

__global__ void k_pose_graph_linearize(DevPoseGraph g, DevProblem p, const double* __restrict__ poses,
                                       double nu, double* __restrict__ C, double* __restrict__ rhs_nat,
                                       double* __restrict__ grad_cam) {
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
  const int* dst = g.dest + 3 * t;
  const double damp = g.first[t] ? nu : 0.0;
  auto add = [&](int code, double sign, double diag) {
    if (code < 0) return;
    double* B = C + 36 * (code >> 1);
    for (int e = 0; e < 36; ++e) {
      const int idx = (code & 1) ? (e % 6) * 6 + e / 6 : e;
      double v = sign * DtD[e];
      if (e % 7 == 0) v += diag;
      if (v != 0.0) atomic_add_d(B + idx, v);
    }
  };
  add(dst[0], 1.0, damp);
  add(dst[1], 1.0, damp);
  add(dst[2], -1.0, 0.0);
  // neg_grad_m -= J_m^T r~ = +D^T r~ ; neg_grad_n -= D^T r~
  if (cm >= 0)
    for (int k = 0; k < 6; ++k) {
      atomic_add_d(rhs_nat + 6 * cm + k, Dtr[k]);
      atomic_add_d(grad_cam + 6 * cm + k, Dtr[k]);
    }
  if (cn >= 0)
    for (int k = 0; k < 6; ++k) {
      atomic_add_d(rhs_nat + 6 * cn + k, -Dtr[k]);
      atomic_add_d(grad_cam + 6 * cn + k, -Dtr[k]);
    }
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
  if (precision == 0)
    hipLaunchKernelGGL(k_point_prep<double>, g, b, 0, s, p, points, intr, cw, nu, w.hinv, w.pt_s,
                       w.pt_y, sc);
  else
    hipLaunchKernelGGL(k_point_prep<float>, g, b, 0, s, p, points, intr, cw, nu, w.hinv, w.pt_s,
                       w.pt_y, sc);
}

void launch_camera_rows(hipStream_t s, int precision, const DevProblem& p, const double* points,
                        const double* intr, const double* cw, double nu, DevLin& w, DevFactor& f,
                        DevScalars* sc) {
  (void)sc;
  if (p.nb == 0) return;
  const int32_t* obs_point = p.obs_point;
  if (precision == 0)
    hipLaunchKernelGGL(k_camera_rows<double>, dim3(p.nb), dim3(kThreads), 0, s, p, obs_point,
                       points, intr, cw, nu, w.hinv, w.pt_y, f.C, w.rhs_nat, w.grad_cam, sc);
  else
    hipLaunchKernelGGL(k_camera_rows<float>, dim3(p.nb), dim3(kThreads), 0, s, p, obs_point,
                       points, intr, cw, nu, w.hinv, w.pt_y, f.C, w.rhs_nat, w.grad_cam, sc);
}

void launch_fill(hipStream_t s, double* ptr, int64_t n, double v) {
  if (n == 0) return;
  const int g = int(std::min<int64_t>((n + kThreads - 1) / kThreads, 4096));
  hipLaunchKernelGGL(k_fill, dim3(g), dim3(kThreads), 0, s, ptr, n, v);
}

void launch_zero_scalars(hipStream_t s, DevScalars* sc) {
  hipLaunchKernelGGL(k_zero_scalars, dim3(1), dim3(64), 0, s, sc);
}

void launch_factor_level(hipStream_t s, const DevFactor& f, int level, DevScalars* sc) {
  const int ncol = f.h_lvl_ptr[level + 1] - f.h_lvl_ptr[level];
  if (ncol == 0) return;
  hipLaunchKernelGGL(k_factor_level, dim3(ncol), dim3(kThreads), 0, s, f, level, sc);
}

void launch_triangular_solve(hipStream_t s, const DevFactor& f, const double* rhs_nat,
                             double* x_perm, double* x_aux, double* dx_cam) {
  (void)x_aux;
  hipLaunchKernelGGL(k_triangular_solve, dim3(1), dim3(1024), 0, s, f, rhs_nat, x_perm, dx_cam);
}

void launch_backsub_points(hipStream_t s, int precision, const DevProblem& p, const double* points,
                           const double* intr, const double* cw, const DevLin& w,
                           double* points_new, DevScalars* sc) {
  const int g = std::max(1, std::min(blocks_for(p.npt, kThreads), kCostBlocks));
  if (precision == 0)
    hipLaunchKernelGGL(k_backsub<double>, dim3(g), dim3(kThreads), 0, s, p, points, intr, cw,
                       w.hinv, w.pt_s, w.dx_cam, points_new, w.partial, sc);
  else
    hipLaunchKernelGGL(k_backsub<float>, dim3(g), dim3(kThreads), 0, s, p, points, intr, cw,
                       w.hinv, w.pt_s, w.dx_cam, points_new, w.partial, sc);
}

void launch_update_poses(hipStream_t s, const DevProblem& p, const double* poses, const DevLin& w,
                         double* poses_new, DevScalars* sc) {
  const int gpt = std::max(1, std::min(blocks_for(p.npt, kThreads), kCostBlocks));
  const int g = std::max(1, std::min(blocks_for(p.nb, kThreads), kCostBlocks));
  hipLaunchKernelGGL(k_update_poses, dim3(g), dim3(kThreads), 0, s, p, poses, w.dx_cam,
                     w.grad_cam, poses_new, w.partial + 2 * gpt, sc);
}

void launch_finalize_step(hipStream_t s, const DevLin& w, int nb, int npt, DevScalars* sc) {
  const int gpt = std::max(1, std::min(blocks_for(npt, kThreads), kCostBlocks));
  const int gpo = std::max(1, std::min(blocks_for(nb, kThreads), kCostBlocks));
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, w.partial, gpt, gpo, 0, nullptr, 0, sc,
                     0);
}

void launch_cost(hipStream_t s, const DevProblem& p, const double* poses, const double* points,
                 const double* intr, double* partial, int32_t partial_cap, DevScalars* sc) {
  (void)partial_cap;
  hipLaunchKernelGGL(k_cost, dim3(kCostBlocks), dim3(kThreads), 0, s, p, p.obs_point, poses,
                     points, intr, partial);
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
                                 const double* poses, double nu, DevLin& w, DevFactor& f) {
  if (g.n == 0) return;
  hipLaunchKernelGGL(k_pose_graph_linearize, dim3(blocks_for(g.n, 128)), dim3(128), 0, s, g, p,
                     poses, nu, f.C, w.rhs_nat, w.grad_cam);
}

void launch_pose_graph_cost(hipStream_t s, const DevPoseGraph& g, const double* poses,
                            double* partial, int32_t partial_cap, DevScalars* sc) {
  (void)partial_cap;
  (void)sc;
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

void launch_eval_pose_graph(hipStream_t s, int64_t n, const double* m, const double* nn,
                            const double* mn, double* res, double* jm, double* jn) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_eval_pose_graph, dim3(blocks_for(n, 128)), dim3(128), 0, s, n, m, nn, mn,
                     res, jm, jn);
}

}  // namespace sgpu
