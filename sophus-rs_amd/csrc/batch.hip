// Batched Schur LM for many small, independent bundle-adjustment windows.
//
// Every window (one shared pinhole, SE(3) poses with pose 0 constant, marginalized
// points, Huber reprojection terms) has at most kBatchMaxCams cameras, so its reduced
// camera system (n = 6 (cameras - 1) <= 180) lives in LDS as a packed lower triangle.
// One workgroup runs one complete LM iteration of one window:
//   per point (threads): A_o, M_o = A^T A, H_ll + nu n I, H_ll^-1, s, y; the point's
//     Schur terms go into S with LDS atomics (diagonal own term M H^-1 (H - M),
//     off-diagonal -M_i H^-1 M_j, lifted by P = [hat(-x) | I]), rhs, gradient;
//   dense block LDLt of S in LDS (6x6 pivots, the reference's PSD rule), solves;
//   back-substitution, T <- exp(delta) T, merit of the proposal (FP64), and the LM
//   decision with the same rules and check order as the single-problem optimizer.
// One launch advances every unfinished window by one iteration; the host captures a
// sequence of launches in a hipGraph.  Precision T = float computes the Jacobians and
// Schur products in FP32 (S, the factorization, the merit and the decisions stay FP64).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <numeric>
#include <vector>

#include "api_internal.hpp"
#include "block6.hpp"
#include "lie.hpp"

namespace sgpu {

constexpr int kBatchMaxCams = 31;

struct BatchDev {
  int32_t nprob;
  const int32_t* order;      // launch slot -> problem (largest first)
  const int32_t* cam0;       // problem -> first camera (prefix, nprob + 1)
  const int32_t* pt0;        // problem -> first point
  const int32_t* obs0;       // problem -> first observation
  const double* intr;        // 4 per problem
  const double* poses_init;  // 7 per camera
  const double* points_init; // 3 per point
  double* poses[2];
  double* points[2];
  const int32_t* obs_cam;    // camera within the problem, sorted by point
  const double* obs_uv;
  const int32_t* pt_obs;     // global point -> first observation (npt_total + 1)
  const int32_t* pt_nu;      // reduced terms per point
  const int32_t* cam_nu;     // reduced terms per camera
  double* hinv;              // 9 per point
  double* pts;               // 3 per point (the point's neg-gradient s)
  // per-problem LM state
  int32_t* cur;
  double* nu;
  double* merit;
  double* init_merit;
  int32_t* rejects;
  int32_t* steps;
  int32_t* done;
  int32_t* term;
  // parameters
  int32_t robust;
  double delta;
  double nu0;
  int64_t num_iterations;
  double function_tolerance, gradient_tolerance, parameter_tolerance;
  int64_t max_consecutive_rejects;
  int32_t lds_doubles;  // dynamic LDS per workgroup
  int32_t n_max;        // largest RCS dimension of the batch
};

namespace {

__host__ __device__ __forceinline__ int64_t pk(int r, int c) { return int64_t(r) * (r + 1) / 2 + c; }  // r >= c

template <int NT>
__device__ double wg_sum(double v, double* red) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < NT / 64; ++w) t += red[w];
  __syncthreads();
  return t;
}

template <int NT>
__device__ double wg_max(double v, double* red) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v = fmax(v, __shfl_xor(v, m, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < NT / 64; ++w) t = fmax(t, red[w]);
  __syncthreads();
  return t;
}

__device__ __forceinline__ double huber_weight(int robust, double delta, double nrm) {
  if (robust != 1) return 1.0;
  if (nrm == 0.0) return 0.0;
  const double a = nrm <= delta ? nrm * nrm : 2.0 * delta * nrm - delta * delta;
  return sqrt(a) / nrm;
}

// camera_from_world R (row-major), t and the camera centre c, per camera, into LDS.
__device__ void cam_cache(const double* pose, double* out) {
  double inv[7], R[9];
  se3_inverse(pose, inv);
  so3_matrix(inv, R);
  for (int k = 0; k < 9; ++k) out[k] = R[k];
  for (int k = 0; k < 3; ++k) out[9 + k] = inv[4 + k];
  for (int k = 0; k < 3; ++k) out[12 + k] = pose[4 + k];
}

// Residual r = uv - K pi(R x + t) (FP64) and the Jacobian A = diag(f) dpi R (2x3, in T).
template <typename T>
__device__ __forceinline__ void obs_eval(const double* cc, const double* K, const double* uv,
                                         const double x[3], double r[2], T A[6]) {
  const double xc0 = cc[0] * x[0] + cc[1] * x[1] + cc[2] * x[2] + cc[9];
  const double xc1 = cc[3] * x[0] + cc[4] * x[1] + cc[5] * x[2] + cc[10];
  const double xc2 = cc[6] * x[0] + cc[7] * x[1] + cc[8] * x[2] + cc[11];
  const double iz = 1.0 / xc2;
  r[0] = uv[0] - (K[0] * xc0 * iz + K[2]);
  r[1] = uv[1] - (K[1] * xc1 * iz + K[3]);
  const T d00 = T(K[0] * iz), d02 = T(-K[0] * xc0 * iz * iz);
  const T d11 = T(K[1] * iz), d12 = T(-K[1] * xc1 * iz * iz);
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    A[j] = d00 * T(cc[j]) + d02 * T(cc[6 + j]);
    A[3 + j] = d11 * T(cc[3 + j]) + d12 * T(cc[6 + j]);
  }
}

template <typename T>
__device__ __forceinline__ void gram3(const T A[6], T M[9]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) M[3 * i + j] = A[i] * A[j] + A[3 + i] * A[3 + j];
}

template <typename T>
__device__ __forceinline__ void mul3(const T* A, const T* B, T* C) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}

// S(ci, cj) += sign * P_i^T N P_j (P = [hat(-x) | I], the same point x), lower packed,
// block row ci >= block column cj.
template <typename T>
__device__ void lift_add(double* S, int ci, int cj, const T x[3], const T N[9], double sign) {
  T HN[9], NHt[9], HNHt[9];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    HN[j] = -x[2] * N[3 + j] + x[1] * N[6 + j];
    HN[3 + j] = x[2] * N[j] - x[0] * N[6 + j];
    HN[6 + j] = -x[1] * N[j] + x[0] * N[3 + j];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    NHt[3 * i + 0] = -N[3 * i + 1] * x[2] + N[3 * i + 2] * x[1];
    NHt[3 * i + 1] = N[3 * i + 0] * x[2] - N[3 * i + 2] * x[0];
    NHt[3 * i + 2] = -N[3 * i + 0] * x[1] + N[3 * i + 1] * x[0];
    HNHt[3 * i + 0] = -HN[3 * i + 1] * x[2] + HN[3 * i + 2] * x[1];
    HNHt[3 * i + 1] = HN[3 * i + 0] * x[2] - HN[3 * i + 2] * x[0];
    HNHt[3 * i + 2] = -HN[3 * i + 0] * x[1] + HN[3 * i + 1] * x[0];
  }
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) {
      const int r = 6 * ci + a, c = 6 * cj + b;
      if (ci == cj && c > r) continue;  // lower triangle of a diagonal block
      T v;
      if (a < 3 && b < 3) v = HNHt[3 * a + b];
      else if (a < 3) v = HN[3 * a + b - 3];
      else if (b < 3) v = NHt[3 * (a - 3) + b];
      else v = N[3 * (a - 3) + b - 3];
      if (v != T(0)) atomicAdd(&S[pk(r, c)], sign * double(v));
    }
}

__device__ __forceinline__ double sym_at(const double* S, int r, int c) {
  return r >= c ? S[pk(r, c)] : S[pk(c, r)];
}

// Solve D x = v with D = L6 diag(d) L6^T stored in the packed diagonal block of k
// (strict lower: L6, diagonal: d); zero pivots map to 0.
__device__ void d_solve(const double* S, int k, double v[6]) {
  const int o = 6 * k;
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < i; ++j) v[i] -= S[pk(o + i, o + j)] * v[j];
  for (int i = 0; i < 6; ++i) {
    const double d = S[pk(o + i, o + i)];
    v[i] = d == 0.0 ? 0.0 : v[i] / d;
  }
  for (int i = 5; i >= 0; --i)
    for (int j = i + 1; j < 6; ++j) v[i] -= S[pk(o + j, o + i)] * v[j];
}

// One LM iteration of unfinished window p by the whole workgroup (lds: the dynamic LDS, red
// and flags (0: singular marginal block, 1: bad pivot) its static scratch).  Thread 0 takes
// the LM decision and writes the window's state; the other threads return before it, so a
// caller that goes on synchronizes first.
template <typename T, int NT>
__device__ void batch_iteration(const BatchDev& d, int p, double* lds, double* red, int* flags) {
  const int c0 = d.cam0[p], nc = d.cam0[p + 1] - c0;
  const int q0 = d.pt0[p], np = d.pt0[p + 1] - q0;
  const int o0 = d.obs0[p], no = d.obs0[p + 1] - o0;
  const int nbk = nc - 1, n = 6 * nbk;  // camera 0 is constant
  const int cur = d.cur[p], nxt = 1 - cur;
  const double nu = d.nu[p];
  const double* K = d.intr + 4 * p;
  double* S = lds;
  double* W = S + pk(d.n_max, 0);
  double* rhs = W + 6 * d.n_max;
  double* grad = rhs + d.n_max;
  double* xs = grad + d.n_max;
  double* cams = xs + d.n_max;  // 15 per camera
  const int tid = threadIdx.x;
  const double* poses = d.poses[cur] + 7 * int64_t(c0);
  const double* points = d.points[cur] + 3 * int64_t(q0);
  // ---- linearize + Schur -------------------------------------------------------------------
  for (int c = tid; c < nc; c += NT) cam_cache(poses + 7 * c, cams + 15 * c);
  for (int64_t e = tid; e < pk(n, 0); e += NT) S[e] = 0.0;
  for (int e = tid; e < n; e += NT) {
    rhs[e] = 0.0;
    grad[e] = 0.0;
  }
  if (tid < 4) flags[tid] = 0;
  __syncthreads();
  double gmax = 0.0;
  for (int b = tid; b < np; b += NT) {
    const int gb = q0 + b;
    const double x[3] = {points[3 * b], points[3 * b + 1], points[3 * b + 2]};
    const int ob = d.pt_obs[gb], oe = d.pt_obs[gb + 1];
    double H[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, s[3] = {0, 0, 0};
    for (int o = ob; o < oe; ++o) {
      double r[2];
      T A[6], M[9];
      obs_eval<T>(cams + 15 * d.obs_cam[o], K, d.obs_uv + 2 * int64_t(o), x, r, A);
      const double w = huber_weight(d.robust, d.delta, sqrt(r[0] * r[0] + r[1] * r[1]));
      const double rw0 = d.robust == 1 ? w * r[0] : r[0], rw1 = d.robust == 1 ? w * r[1] : r[1];
      gram3(A, M);
      for (int k = 0; k < 9; ++k) H[k] += double(M[k]);
      for (int k = 0; k < 3; ++k) s[k] += double(A[k]) * rw0 + double(A[3 + k]) * rw1;
    }
    const double damp = nu * double(d.pt_nu[gb]);
    H[0] += damp;
    H[4] += damp;
    H[8] += damp;
    const double a0 = H[4] * H[8] - H[7] * H[5], a1 = H[3] * H[8] - H[6] * H[5], a2 = H[3] * H[7] - H[6] * H[4];
    const double det = H[0] * a0 - H[1] * a1 + H[2] * a2;
    double Hi[9];
    if (det == 0.0 || !isfinite(det)) {
      flags[0] = 1;
      for (int k = 0; k < 9; ++k) Hi[k] = 0.0;
    } else {
      Hi[0] = a0 / det;
      Hi[1] = (H[2] * H[7] - H[8] * H[1]) / det;
      Hi[2] = (H[1] * H[5] - H[4] * H[2]) / det;
      Hi[3] = -a1 / det;
      Hi[4] = (H[0] * H[8] - H[6] * H[2]) / det;
      Hi[5] = (H[2] * H[3] - H[5] * H[0]) / det;
      Hi[6] = a2 / det;
      Hi[7] = (H[1] * H[6] - H[7] * H[0]) / det;
      Hi[8] = (H[0] * H[4] - H[3] * H[1]) / det;
    }
    for (int k = 0; k < 9; ++k) d.hinv[9 * int64_t(gb) + k] = Hi[k];
    double y[3];
    for (int k = 0; k < 3; ++k) {
      d.pts[3 * int64_t(gb) + k] = s[k];
      gmax = fmax(gmax, fabs(s[k]));
      y[k] = Hi[3 * k] * s[0] + Hi[3 * k + 1] * s[1] + Hi[3 * k + 2] * s[2];
    }
    const T xt[3] = {T(x[0]), T(x[1]), T(x[2])};
    T HiT[9];
    for (int k = 0; k < 9; ++k) HiT[k] = T(Hi[k]);
    for (int oi = ob; oi < oe; ++oi) {
      const int ci = d.obs_cam[oi] - 1;
      if (ci < 0) continue;
      double r[2];
      T Ai[6], Mi[9];
      obs_eval<T>(cams + 15 * (ci + 1), K, d.obs_uv + 2 * int64_t(oi), x, r, Ai);
      const double w = huber_weight(d.robust, d.delta, sqrt(r[0] * r[0] + r[1] * r[1]));
      const double rw0 = d.robust == 1 ? w * r[0] : r[0], rw1 = d.robust == 1 ? w * r[1] : r[1];
      gram3(Ai, Mi);
      double v[3], t[3];
      for (int k = 0; k < 3; ++k) v[k] = double(Ai[k]) * rw0 + double(Ai[3 + k]) * rw1;
      for (int k = 0; k < 3; ++k)
        t[k] = double(Mi[3 * k]) * y[0] + double(Mi[3 * k + 1]) * y[1] + double(Mi[3 * k + 2]) * y[2] - v[k];
      // rhs += P^T t = [x cross t; t]; gradient -= P^T v
      const double pt[6] = {x[1] * t[2] - x[2] * t[1], x[2] * t[0] - x[0] * t[2], x[0] * t[1] - x[1] * t[0],
                            t[0], t[1], t[2]};
      const double pv[6] = {x[1] * v[2] - x[2] * v[1], x[2] * v[0] - x[0] * v[2], x[0] * v[1] - x[1] * v[0],
                            v[0], v[1], v[2]};
      for (int k = 0; k < 6; ++k) {
        atomicAdd(&rhs[6 * ci + k], pt[k]);
        atomicAdd(&grad[6 * ci + k], -pv[k]);
      }
      // own term M_i H^-1 (H - M_i)
      T R[9], Z[9], N[9];
      for (int k = 0; k < 9; ++k) R[k] = T(H[k] - double(Mi[k]));
      mul3(Mi, HiT, Z);
      mul3(Z, R, N);
      lift_add(S, ci, ci, xt, N, 1.0);
      for (int oj = ob; oj < oe; ++oj) {
        const int cj = d.obs_cam[oj] - 1;
        if (oj == oi || cj < 0 || cj > ci) continue;
        double rj[2];
        T Aj[6], Mj[9];
        obs_eval<T>(cams + 15 * (cj + 1), K, d.obs_uv + 2 * int64_t(oj), x, rj, Aj);
        gram3(Aj, Mj);
        mul3(Z, Mj, N);
        lift_add(S, ci, cj, xt, N, -1.0);
      }
    }
  }
  __syncthreads();
  // per-term damping on the camera diagonal
  for (int e = tid; e < n; e += NT) S[pk(e, e)] += nu * double(d.cam_nu[c0 + 1 + e / 6]);
  for (int e = tid; e < n; e += NT) gmax = fmax(gmax, fabs(grad[e]));
  const double max_grad = wg_max<NT>(gmax, red);
  // ---- dense block LDLt in LDS --------------------------------------------------------------
  for (int k = 0; k < nbk; ++k) {
    const int o = 6 * k;
    if (tid == 0) {
      double Dk[36], L6[36], dk[6];
      for (int a = 0; a < 6; ++a)
        for (int b = 0; b < 6; ++b) Dk[6 * a + b] = sym_at(S, o + a, o + b);
      if (!ldlt6(Dk, L6, dk)) {
        flags[1] = 1;
        for (int a = 0; a < 6; ++a) dk[a] = 0.0;
      }
      for (int a = 0; a < 6; ++a) {
        for (int b = 0; b < a; ++b) S[pk(o + a, o + b)] = L6[6 * a + b];
        S[pk(o + a, o + a)] = dk[a];
      }
    }
    __syncthreads();
    // W = S(:, k) below the diagonal block; L(:, k) = W D_k^+
    for (int r = o + 6 + tid; r < n; r += NT) {
      double v[6];
      for (int t = 0; t < 6; ++t) {
        v[t] = S[pk(r, o + t)];
        W[6 * r + t] = v[t];
      }
      d_solve(S, k, v);
      for (int t = 0; t < 6; ++t) S[pk(r, o + t)] = v[t];
    }
    __syncthreads();
    // trailing update S(r, c) -= L(r, k) . W(c, k) for o + 6 <= c <= r
    for (int r = o + 6 + tid; r < n; r += NT) {
      double l[6];
      for (int t = 0; t < 6; ++t) l[t] = S[pk(r, o + t)];
      for (int c = o + 6; c <= r; ++c) {
        const double* w = W + 6 * c;
        S[pk(r, c)] -= l[0] * w[0] + l[1] * w[1] + l[2] * w[2] + l[3] * w[3] + l[4] * w[4] + l[5] * w[5];
      }
    }
    __syncthreads();
  }
  // ---- solves -----------------------------------------------------------------------------
  for (int e = tid; e < n; e += NT) xs[e] = rhs[e];
  __syncthreads();
  for (int k = 0; k < nbk; ++k) {
    const int o = 6 * k;
    for (int r = o + 6 + tid; r < n; r += NT) {
      double s = 0.0;
      for (int t = 0; t < 6; ++t) s += S[pk(r, o + t)] * xs[o + t];
      xs[r] -= s;
    }
    __syncthreads();
  }
  for (int k = tid; k < nbk; k += NT) {
    double v[6];
    for (int t = 0; t < 6; ++t) v[t] = xs[6 * k + t];
    d_solve(S, k, v);
    for (int t = 0; t < 6; ++t) xs[6 * k + t] = v[t];
  }
  __syncthreads();
  for (int k = nbk - 1; k > 0; --k) {
    const int o = 6 * k;
    for (int r = tid; r < o; r += NT) {
      double s = 0.0;
      for (int a = 0; a < 6; ++a) s += S[pk(o + a, r)] * xs[o + a];
      xs[r] -= s;
    }
    __syncthreads();
  }
  // ---- back-substitution, update, proposal merit ----------------------------------------------
  double dd = 0.0, dg = 0.0, dmax = 0.0;
  double* pnew = d.points[nxt] + 3 * int64_t(q0);
  for (int b = tid; b < np; b += NT) {
    const int gb = q0 + b;
    const double x[3] = {points[3 * b], points[3 * b + 1], points[3 * b + 2]};
    double acc[3] = {d.pts[3 * int64_t(gb)], d.pts[3 * int64_t(gb) + 1], d.pts[3 * int64_t(gb) + 2]};
    for (int o = d.pt_obs[gb]; o < d.pt_obs[gb + 1]; ++o) {
      const int ci = d.obs_cam[o] - 1;
      if (ci < 0) continue;
      double r[2];
      T A[6], M[9];
      obs_eval<T>(cams + 15 * (ci + 1), K, d.obs_uv + 2 * int64_t(o), x, r, A);
      gram3(A, M);
      const double* dx = xs + 6 * ci;
      const double q[3] = {dx[1] * x[2] - dx[2] * x[1] + dx[3], dx[2] * x[0] - dx[0] * x[2] + dx[4],
                           dx[0] * x[1] - dx[1] * x[0] + dx[5]};
      for (int k = 0; k < 3; ++k)
        acc[k] += double(M[3 * k]) * q[0] + double(M[3 * k + 1]) * q[1] + double(M[3 * k + 2]) * q[2];
    }
    const double* Hi = d.hinv + 9 * int64_t(gb);
    for (int k = 0; k < 3; ++k) {
      const double dk = Hi[3 * k] * acc[0] + Hi[3 * k + 1] * acc[1] + Hi[3 * k + 2] * acc[2];
      pnew[3 * b + k] = x[k] + dk;
      dd += dk * dk;
      dg += dk * d.pts[3 * int64_t(gb) + k];
      dmax = fmax(dmax, fabs(dk));
    }
  }
  double* tnew = d.poses[nxt] + 7 * int64_t(c0);
  for (int c = tid; c < nc; c += NT) {
    if (c == 0) {
      for (int k = 0; k < 7; ++k) tnew[k] = poses[k];
      continue;
    }
    double t[6], T1[7];
    for (int k = 0; k < 6; ++k) {
      t[k] = xs[6 * (c - 1) + k];
      dd += t[k] * t[k];
      dg += t[k] * grad[6 * (c - 1) + k];
      dmax = fmax(dmax, fabs(t[k]));
    }
    se3_oplus(poses + 7 * c, t, T1);
    for (int k = 0; k < 7; ++k) tnew[7 * c + k] = T1[k];
  }
  __syncthreads();  // proposal poses are read below
  // cameras of the proposal into the cache
  for (int c = tid; c < nc; c += NT) cam_cache(tnew + 7 * c, cams + 15 * c);
  __syncthreads();
  double cost = 0.0;
  for (int o = tid; o < no; o += NT) {
    const int go = o0 + o;
    // point of the observation: observations are sorted by point; find by binary search
    int lo = q0, hi = q0 + np - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) / 2;
      if (d.pt_obs[mid] <= go) lo = mid;
      else hi = mid - 1;
    }
    const double* xp = pnew + 3 * (lo - q0);
    const double x[3] = {xp[0], xp[1], xp[2]};
    double r[2];
    double A[6];
    obs_eval<double>(cams + 15 * d.obs_cam[go], K, d.obs_uv + 2 * int64_t(go), x, r, A);
    const double nrm2 = r[0] * r[0] + r[1] * r[1];
    if (d.robust == 1) {
      const double w = huber_weight(1, d.delta, sqrt(nrm2));
      cost += w * w * nrm2;
    } else {
      cost += nrm2;
    }
  }
  const double new_merit = wg_sum<NT>(cost, red);
  const double dd_s = wg_sum<NT>(dd, red);
  const double dg_s = wg_sum<NT>(dg, red);
  const double max_step = wg_max<NT>(dmax, red);
  if (tid != 0) return;
  // ---- LM decision (step_nlls_inner order, optimizer_step.rs:366-443) -------------------------
  int reason = 0, accepted = 0;
  double nu_next = nu, merit = d.merit[p];
  int rejects = d.rejects[p];
  if (max_grad < d.gradient_tolerance) {
    reason = 3;  // GradientTolerance
  } else if (flags[0] || flags[1]) {
    nu_next = 2.0 * nu;
    if (++rejects >= d.max_consecutive_rejects) reason = 5;
  } else if (max_step < d.parameter_tolerance) {
    if (new_merit < merit) {
      accepted = 1;
      merit = new_merit;
    }
    reason = 4;
  } else {
    const double pred = nu * dd_s + dg_s;
    const double actual = merit - new_merit;
    const double rho = fabs(pred) > DBL_MIN ? actual / pred : (actual > 0 ? 1.0 : 0.0);
    if (rho > 0.0 && new_merit < merit) {
      const double f = 1.0 - pow(2.0 * rho - 1.0, 3.0);
      nu_next = nu * fmax(f, 1.0 / 3.0);
      rejects = 0;
      if ((merit - new_merit) / (fabs(merit) + DBL_MIN) < d.function_tolerance) reason = 2;
      merit = new_merit;
      accepted = 1;
    } else {
      nu_next = 2.0 * nu;
      if (++rejects >= d.max_consecutive_rejects) reason = 5;
    }
  }
  const int steps = d.steps[p] + 1;
  if (steps >= d.num_iterations && reason == 0) reason = 1;
  if (accepted) d.cur[p] = nxt;
  d.nu[p] = nu_next;
  d.merit[p] = merit;
  d.rejects[p] = rejects;
  d.steps[p] = steps;
  if (reason) {
    d.done[p] = 1;
    d.term[p] = reason;
  }
}

// One LM iteration of every unfinished window: one workgroup per window, largest first.
template <typename T, int NT>
__global__ void __launch_bounds__(NT) k_batch_step(BatchDev d) {
  extern __shared__ double lds[];
  __shared__ double red[NT / 64];
  __shared__ int flags[4];
  const int p = d.order[blockIdx.x];
  if (d.done[p]) return;  // uniform per workgroup
  batch_iteration<T, NT>(d, p, lds, red, flags);
}

// Every window's LM run in one launch: each workgroup takes the next window (largest first)
// from a queue counter and runs up to `iterations` LM iterations of it until it terminates,
// then takes the next; the grid is sized to the resident workgroups.  Nothing waits on
// another workgroup, and every loop is bounded (windows, iterations), so the grid always
// drains.  t_start / t_end: the device wall clock when a window was taken and when its run
// ended.
template <typename T, int NT>
__global__ void __launch_bounds__(NT)
    k_batch_persistent(BatchDev d, int32_t iterations, int32_t* __restrict__ queue,
                       unsigned long long* __restrict__ t_start, unsigned long long* __restrict__ t_end) {
  extern __shared__ double lds[];
  __shared__ double red[NT / 64];
  __shared__ int flags[4];
  __shared__ int s_slot, s_done;
  for (;;) {
    if (threadIdx.x == 0) s_slot = atomicAdd(queue, 1);
    __syncthreads();
    const int slot = s_slot;
    if (slot >= d.nprob) return;  // uniform per workgroup
    const int p = d.order[slot];
    if (threadIdx.x == 0) {
      t_start[p] = (unsigned long long)wall_clock64();
      s_done = d.done[p];
    }
    __syncthreads();
    for (int it = 0; it < iterations && !s_done; ++it) {
      // p is uniform, but read from LDS the compiler cannot know it: made a scalar so that
      // the addresses derived from it stay in scalar registers (in vector registers the body
      // takes 284 of them against 165, with spills at 512 threads)
      batch_iteration<T, NT>(d, __builtin_amdgcn_readfirstlane(p), lds, red, flags);
      __syncthreads();
      if (threadIdx.x == 0) s_done = d.done[p];
      __syncthreads();
    }
    if (threadIdx.x == 0) t_end[p] = (unsigned long long)wall_clock64();
    __syncthreads();
  }
}

__global__ void k_batch_zero_queue(int32_t* queue) {
  if (threadIdx.x == 0) *queue = 0;
}

// State from the initial values and the initial merit.
template <int NT>
__global__ void __launch_bounds__(NT) k_batch_init(BatchDev d) {
  __shared__ double red[NT / 64];
  __shared__ double cams[15 * kBatchMaxCams];
  const int p = blockIdx.x;
  const int c0 = d.cam0[p], nc = d.cam0[p + 1] - c0;
  const int q0 = d.pt0[p], np = d.pt0[p + 1] - q0;
  const int o0 = d.obs0[p], no = d.obs0[p + 1] - o0;
  for (int e = threadIdx.x; e < 7 * nc; e += NT) d.poses[0][7 * int64_t(c0) + e] = d.poses_init[7 * int64_t(c0) + e];
  for (int e = threadIdx.x; e < 3 * np; e += NT) d.points[0][3 * int64_t(q0) + e] = d.points_init[3 * int64_t(q0) + e];
  for (int c = threadIdx.x; c < nc; c += NT) cam_cache(d.poses_init + 7 * int64_t(c0 + c), cams + 15 * c);
  __syncthreads();
  const double* K = d.intr + 4 * p;
  double cost = 0.0;
  for (int o = threadIdx.x; o < no; o += NT) {
    const int go = o0 + o;
    int lo = q0, hi = q0 + np - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) / 2;
      if (d.pt_obs[mid] <= go) lo = mid;
      else hi = mid - 1;
    }
    const double* xp = d.points_init + 3 * int64_t(lo);
    const double x[3] = {xp[0], xp[1], xp[2]};
    double r[2], A[6];
    obs_eval<double>(cams + 15 * d.obs_cam[go], K, d.obs_uv + 2 * int64_t(go), x, r, A);
    const double nrm2 = r[0] * r[0] + r[1] * r[1];
    if (d.robust == 1) {
      const double w = huber_weight(1, d.delta, sqrt(nrm2));
      cost += w * w * nrm2;
    } else {
      cost += nrm2;
    }
  }
  const double m = wg_sum<NT>(cost, red);
  if (threadIdx.x == 0) {
    d.cur[p] = 0;
    d.nu[p] = d.nu0;
    d.merit[p] = m;
    d.init_merit[p] = m;
    d.rejects[p] = 0;
    d.steps[p] = 0;
    d.done[p] = d.num_iterations <= 0 ? 1 : 0;
    d.term[p] = d.num_iterations <= 0 ? 1 : 0;
  }
}

}  // namespace
}  // namespace sgpu

using namespace sgpu;

struct sgpu_batch {
  BatchDev d{};
  DeviceArena arena;
  hipStream_t stream = nullptr;
  int precision = 0;
  int threads = 256;           // chosen by the autotuner
  double tune_ms[3] = {0, 0, 0};
  size_t lds_bytes = 0;
  int64_t ncam = 0, npt = 0, nobs = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int graph_iters = 0;
  // the persistent launch: queue counter, per-window wall-clock stamps, its own graph
  int32_t* queue = nullptr;
  unsigned long long* t_start = nullptr;
  unsigned long long* t_end = nullptr;
  int persistent_grid = 0;     // resident workgroups of the persistent kernel's size
  int persistent_threads = 256;  // the persistent kernel's workgroup size (its own autotuner)
  double ptune_ms[3] = {0, 0, 0};
  double clock_khz = 0.0;   // device wall-clock rate
  hipGraph_t pgraph = nullptr;
  hipGraphExec_t pgraph_exec = nullptr;
  int pgraph_iters = -1;

  ~sgpu_batch() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (pgraph_exec) (void)hipGraphExecDestroy(pgraph_exec);
    if (pgraph) (void)hipGraphDestroy(pgraph);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void launch_persistent(int iterations) {
    hipLaunchKernelGGL(k_batch_zero_queue, dim3(1), dim3(64), 0, stream, queue);
    const dim3 g(persistent_grid);
    const int nt = persistent_threads;
    if (precision == SGPU_PRECISION_MIXED) {
      if (nt == 128) hipLaunchKernelGGL((k_batch_persistent<float, 128>), g, dim3(128), lds_bytes, stream, d, iterations, queue, t_start, t_end);
      else if (nt == 512) hipLaunchKernelGGL((k_batch_persistent<float, 512>), g, dim3(512), lds_bytes, stream, d, iterations, queue, t_start, t_end);
      else hipLaunchKernelGGL((k_batch_persistent<float, 256>), g, dim3(256), lds_bytes, stream, d, iterations, queue, t_start, t_end);
    } else {
      if (nt == 128) hipLaunchKernelGGL((k_batch_persistent<double, 128>), g, dim3(128), lds_bytes, stream, d, iterations, queue, t_start, t_end);
      else if (nt == 512) hipLaunchKernelGGL((k_batch_persistent<double, 512>), g, dim3(512), lds_bytes, stream, d, iterations, queue, t_start, t_end);
      else hipLaunchKernelGGL((k_batch_persistent<double, 256>), g, dim3(256), lds_bytes, stream, d, iterations, queue, t_start, t_end);
    }
  }

  void launch_init() {
    hipLaunchKernelGGL((k_batch_init<256>), dim3(d.nprob), dim3(256), 0, stream, d);
  }
  void launch_step(int nt) {
    const dim3 g(d.nprob);
    if (precision == SGPU_PRECISION_MIXED) {
      if (nt == 128) hipLaunchKernelGGL((k_batch_step<float, 128>), g, dim3(128), lds_bytes, stream, d);
      else if (nt == 512) hipLaunchKernelGGL((k_batch_step<float, 512>), g, dim3(512), lds_bytes, stream, d);
      else hipLaunchKernelGGL((k_batch_step<float, 256>), g, dim3(256), lds_bytes, stream, d);
    } else {
      if (nt == 128) hipLaunchKernelGGL((k_batch_step<double, 128>), g, dim3(128), lds_bytes, stream, d);
      else if (nt == 512) hipLaunchKernelGGL((k_batch_step<double, 512>), g, dim3(512), lds_bytes, stream, d);
      else hipLaunchKernelGGL((k_batch_step<double, 256>), g, dim3(256), lds_bytes, stream, d);
    }
  }
};

namespace {
template <typename F>
void set_lds(F kernel, size_t bytes) {
  SGPU_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes)));
}
}  // namespace

extern "C" {

sgpu_status sgpu_batch_create(const sgpu_batch_desc* desc, const sgpu_opt_params* params,
                              sgpu_batch** out) {
  return guarded([&] {
    if (!desc || !params || !out) throw Error(SGPU_ERR_INVALID_ARGUMENT, "desc, params and out required");
    *out = nullptr;
    const int64_t P = desc->num_problems;
    if (P <= 0 || P > INT32_MAX / 2 || !desc->cam_ptr || !desc->point_ptr || !desc->obs_ptr ||
        !desc->intrinsics || !desc->poses || !desc->points || !desc->obs_cam || !desc->obs_point ||
        !desc->obs_uv)
      throw Error(SGPU_ERR_INVALID_ARGUMENT, "incomplete batch description");
    if (params->precision != SGPU_PRECISION_F64 && params->precision != SGPU_PRECISION_MIXED)
      throw Error(SGPU_ERR_INVALID_ARGUMENT, "unknown precision");
    if (params->solver != SGPU_SOLVER_SCHUR_BLOCK_SPARSE_LDLT)
      throw Error(SGPU_ERR_UNSUPPORTED, "the batched path solves each window directly");
    const int64_t C = desc->cam_ptr[P], Q = desc->point_ptr[P], O = desc->obs_ptr[P];
    if (desc->cam_ptr[0] != 0 || desc->point_ptr[0] != 0 || desc->obs_ptr[0] != 0 || O > INT32_MAX ||
        Q > INT32_MAX || C > INT32_MAX)
      throw Error(SGPU_ERR_INVALID_ARGUMENT, "bad prefix arrays");
    int n_max = 0;
    std::vector<int32_t> cam0(P + 1), pt0(P + 1), obs0(P + 1), obs_cam(O), pt_obs(Q + 1), pt_nu(Q, 0),
        cam_nu(C, 0);
    std::vector<double> uv(2 * O);
    for (int64_t p = 0; p <= P; ++p) {
      cam0[p] = int32_t(desc->cam_ptr[p]);
      pt0[p] = int32_t(desc->point_ptr[p]);
      obs0[p] = int32_t(desc->obs_ptr[p]);
    }
    for (int64_t p = 0; p < P; ++p) {
      const int64_t nc = cam0[p + 1] - cam0[p], np = pt0[p + 1] - pt0[p], no = obs0[p + 1] - obs0[p];
      if (nc < 2 || nc > kBatchMaxCams || np < 1 || no < 1)
        throw Error(nc > kBatchMaxCams ? SGPU_ERR_UNSUPPORTED : SGPU_ERR_INVALID_ARGUMENT,
                    "each window needs 2.." + std::to_string(kBatchMaxCams) +
                        " cameras, points and observations");
      n_max = std::max<int>(n_max, int(6 * (nc - 1)));
      // observations of the window sorted by (point, camera) (CostFn::sort order)
      std::vector<int32_t> ord(no);
      std::iota(ord.begin(), ord.end(), 0);
      const int64_t ob = obs0[p];
      for (int64_t o = ob; o < ob + no; ++o) {
        if (desc->obs_cam[o] < 0 || desc->obs_cam[o] >= nc || desc->obs_point[o] < 0 || desc->obs_point[o] >= np)
          throw Error(SGPU_ERR_INDEX_OUT_OF_RANGE, "observation index out of range");
      }
      std::stable_sort(ord.begin(), ord.end(), [&](int32_t a, int32_t b) {
        const int64_t pa = desc->obs_point[ob + a], pb = desc->obs_point[ob + b];
        if (pa != pb) return pa < pb;
        return desc->obs_cam[ob + a] < desc->obs_cam[ob + b];
      });
      std::vector<int32_t> cnt(np + 1, 0);
      for (int64_t k = 0; k < no; ++k) {
        const int64_t o = ob + ord[k];
        obs_cam[ob + k] = desc->obs_cam[o];
        uv[2 * (ob + k)] = desc->obs_uv[2 * o];
        uv[2 * (ob + k) + 1] = desc->obs_uv[2 * o + 1];
        ++cnt[desc->obs_point[o] + 1];
      }
      for (int64_t b = 0; b < np; ++b) cnt[b + 1] += cnt[b];
      for (int64_t b = 0; b <= np; ++b) pt_obs[pt0[p] + b] = int32_t(ob + cnt[b]);
      // reduced terms: distinct (camera, point) pairs
      for (int64_t b = 0; b < np; ++b) {
        int32_t last = -1;
        for (int32_t o = pt_obs[pt0[p] + b]; o < pt_obs[pt0[p] + b + 1]; ++o) {
          if (obs_cam[o] == last) continue;
          last = obs_cam[o];
          ++pt_nu[pt0[p] + b];
          ++cam_nu[cam0[p] + last];
        }
      }
    }
    const int nt_max = 512;
    const size_t lds_doubles = size_t(pk(n_max, 0)) + 6 * size_t(n_max) + 3 * size_t(n_max) +
                               15 * size_t(kBatchMaxCams);
    const size_t lds_bytes = lds_doubles * 8;
    if (lds_bytes + 8 * (nt_max / 64) + 16 + 8 > 160 * 1024)  // + red, flags, queue slot
      throw Error(SGPU_ERR_UNSUPPORTED, "window too large for the LDS-resident solve");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw Error(SGPU_ERR_NO_DEVICE, "no HIP device");
    if (params->device < 0 || params->device >= ndev) throw Error(SGPU_ERR_NO_DEVICE, "device out of range");
    SGPU_HIP(hipSetDevice(params->device));
    auto* b = new sgpu_batch();
    try {
      b->precision = params->precision;
      b->lds_bytes = lds_bytes;
      b->ncam = C;
      b->npt = Q;
      b->nobs = O;
      SGPU_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
      // launch order: largest windows first (load balance)
      std::vector<int32_t> order(P);
      std::iota(order.begin(), order.end(), 0);
      std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t c) {
        return obs0[a + 1] - obs0[a] > obs0[c + 1] - obs0[c];
      });
      BatchDev& d = b->d;
      DeviceArena& A = b->arena;
      d.nprob = int32_t(P);
      d.order = A.upload(order, nullptr);
      d.cam0 = A.upload(cam0, nullptr);
      d.pt0 = A.upload(pt0, nullptr);
      d.obs0 = A.upload(obs0, nullptr);
      d.intr = A.upload(std::vector<double>(desc->intrinsics, desc->intrinsics + 4 * P), nullptr);
      d.poses_init = A.upload(std::vector<double>(desc->poses, desc->poses + 7 * C), nullptr);
      d.points_init = A.upload(std::vector<double>(desc->points, desc->points + 3 * Q), nullptr);
      for (int i = 0; i < 2; ++i) {
        d.poses[i] = A.alloc<double>(size_t(7) * C);
        d.points[i] = A.alloc<double>(size_t(3) * Q);
      }
      d.obs_cam = A.upload(obs_cam, nullptr);
      d.obs_uv = A.upload(uv, nullptr);
      d.pt_obs = A.upload(pt_obs, nullptr);
      d.pt_nu = A.upload(pt_nu, nullptr);
      d.cam_nu = A.upload(cam_nu, nullptr);
      d.hinv = A.alloc<double>(size_t(9) * Q);
      d.pts = A.alloc<double>(size_t(3) * Q);
      d.cur = A.alloc<int32_t>(P);
      d.nu = A.alloc<double>(P);
      d.merit = A.alloc<double>(P);
      d.init_merit = A.alloc<double>(P);
      d.rejects = A.alloc<int32_t>(P);
      d.steps = A.alloc<int32_t>(P);
      d.done = A.alloc<int32_t>(P);
      d.term = A.alloc<int32_t>(P);
      d.robust = desc->huber_delta > 0.0 ? 1 : 0;
      d.delta = desc->huber_delta;
      d.nu0 = params->initial_lm_damping;
      d.num_iterations = params->num_iterations;
      d.function_tolerance = params->function_tolerance;
      d.gradient_tolerance = params->gradient_tolerance;
      d.parameter_tolerance = params->parameter_tolerance;
      d.max_consecutive_rejects = params->max_consecutive_rejects;
      d.lds_doubles = int32_t(lds_doubles);
      d.n_max = n_max;
      set_lds(k_batch_step<double, 128>, lds_bytes);
      set_lds(k_batch_step<double, 256>, lds_bytes);
      set_lds(k_batch_step<double, 512>, lds_bytes);
      set_lds(k_batch_step<float, 128>, lds_bytes);
      set_lds(k_batch_step<float, 256>, lds_bytes);
      set_lds(k_batch_step<float, 512>, lds_bytes);
      set_lds(k_batch_persistent<double, 128>, lds_bytes);
      set_lds(k_batch_persistent<double, 256>, lds_bytes);
      set_lds(k_batch_persistent<double, 512>, lds_bytes);
      set_lds(k_batch_persistent<float, 128>, lds_bytes);
      set_lds(k_batch_persistent<float, 256>, lds_bytes);
      set_lds(k_batch_persistent<float, 512>, lds_bytes);
      b->queue = A.alloc<int32_t>(1);
      b->t_start = A.alloc<unsigned long long>(P);
      b->t_end = A.alloc<unsigned long long>(P);
      int khz = 0;
      SGPU_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, params->device));
      if (khz <= 0) throw Error(SGPU_ERR_HIP, "device wall-clock rate unavailable");
      b->clock_khz = double(khz);
      // autotuner: one LM iteration of the whole batch per candidate workgroup size, from
      // the same initial state; the fastest is kept
      hipEvent_t e0, e1;
      SGPU_HIP(hipEventCreate(&e0));
      SGPU_HIP(hipEventCreate(&e1));
      const int cand[3] = {128, 256, 512};
      float best = 1e30f;
      for (int i = 0; i < 3; ++i) {  // first launches of each variant, untimed
        b->launch_init();
        b->launch_step(cand[i]);
      }
      for (int i = 0; i < 3; ++i) {
        b->launch_init();
        SGPU_HIP(hipEventRecord(e0, b->stream));
        b->launch_step(cand[i]);
        SGPU_HIP(hipEventRecord(e1, b->stream));
        SGPU_HIP(hipGetLastError());
        SGPU_HIP(hipEventSynchronize(e1));
        float ms = 0;
        SGPU_HIP(hipEventElapsedTime(&ms, e0, e1));
        b->tune_ms[i] = ms;
        if (ms < best) {
          best = ms;
          b->threads = cand[i];
        }
      }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      // the persistent kernel keeps more registers live than the per-iteration one, so it
      // has its own choice: for each size, the grid of every resident workgroup and one
      // timed persistent launch of one LM iteration per window; the fastest is kept
      {
        int cus = 0;
        SGPU_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, params->device));
        int grids[3] = {0, 0, 0};
        for (int i = 0; i < 3; ++i) {
          const int nt = cand[i];
          const void* fn = nullptr;
          if (b->precision == SGPU_PRECISION_MIXED)
            fn = nt == 128 ? reinterpret_cast<const void*>(k_batch_persistent<float, 128>)
                 : nt == 512 ? reinterpret_cast<const void*>(k_batch_persistent<float, 512>)
                             : reinterpret_cast<const void*>(k_batch_persistent<float, 256>);
          else
            fn = nt == 128 ? reinterpret_cast<const void*>(k_batch_persistent<double, 128>)
                 : nt == 512 ? reinterpret_cast<const void*>(k_batch_persistent<double, 512>)
                             : reinterpret_cast<const void*>(k_batch_persistent<double, 256>);
          int per_cu = 0;
          SGPU_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, nt, lds_bytes));
          grids[i] = int(std::max<int64_t>(1, std::min<int64_t>(P, int64_t(std::max(per_cu, 1)) * cus)));
        }
        hipEvent_t p0, p1;
        SGPU_HIP(hipEventCreate(&p0));
        SGPU_HIP(hipEventCreate(&p1));
        float pbest = 1e30f;
        for (int rep = 0; rep < 2; ++rep)  // the first pass is untimed
          for (int i = 0; i < 3; ++i) {
            b->persistent_threads = cand[i];
            b->persistent_grid = grids[i];
            b->launch_init();
            SGPU_HIP(hipEventRecord(p0, b->stream));
            b->launch_persistent(1);
            SGPU_HIP(hipEventRecord(p1, b->stream));
            SGPU_HIP(hipGetLastError());
            SGPU_HIP(hipEventSynchronize(p1));
            float ms = 0;
            SGPU_HIP(hipEventElapsedTime(&ms, p0, p1));
            if (rep == 1) b->ptune_ms[i] = ms;
          }
        int pick = 0;
        for (int i = 0; i < 3; ++i)
          if (b->ptune_ms[i] < pbest) {
            pbest = float(b->ptune_ms[i]);
            pick = i;
          }
        b->persistent_threads = cand[pick];
        b->persistent_grid = grids[pick];
        (void)hipEventDestroy(p0);
        (void)hipEventDestroy(p1);
      }
      b->launch_init();
      SGPU_HIP(hipStreamSynchronize(b->stream));
    } catch (...) {
      delete b;
      throw;
    }
    *out = b;
  });
}

sgpu_status sgpu_batch_run(sgpu_batch* b, int32_t iterations, int32_t use_graph, double* seconds) {
  return guarded([&] {
    if (!b || iterations < 0) throw Error(SGPU_ERR_INVALID_ARGUMENT, "bad arguments");
    const auto t0 = std::chrono::steady_clock::now();
    if (use_graph) {
      if (!b->graph_exec || b->graph_iters != iterations) {
        if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->graph_exec = nullptr;
        b->graph = nullptr;
        SGPU_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < iterations; ++i) b->launch_step(b->threads);
        SGPU_HIP(hipStreamEndCapture(b->stream, &b->graph));
        SGPU_HIP(hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
        b->graph_iters = iterations;
      }
      SGPU_HIP(hipGraphLaunch(b->graph_exec, b->stream));
    } else {
      for (int i = 0; i < iterations; ++i) b->launch_step(b->threads);
    }
    SGPU_HIP(hipGetLastError());
    SGPU_HIP(hipStreamSynchronize(b->stream));
    if (seconds)
      *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  });
}

sgpu_status sgpu_batch_run_persistent(sgpu_batch* b, int32_t iterations, int32_t use_graph,
                                      double* seconds) {
  return guarded([&] {
    if (!b || iterations < 0) throw Error(SGPU_ERR_INVALID_ARGUMENT, "bad arguments");
    const auto t0 = std::chrono::steady_clock::now();
    if (use_graph) {
      if (!b->pgraph_exec || b->pgraph_iters != iterations) {
        if (b->pgraph_exec) (void)hipGraphExecDestroy(b->pgraph_exec);
        if (b->pgraph) (void)hipGraphDestroy(b->pgraph);
        b->pgraph_exec = nullptr;
        b->pgraph = nullptr;
        SGPU_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        b->launch_persistent(iterations);
        SGPU_HIP(hipStreamEndCapture(b->stream, &b->pgraph));
        SGPU_HIP(hipGraphInstantiate(&b->pgraph_exec, b->pgraph, nullptr, nullptr, 0));
        b->pgraph_iters = iterations;
      }
      SGPU_HIP(hipGraphLaunch(b->pgraph_exec, b->stream));
    } else {
      b->launch_persistent(iterations);
    }
    SGPU_HIP(hipGetLastError());
    SGPU_HIP(hipStreamSynchronize(b->stream));
    if (seconds)
      *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  });
}

sgpu_status sgpu_batch_latencies(const sgpu_batch* b, double* start_ms, double* end_ms) {
  return guarded([&] {
    if (!b || !start_ms || !end_ms) throw Error(SGPU_ERR_INVALID_ARGUMENT, "bad arguments");
    const size_t P = size_t(b->d.nprob);
    std::vector<unsigned long long> ts(P), te(P);
    SGPU_HIP(hipMemcpy(ts.data(), b->t_start, P * 8, hipMemcpyDeviceToHost));
    SGPU_HIP(hipMemcpy(te.data(), b->t_end, P * 8, hipMemcpyDeviceToHost));
    const unsigned long long t0 = *std::min_element(ts.begin(), ts.end());
    for (size_t p = 0; p < P; ++p) {
      start_ms[p] = double(ts[p] - t0) / b->clock_khz;
      end_ms[p] = double(te[p] - t0) / b->clock_khz;
    }
  });
}

sgpu_status sgpu_batch_reset(sgpu_batch* b) {
  return guarded([&] {
    if (!b) throw Error(SGPU_ERR_INVALID_ARGUMENT, "null batch");
    b->launch_init();
    SGPU_HIP(hipStreamSynchronize(b->stream));
  });
}

sgpu_status sgpu_batch_results(sgpu_batch* b, double* poses_out, double* points_out, double* initial_cost,
                               double* final_cost, int32_t* num_steps, int32_t* termination) {
  return guarded([&] {
    if (!b) throw Error(SGPU_ERR_INVALID_ARGUMENT, "null batch");
    const BatchDev& d = b->d;
    const size_t P = size_t(d.nprob);
    std::vector<int32_t> cur(P);
    SGPU_HIP(hipMemcpy(cur.data(), d.cur, P * 4, hipMemcpyDeviceToHost));
    if (poses_out || points_out) {
      std::vector<double> p0(7 * b->ncam), p1(7 * b->ncam), x0(3 * b->npt), x1(3 * b->npt);
      std::vector<int32_t> cam0(P + 1), pt0(P + 1);
      SGPU_HIP(hipMemcpy(cam0.data(), d.cam0, (P + 1) * 4, hipMemcpyDeviceToHost));
      SGPU_HIP(hipMemcpy(pt0.data(), d.pt0, (P + 1) * 4, hipMemcpyDeviceToHost));
      SGPU_HIP(hipMemcpy(p0.data(), d.poses[0], p0.size() * 8, hipMemcpyDeviceToHost));
      SGPU_HIP(hipMemcpy(p1.data(), d.poses[1], p1.size() * 8, hipMemcpyDeviceToHost));
      SGPU_HIP(hipMemcpy(x0.data(), d.points[0], x0.size() * 8, hipMemcpyDeviceToHost));
      SGPU_HIP(hipMemcpy(x1.data(), d.points[1], x1.size() * 8, hipMemcpyDeviceToHost));
      for (size_t p = 0; p < P; ++p) {
        const auto& ps = cur[p] ? p1 : p0;
        const auto& xs = cur[p] ? x1 : x0;
        if (poses_out)
          std::copy(ps.begin() + 7 * cam0[p], ps.begin() + 7 * cam0[p + 1], poses_out + 7 * cam0[p]);
        if (points_out)
          std::copy(xs.begin() + 3 * pt0[p], xs.begin() + 3 * pt0[p + 1], points_out + 3 * pt0[p]);
      }
    }
    if (initial_cost) SGPU_HIP(hipMemcpy(initial_cost, d.init_merit, P * 8, hipMemcpyDeviceToHost));
    if (final_cost) SGPU_HIP(hipMemcpy(final_cost, d.merit, P * 8, hipMemcpyDeviceToHost));
    if (num_steps) SGPU_HIP(hipMemcpy(num_steps, d.steps, P * 4, hipMemcpyDeviceToHost));
    if (termination) SGPU_HIP(hipMemcpy(termination, d.term, P * 4, hipMemcpyDeviceToHost));
  });
}

sgpu_status sgpu_batch_info(const sgpu_batch* b, int32_t* threads, double* tune_ms3, int64_t* lds_bytes,
                            int32_t* persistent_threads, double* persistent_tune_ms3,
                            int32_t* persistent_grid) {
  return guarded([&] {
    if (!b) throw Error(SGPU_ERR_INVALID_ARGUMENT, "null batch");
    if (threads) *threads = b->threads;
    if (persistent_threads) *persistent_threads = b->persistent_threads;
    if (persistent_tune_ms3)
      for (int i = 0; i < 3; ++i) persistent_tune_ms3[i] = b->ptune_ms[i];
    if (persistent_grid) *persistent_grid = b->persistent_grid;
    if (tune_ms3)
      for (int i = 0; i < 3; ++i) tune_ms3[i] = b->tune_ms[i];
    if (lds_bytes) *lds_bytes = int64_t(b->lds_bytes);
  });
}

void sgpu_batch_destroy(sgpu_batch* b) { delete b; }

}  // extern "C"
