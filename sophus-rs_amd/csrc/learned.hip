// Learned block-Jacobi correction: features, MLP forward/backward, loss, Adam, and the
// preconditioner blocks (see learned.hpp for the model).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "api_internal.hpp"
#include "learned.hpp"

namespace sgpu {
namespace {

constexpr int kT = 256;
constexpr int kRows = kT / 64;  // one wavefront per block row

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

// D^-1/2-scaled Cholesky whitening of a symmetric 6x6 block: B = J L^-T with
// J = diag(D)^-1/2, J D J = L L^T.  Returns false if D is not positive definite.
__device__ bool whiten6(const double* __restrict__ D, double* __restrict__ B) {
  double J[6], L[36], Li[36];
  for (int i = 0; i < 6; ++i) {
    const double d = D[7 * i];
    if (!(d > 0.0) || !isfinite(d)) return false;
    J[i] = 1.0 / sqrt(d);
  }
  for (int j = 0; j < 6; ++j) {
    double s = D[7 * j] * J[j] * J[j];
    for (int k = 0; k < j; ++k) s -= L[6 * j + k] * L[6 * j + k];
    if (!(s > 1e-14)) return false;
    const double ljj = sqrt(s);
    L[6 * j + j] = ljj;
    for (int i = j + 1; i < 6; ++i) {
      double t = D[6 * i + j] * J[i] * J[j];  // lower triangle of D
      for (int k = 0; k < j; ++k) t -= L[6 * i + k] * L[6 * j + k];
      L[6 * i + j] = t / ljj;
    }
    for (int i = 0; i < j; ++i) L[6 * i + j] = 0.0;
  }
  // Li = L^-1 (lower)
  for (int c = 0; c < 6; ++c)
    for (int i = 0; i < 6; ++i) {
      if (i < c) {
        Li[6 * i + c] = 0.0;
        continue;
      }
      double s = i == c ? 1.0 : 0.0;
      for (int k = c; k < i; ++k) s -= L[6 * i + k] * Li[6 * k + c];
      Li[6 * i + c] = s / L[6 * i + i];
    }
  // B = J L^-T : B[i][j] = J_i Li[j][i]
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) B[6 * i + j] = J[i] * Li[6 * j + i];
  return true;
}

// G from the 21 outputs (row-major lower triangle, diagonal exp(o)).
__device__ __forceinline__ void make_g(const float* __restrict__ o, double G[36]) {
  int idx = 0;
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      if (j > i) {
        G[6 * i + j] = 0.0;
      } else {
        const double v = double(o[idx++]);
        G[6 * i + j] = i == j ? exp(v) : v;
      }
    }
}

// M = B G G^T B^T
__device__ void assemble_minv(const double* __restrict__ B, const double G[36], double* __restrict__ M) {
  double BG[36];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      double s = 0.0;
      for (int k = j; k < 6; ++k) s += B[6 * i + k] * G[6 * k + j];
      BG[6 * i + j] = s;
    }
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      double s = 0.0;
      for (int k = 0; k < 6; ++k) s += BG[6 * i + k] * BG[6 * j + k];
      M[6 * i + j] = s;
    }
}

__device__ __forceinline__ void load_diag_block(const DevPcg& g, int b, double D[36]) {
  const double* S = g.bsr + 36 * int64_t(g.diag_slot[b]);
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j <= i; ++j) {
      D[6 * i + j] = S[6 * i + j];
      D[6 * j + i] = S[6 * i + j];
    }
}

// ---- features ------------------------------------------------------------------------
// Per block: the 15 correlations of J D J above the diagonal, log(d_i) - mean log(d),
// log(1 + neighbours) / 4, and log1p of the sum, max and sum of squares of the scaled
// off-diagonal block norms |J_b S_bc J_c|_F.  Entries 25..31 are zero.
__global__ void __launch_bounds__(kT) k_features(DevPcg g, float* __restrict__ feat) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * kRows + (threadIdx.x >> 6);
  if (row >= g.nb) return;  // whole wavefront leaves together
  double Jr[6];
  {
    const double* S = g.bsr + 36 * int64_t(g.diag_slot[row]);
    for (int i = 0; i < 6; ++i) Jr[i] = 1.0 / sqrt(fmax(S[7 * i], 1e-300));
  }
  const int start = g.mv_ptr[row], len = g.mv_ptr[row + 1] - start;
  const double* Brow = g.fbsr + 36 * int64_t(start);
  double csum = 0.0, cmax = 0.0, csq = 0.0;
  for (int j = lane; j < len; j += 64) {
    const int col = g.fcol[start + j];
    if (col == row) continue;
    const double* Sc = g.bsr + 36 * int64_t(g.diag_slot[col]);
    double Jc[6];
    for (int k = 0; k < 6; ++k) Jc[k] = 1.0 / sqrt(fmax(Sc[7 * k], 1e-300));
    double f2 = 0.0;
    for (int a = 0; a < 6; ++a)
      for (int k = 0; k < 6; ++k) {
        const double v = Jr[a] * Brow[int64_t(6 * a + k) * len + j] * Jc[k];
        f2 += v * v;
      }
    const double c = sqrt(f2);
    csum += c;
    cmax = fmax(cmax, c);
    csq += f2;
  }
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) {
    csum += __shfl_xor(csum, m, 64);
    csq += __shfl_xor(csq, m, 64);
    cmax = fmax(cmax, __shfl_xor(cmax, m, 64));
  }
  if (lane != 0) return;
  double D[36];
  load_diag_block(g, row, D);
  float* f = feat + int64_t(kPcFeat) * row;
  int idx = 0;
  for (int i = 0; i < 6; ++i)
    for (int j = i + 1; j < 6; ++j) f[idx++] = float(D[6 * j + i] * Jr[i] * Jr[j]);
  double ld[6], mean = 0.0;
  for (int i = 0; i < 6; ++i) {
    ld[i] = log(fmax(D[7 * i], 1e-300));
    mean += ld[i] / 6.0;
  }
  for (int i = 0; i < 6; ++i) f[idx++] = float(ld[i] - mean);
  f[idx++] = float(log1p(double(len - 1)) / 4.0);
  f[idx++] = float(log1p(csum));
  f[idx++] = float(log1p(cmax));
  f[idx++] = float(log1p(csq));
  while (idx < kPcFeat) f[idx++] = 0.0f;
}

// ---- MLP -----------------------------------------------------------------------------
// Y[n x M] = act(X[n x K] W^T + b), one thread per output; FP32 accumulation.
template <bool kBf16, bool kTanh>
__global__ void __launch_bounds__(kT)
    k_dense(int64_t n, int K, int M, const float* __restrict__ X, const float* __restrict__ W32,
            const uint16_t* __restrict__ W16, int off_w, int off_b, float* __restrict__ Y) {
  const int64_t t = int64_t(blockIdx.x) * kT + threadIdx.x;
  if (t >= n * M) return;
  const int64_t b = t / M;
  const int m = int(t - b * M);
  const float* x = X + b * K;
  float acc = kBf16 ? bf16_to_f32(W16[off_b + m]) : W32[off_b + m];
  for (int k = 0; k < K; ++k) {
    const float w = kBf16 ? bf16_to_f32(W16[off_w + m * K + k]) : W32[off_w + m * K + k];
    acc += w * x[k];
  }
  Y[t] = kTanh ? tanhf(acc) : acc;
}

// dA[b][m] = (sum_i dAn[b][i] Wn[i][m]) (1 - H[b][m]^2)
__global__ void __launch_bounds__(kT)
    k_back_hidden(int64_t n, int Mn, int M, const float* __restrict__ dAn, const float* __restrict__ W32,
                  int off_wn, const float* __restrict__ H, float* __restrict__ dA) {
  const int64_t t = int64_t(blockIdx.x) * kT + threadIdx.x;
  if (t >= n * M) return;
  const int64_t b = t / M;
  const int m = int(t - b * M);
  float s = 0.0f;
  for (int i = 0; i < Mn; ++i) s += dAn[b * Mn + i] * W32[off_wn + i * M + m];
  const float h = H[t];
  dA[t] = s * (1.0f - h * h);
}

// grad W[m][k] = sum_b dA[b][m] X[b][k]  (k == K: the bias), fixed order over b.
__global__ void __launch_bounds__(kT)
    k_weight_grad(int64_t n, int M, int K, const float* __restrict__ dA, const float* __restrict__ X,
                  int off_w, int off_b, float* __restrict__ grad) {
  const int t = blockIdx.x * kT + threadIdx.x;
  if (t >= M * (K + 1)) return;
  const int m = t / (K + 1), k = t - m * (K + 1);
  float s = 0.0f;
  if (k < K) {
    for (int64_t b = 0; b < n; ++b) s += dA[b * M + m] * X[b * K + k];
    grad[off_w + m * K + k] = s;
  } else {
    for (int64_t b = 0; b < n; ++b) s += dA[b * M + m];
    grad[off_b + m] = s;
  }
}

__global__ void __launch_bounds__(kT)
    k_adam(int n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
           float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2,
           float scale) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] * scale;
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---- preconditioner blocks and the training loss ----------------------------------------
__global__ void __launch_bounds__(kT) k_learned_minv(DevPcg g, const float* __restrict__ out) {
  const int b = blockIdx.x * kT + threadIdx.x;
  if (b >= g.nb) return;
  double D[36], B[36], G[36], M[36];
  load_diag_block(g, b, D);
  if (!whiten6(D, B)) return;  // keeps block-Jacobi
  make_g(out + int64_t(kPcOut) * b, G);
  assemble_minv(B, G, M);
  for (int e = 0; e < 36; ++e) g.minv[36 * int64_t(b) + e] = M[e];
}

__global__ void __launch_bounds__(kT)
    k_blocks_from_diag(int64_t n, const double* __restrict__ Dg, const float* __restrict__ out,
                       double* __restrict__ minv) {
  const int64_t b = int64_t(blockIdx.x) * kT + threadIdx.x;
  if (b >= n) return;
  double D[36], B[36], G[36], M[36];
  for (int e = 0; e < 36; ++e) D[e] = Dg[36 * b + e];
  if (!whiten6(D, B)) {
    for (int e = 0; e < 36; ++e) minv[36 * b + e] = NAN;
    return;
  }
  make_g(out + int64_t(kPcOut) * b, G);
  assemble_minv(B, G, M);
  for (int e = 0; e < 36; ++e) minv[36 * b + e] = M[e];
}

// Whitening of the diagonal blocks of a snapshot (ok = 0: D not positive definite).
__global__ void __launch_bounds__(kT) k_whiten(DevPcg g, double* __restrict__ Bm, int32_t* __restrict__ ok) {
  const int b = blockIdx.x * kT + threadIdx.x;
  if (b >= g.nb) return;
  double D[36], B[36];
  load_diag_block(g, b, D);
  const bool good = whiten6(D, B);
  ok[b] = good ? 1 : 0;
  for (int e = 0; e < 36; ++e) Bm[36 * int64_t(b) + e] = good ? B[e] : 0.0;
}

__global__ void __launch_bounds__(kT)
    k_whiten_diag(int64_t n, const double* __restrict__ Dg, double* __restrict__ Bm, int32_t* __restrict__ ok) {
  const int64_t b = int64_t(blockIdx.x) * kT + threadIdx.x;
  if (b >= n) return;
  double D[36], B[36];
  for (int e = 0; e < 36; ++e) D[e] = Dg[36 * b + e];
  const bool good = whiten6(D, B);
  ok[b] = good ? 1 : 0;
  for (int e = 0; e < 36; ++e) Bm[36 * b + e] = good ? B[e] : 0.0;
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Probe w ~ N(0, I) per block (hashed from seed and the block index), v_b = B_b w_b.
__global__ void __launch_bounds__(kT)
    k_probe(int64_t n, const double* __restrict__ Bm, uint64_t seed, double* __restrict__ w,
            double* __restrict__ v) {
  const int64_t b = int64_t(blockIdx.x) * kT + threadIdx.x;
  if (b >= n) return;
  double z[6];
  for (int k = 0; k < 6; k += 2) {
    const uint64_t h1 = mix64(seed ^ mix64(uint64_t(6 * b + k)));
    const uint64_t h2 = mix64(h1 ^ 0xD1B54A32D192ED03ull);
    const double u1 = (double(h1 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    const double u2 = double(h2 >> 11) * (1.0 / 9007199254740992.0);
    const double r = sqrt(-2.0 * log(u1));
    z[k] = r * cos(6.283185307179586 * u2);
    z[k + 1] = r * sin(6.283185307179586 * u2);
  }
  for (int i = 0; i < 6; ++i) {
    w[6 * b + i] = z[i];
    double s = 0.0;
    for (int k = 0; k < 6; ++k) s += Bm[36 * b + 6 * i + k] * z[k];
    v[6 * b + i] = s;
  }
}

// u = S v over the full block rows (one wavefront per row).
__global__ void __launch_bounds__(kT) k_rcs_matvec(DevPcg g, const double* __restrict__ v, double* __restrict__ u) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * kRows + (threadIdx.x >> 6);
  if (row >= g.nb) return;
  const int start = g.mv_ptr[row], len = g.mv_ptr[row + 1] - start;
  const double* Brow = g.fbsr + 36 * int64_t(start);
  double y[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int j = lane; j < len; j += 64) {
    const int col = g.fcol[start + j];
    for (int a = 0; a < 6; ++a)
      for (int k = 0; k < 6; ++k) y[a] += Brow[int64_t(6 * a + k) * len + j] * v[6 * col + k];
  }
#pragma unroll
  for (int m = 32; m > 0; m >>= 1)
#pragma unroll
    for (int a = 0; a < 6; ++a) y[a] += __shfl_xor(y[a], m, 64);
  if (lane == 0)
    for (int a = 0; a < 6; ++a) u[6 * row + a] = y[a];
}

// u~_b = B_b^T u_b
__global__ void __launch_bounds__(kT)
    k_utilde(int64_t n, const double* __restrict__ Bm, const double* __restrict__ u, double* __restrict__ ut) {
  const int64_t b = int64_t(blockIdx.x) * kT + threadIdx.x;
  if (b >= n) return;
  for (int i = 0; i < 6; ++i) {
    double s = 0.0;
    for (int k = 0; k < 6; ++k) s += Bm[36 * b + 6 * k + i] * u[6 * b + k];
    ut[6 * b + i] = s;
  }
}

// loss_b = |G G^T u~ - w|^2 / |w|^2 and dloss/do (added to dO when `accumulate`).
//   dG = 2/|w|^2 (f (G^T u~)^T + u~ (G^T f)^T), f = G G^T u~ - w; do_ii = dG_ii G_ii.
__global__ void __launch_bounds__(kT)
    k_loss_grad(int64_t n, const int32_t* __restrict__ ok, const double* __restrict__ ut,
                const double* __restrict__ w, const float* __restrict__ out, float* __restrict__ dO,
                double* __restrict__ loss, int accumulate) {
  const int64_t b = int64_t(blockIdx.x) * kT + threadIdx.x;
  if (b >= n) return;
  float* d = dO + int64_t(kPcOut) * b;
  if (!ok[b]) {
    if (!accumulate)
      for (int k = 0; k < kPcOut; ++k) d[k] = 0.0f;
    loss[b] = accumulate ? loss[b] : 0.0;
    return;
  }
  double G[36], t[6], f[6], s[6];
  make_g(out + int64_t(kPcOut) * b, G);
  const double* u = ut + 6 * b;
  const double* wb = w + 6 * b;
  double ww = 0.0;
  for (int i = 0; i < 6; ++i) ww += wb[i] * wb[i];
  for (int i = 0; i < 6; ++i) {  // t = G^T u
    double a = 0.0;
    for (int k = i; k < 6; ++k) a += G[6 * k + i] * u[k];
    t[i] = a;
  }
  double l = 0.0;
  for (int i = 0; i < 6; ++i) {  // f = G t - w
    double a = 0.0;
    for (int k = 0; k <= i; ++k) a += G[6 * i + k] * t[k];
    f[i] = a - wb[i];
    l += f[i] * f[i];
  }
  for (int i = 0; i < 6; ++i) {  // s = G^T f
    double a = 0.0;
    for (int k = i; k < 6; ++k) a += G[6 * k + i] * f[k];
    s[i] = a;
  }
  const double c = 2.0 / ww;
  int idx = 0;
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j <= i; ++j) {
      double gij = c * (f[i] * t[j] + u[i] * s[j]);
      if (i == j) gij *= G[7 * i];
      d[idx] = accumulate ? d[idx] + float(gij) : float(gij);
      ++idx;
    }
  loss[b] = (accumulate ? loss[b] : 0.0) + l / ww;
}

inline unsigned grid_for(int64_t n) { return unsigned((n + kT - 1) / kT); }

}  // namespace

std::vector<uint16_t> weights_to_bf16(const float* w, int64_t n) {
  std::vector<uint16_t> out(size_t(std::max<int64_t>(n, 0)));
  for (int64_t i = 0; i < n; ++i) {
    uint32_t u;
    std::memcpy(&u, &w[i], 4);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x7FFFFFu)) {
      out[i] = uint16_t((u >> 16) | 0x40u);  // quiet NaN
    } else {
      const uint32_t rounding = 0x7FFFu + ((u >> 16) & 1u);
      out[i] = uint16_t((u + rounding) >> 16);
    }
  }
  return out;
}

void launch_precond_features(hipStream_t s, const DevPcg& g, float* feat) {
  hipLaunchKernelGGL(k_features, dim3((g.nb + kRows - 1) / kRows), dim3(kT), 0, s, g, feat);
}

void launch_mlp_forward(hipStream_t s, int64_t n, const float* feat, const float* w32,
                        const uint16_t* wbf16, float* h1, float* h2, float* out) {
  if (n <= 0) return;
  if (wbf16) {
    hipLaunchKernelGGL((k_dense<true, true>), dim3(grid_for(n * kPcHid)), dim3(kT), 0, s, n, kPcFeat,
                       kPcHid, feat, nullptr, wbf16, kPcOffW1, kPcOffB1, h1);
    hipLaunchKernelGGL((k_dense<true, true>), dim3(grid_for(n * kPcHid)), dim3(kT), 0, s, n, kPcHid,
                       kPcHid, h1, nullptr, wbf16, kPcOffW2, kPcOffB2, h2);
    hipLaunchKernelGGL((k_dense<true, false>), dim3(grid_for(n * kPcOut)), dim3(kT), 0, s, n, kPcHid,
                       kPcOut, h2, nullptr, wbf16, kPcOffW3, kPcOffB3, out);
  } else {
    hipLaunchKernelGGL((k_dense<false, true>), dim3(grid_for(n * kPcHid)), dim3(kT), 0, s, n, kPcFeat,
                       kPcHid, feat, w32, nullptr, kPcOffW1, kPcOffB1, h1);
    hipLaunchKernelGGL((k_dense<false, true>), dim3(grid_for(n * kPcHid)), dim3(kT), 0, s, n, kPcHid,
                       kPcHid, h1, w32, nullptr, kPcOffW2, kPcOffB2, h2);
    hipLaunchKernelGGL((k_dense<false, false>), dim3(grid_for(n * kPcOut)), dim3(kT), 0, s, n, kPcHid,
                       kPcOut, h2, w32, nullptr, kPcOffW3, kPcOffB3, out);
  }
}

void launch_learned_minv(hipStream_t s, const DevPcg& g, const float* out) {
  hipLaunchKernelGGL(k_learned_minv, dim3(grid_for(g.nb)), dim3(kT), 0, s, g, out);
}

void launch_learned_preconditioner(hipStream_t s, const DevPcg& g, const DevLearned& L) {
  launch_precond_features(s, g, L.feat);
  launch_mlp_forward(s, g.nb, L.feat, nullptr, L.wbf16, L.h1, L.h2, L.out);
  launch_learned_minv(s, g, L.out);
}

// ---- training pieces -----------------------------------------------------------------------

void launch_whiten(hipStream_t s, const DevPcg& g, double* Bm, int32_t* ok) {
  hipLaunchKernelGGL(k_whiten, dim3(grid_for(g.nb)), dim3(kT), 0, s, g, Bm, ok);
}
void launch_probe(hipStream_t s, int64_t n, const double* Bm, uint64_t seed, double* w, double* v) {
  hipLaunchKernelGGL(k_probe, dim3(grid_for(n)), dim3(kT), 0, s, n, Bm, seed, w, v);
}
void launch_rcs_matvec(hipStream_t s, const DevPcg& g, const double* v, double* u) {
  hipLaunchKernelGGL(k_rcs_matvec, dim3((g.nb + kRows - 1) / kRows), dim3(kT), 0, s, g, v, u);
}
void launch_utilde(hipStream_t s, int64_t n, const double* Bm, const double* u, double* ut) {
  hipLaunchKernelGGL(k_utilde, dim3(grid_for(n)), dim3(kT), 0, s, n, Bm, u, ut);
}
void launch_loss_grad(hipStream_t s, int64_t n, const int32_t* ok, const double* ut, const double* w,
                      const float* out, float* dO, double* loss, int accumulate) {
  hipLaunchKernelGGL(k_loss_grad, dim3(grid_for(n)), dim3(kT), 0, s, n, ok, ut, w, out, dO, loss,
                     accumulate);
}
void launch_mlp_backward(hipStream_t s, int64_t n, const float* feat, const float* w32, const float* h1,
                         const float* h2, const float* dO, float* dA2, float* dA1, float* grad) {
  // layer 3 (linear): grad W3, b3 from dO and h2
  hipLaunchKernelGGL(k_weight_grad, dim3(grid_for(kPcOut * (kPcHid + 1))), dim3(kT), 0, s, n, kPcOut,
                     kPcHid, dO, h2, kPcOffW3, kPcOffB3, grad);
  hipLaunchKernelGGL(k_back_hidden, dim3(grid_for(n * kPcHid)), dim3(kT), 0, s, n, kPcOut, kPcHid, dO,
                     w32, kPcOffW3, h2, dA2);
  hipLaunchKernelGGL(k_weight_grad, dim3(grid_for(kPcHid * (kPcHid + 1))), dim3(kT), 0, s, n, kPcHid,
                     kPcHid, dA2, h1, kPcOffW2, kPcOffB2, grad);
  hipLaunchKernelGGL(k_back_hidden, dim3(grid_for(n * kPcHid)), dim3(kT), 0, s, n, kPcHid, kPcHid, dA2,
                     w32, kPcOffW2, h1, dA1);
  hipLaunchKernelGGL(k_weight_grad, dim3(grid_for(kPcHid * (kPcFeat + 1))), dim3(kT), 0, s, n, kPcHid,
                     kPcFeat, dA1, feat, kPcOffW1, kPcOffB1, grad);
}
void launch_adam(hipStream_t s, float* w, const float* g, float* m, float* v, float lr, int step,
                 float scale) {
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - std::pow(b1, float(step)), c2 = 1.0f - std::pow(b2, float(step));
  hipLaunchKernelGGL(k_adam, dim3(grid_for(kPcWeights)), dim3(kT), 0, s, kPcWeights, w, g, m, v, lr, b1, b2,
                     1e-8f, c1, c2, scale);
}

// ---- host-array stage entry points ----------------------------------------------------------

void precond_blocks_host(int64_t n, const double* diag_blocks, const float* outputs, double* minv) {
  DeviceArena a;
  double* dD = a.alloc<double>(size_t(36) * n);
  float* dO = a.alloc<float>(size_t(kPcOut) * n);
  double* dM = a.alloc<double>(size_t(36) * n);
  SGPU_HIP(hipMemcpy(dD, diag_blocks, size_t(36) * n * 8, hipMemcpyHostToDevice));
  SGPU_HIP(hipMemcpy(dO, outputs, size_t(kPcOut) * n * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_blocks_from_diag, dim3(grid_for(n)), dim3(kT), 0, nullptr, n, dD, dO, dM);
  SGPU_HIP(hipGetLastError());
  SGPU_HIP(hipMemcpy(minv, dM, size_t(36) * n * 8, hipMemcpyDeviceToHost));
}

void precond_forward_host(int64_t n, const float* features, const float* weights, int use_bf16,
                          float* outputs) {
  DeviceArena a;
  float* dF = a.alloc<float>(size_t(kPcFeat) * n);
  float* h1 = a.alloc<float>(size_t(kPcHid) * n);
  float* h2 = a.alloc<float>(size_t(kPcHid) * n);
  float* dO = a.alloc<float>(size_t(kPcOut) * n);
  SGPU_HIP(hipMemcpy(dF, features, size_t(kPcFeat) * n * 4, hipMemcpyHostToDevice));
  float* w32 = nullptr;
  uint16_t* w16 = nullptr;
  if (use_bf16) {
    const std::vector<uint16_t> hb = weights_to_bf16(weights, kPcWeights);
    w16 = a.upload(hb, nullptr);
  } else {
    w32 = a.alloc<float>(kPcWeights);
    SGPU_HIP(hipMemcpy(w32, weights, size_t(kPcWeights) * 4, hipMemcpyHostToDevice));
  }
  launch_mlp_forward(nullptr, n, dF, w32, w16, h1, h2, dO);
  SGPU_HIP(hipGetLastError());
  SGPU_HIP(hipMemcpy(outputs, dO, size_t(kPcOut) * n * 4, hipMemcpyDeviceToHost));
}

void precond_loss_grad_host(int64_t n, const float* features, const double* diag_blocks,
                            const double* u, const double* w, const float* weights, double* loss,
                            float* grad) {
  DeviceArena a;
  float* dF = a.alloc<float>(size_t(kPcFeat) * n);
  double* dD = a.alloc<double>(size_t(36) * n);
  double* dU = a.alloc<double>(size_t(6) * n);
  double* dW = a.alloc<double>(size_t(6) * n);
  double* dUt = a.alloc<double>(size_t(6) * n);
  double* dB = a.alloc<double>(size_t(36) * n);
  int32_t* ok = a.alloc<int32_t>(size_t(n));
  float* w32 = a.alloc<float>(kPcWeights);
  float* h1 = a.alloc<float>(size_t(kPcHid) * n);
  float* h2 = a.alloc<float>(size_t(kPcHid) * n);
  float* out = a.alloc<float>(size_t(kPcOut) * n);
  float* dO = a.alloc<float>(size_t(kPcOut) * n);
  float* dA2 = a.alloc<float>(size_t(kPcHid) * n);
  float* dA1 = a.alloc<float>(size_t(kPcHid) * n);
  float* g = a.alloc<float>(kPcWeights);
  double* l = a.alloc<double>(size_t(n));
  SGPU_HIP(hipMemcpy(dF, features, size_t(kPcFeat) * n * 4, hipMemcpyHostToDevice));
  SGPU_HIP(hipMemcpy(dD, diag_blocks, size_t(36) * n * 8, hipMemcpyHostToDevice));
  SGPU_HIP(hipMemcpy(dU, u, size_t(6) * n * 8, hipMemcpyHostToDevice));
  SGPU_HIP(hipMemcpy(dW, w, size_t(6) * n * 8, hipMemcpyHostToDevice));
  SGPU_HIP(hipMemcpy(w32, weights, size_t(kPcWeights) * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_whiten_diag, dim3(grid_for(n)), dim3(kT), 0, nullptr, n, dD, dB, ok);
  launch_utilde(nullptr, n, dB, dU, dUt);
  launch_mlp_forward(nullptr, n, dF, w32, nullptr, h1, h2, out);
  launch_loss_grad(nullptr, n, ok, dUt, dW, out, dO, l, 0);
  launch_mlp_backward(nullptr, n, dF, w32, h1, h2, dO, dA2, dA1, g);
  SGPU_HIP(hipGetLastError());
  std::vector<double> hl(static_cast<size_t>(n));
  SGPU_HIP(hipMemcpy(hl.data(), l, size_t(n) * 8, hipMemcpyDeviceToHost));
  double tot = 0.0;
  for (double v : hl) tot += v;
  *loss = tot;
  SGPU_HIP(hipMemcpy(grad, g, size_t(kPcWeights) * 4, hipMemcpyDeviceToHost));
}

}  // namespace sgpu
