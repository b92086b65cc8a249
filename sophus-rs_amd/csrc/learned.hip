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
// Every layer product runs on the matrix cores, one wavefront per 16 x 16 tile of the
// output (the workgroup's four wavefronts take four neighbouring column tiles):
//   C[i][j] = epi(sum_k A(i, k) B(k, j)),  A(i, k) = A[i sai + k sak],  B(k, j) = B[k sbk + j sbj]
// with the strides choosing the transposes: the forward layers act(X W^T + b), the hidden
// backward products (dA_next W) (1 - H^2), and the weight gradients dA^T X (the long
// reduction over the rows of the batch; the bias gradient against a stride-0 vector of ones).
typedef float mfma_f4 __attribute__((ext_vector_type(4)));
typedef short mfma_s4 __attribute__((ext_vector_type(4)));
enum GemmEpi { kEpiPlain = 0, kEpiTanh = 1, kEpiDtanh = 2 };

// FP32 on v_mfma_f32_16x16x4f32: A lane l holds A(i0 + l % 16, k + l / 16), B lane l holds
// B(k + l / 16, j0 + l % 16); accumulator component q is C(i0 + 4 (l / 16) + q, j0 + l % 16).
template <int kEpi>
__global__ void __launch_bounds__(kT)
    k_gemm_f32(int64_t M, int N, int64_t K, const float* __restrict__ A, int64_t sai, int64_t sak,
               const float* __restrict__ B, int64_t sbk, int64_t sbj, const float* __restrict__ bias,
               const float* __restrict__ H, float* __restrict__ C, int64_t ldc) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, lr = lane & 15, lc = lane >> 4;
  const int64_t i0 = int64_t(blockIdx.x) * 16;
  const int j0 = (int(blockIdx.y) * 4 + wv) * 16;
  if (j0 >= N) return;  // whole wavefront; no barriers here
  const bool iok = i0 + lr < M, jok = j0 + lr < N;
  const float* Ai = A + (iok ? i0 + lr : 0) * sai;
  const float* Bj = B + int64_t(jok ? j0 + lr : 0) * sbj;
  mfma_f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 4
  for (int64_t k0 = 0; k0 < K; k0 += 4) {
    const int64_t k = k0 + lc;
    const float a = iok && k < K ? Ai[k * sak] : 0.0f;
    const float b = jok && k < K ? Bj[k * sbk] : 0.0f;
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int64_t i = i0 + 4 * lc + q;
    const int j = j0 + lr;
    if (i >= M || j >= N) continue;
    float v = acc[q];
    if (kEpi != kEpiDtanh && bias) v += bias[j];
    if (kEpi == kEpiTanh) v = tanhf(v);
    if (kEpi == kEpiDtanh) {
      const float h = H[i * ldc + j];
      v *= 1.0f - h * h;
    }
    C[i * ldc + j] = v;
  }
}

__device__ __forceinline__ short f32_to_bf16(float x) {
  const uint32_t u = __float_as_uint(x);
  return short((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);  // round to nearest even
}

// Inference layer Y = act(X W^T + b) with bf16 weights (W: N x K row-major, b: N) on
// v_mfma_f32_16x16x16bf16_1k, the activations rounded to bf16 as they are loaded, FP32
// accumulation: A lane l holds X(i0 + l % 16, k + 4 (l / 16) .. + 3), B lane l holds
// W(j0 + l % 16, k + 4 (l / 16) .. + 3); accumulator component q is Y(i0 + 4 (l / 16) + q,
// j0 + l % 16).
template <bool kTanh>
__global__ void __launch_bounds__(kT)
    k_dense_bf16(int64_t M, int N, int K, const float* __restrict__ X, const uint16_t* __restrict__ W16,
                 int off_w, int off_b, float* __restrict__ Y) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, lr = lane & 15, lc = lane >> 4;
  const int64_t i0 = int64_t(blockIdx.x) * 16;
  const int j0 = (int(blockIdx.y) * 4 + wv) * 16;
  if (j0 >= N) return;
  const bool iok = i0 + lr < M, jok = j0 + lr < N;
  const float* x = X + (iok ? i0 + lr : 0) * K;
  const uint16_t* w = W16 + off_w + (jok ? j0 + lr : 0) * K;
  mfma_f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    mfma_s4 a, b;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int k = k0 + 4 * lc + e;
      a[e] = iok && k < K ? f32_to_bf16(x[k]) : short(0);
      b[e] = jok && k < K ? short(w[k]) : short(0);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int64_t i = i0 + 4 * lc + q;
    const int j = j0 + lr;
    if (i >= M || j >= N) continue;
    const float v = acc[q] + bf16_to_f32(W16[off_b + j]);
    Y[i * N + j] = kTanh ? tanhf(v) : v;
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

__global__ void __launch_bounds__(kT) k_fill_f32(int64_t n, float* __restrict__ p, float v) {
  const int64_t i = int64_t(blockIdx.x) * kT + threadIdx.x;
  if (i < n) p[i] = v;
}

// Row rowmap[t] of the training data set to row t of the batch buffers (the probe arrays
// are probe-major: probe q of row r at 6 (q cap + r)).
__global__ void __launch_bounds__(kT)
    k_gather_rows(int64_t n, const int64_t* __restrict__ rowmap, int64_t rows_cap, int probes,
                  const float* __restrict__ feat, const int32_t* __restrict__ ok,
                  const double* __restrict__ w, const double* __restrict__ ut, float* __restrict__ feat_out,
                  int32_t* __restrict__ ok_out, double* __restrict__ w_out, double* __restrict__ ut_out,
                  int64_t out_cap) {
  const int64_t t = int64_t(blockIdx.x) * kT + threadIdx.x;
  if (t >= n) return;
  const int64_t r = rowmap[t];
  for (int k = 0; k < kPcFeat; ++k) feat_out[int64_t(kPcFeat) * t + k] = feat[int64_t(kPcFeat) * r + k];
  ok_out[t] = ok[r];
  for (int q = 0; q < probes; ++q)
    for (int k = 0; k < 6; ++k) {
      w_out[6 * (q * out_cap + t) + k] = w[6 * (q * rows_cap + r) + k];
      ut_out[6 * (q * out_cap + t) + k] = ut[6 * (q * rows_cap + r) + k];
    }
}

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

static dim3 gemm_grid(int64_t M, int N) {
  return dim3(unsigned((M + 15) / 16), unsigned((N + 63) / 64));
}

// act(X W^T + b): X n x K row-major, W (at off_w) M x K row-major, b (at off_b).
template <int kEpi>
static void dense_f32(hipStream_t s, int64_t n, int K, int M, const float* X, const float* w32, int off_w,
                      int off_b, float* Y) {
  hipLaunchKernelGGL((k_gemm_f32<kEpi>), gemm_grid(n, M), dim3(kT), 0, s, n, M, int64_t(K), X, int64_t(K),
                     int64_t(1), w32 + off_w, int64_t(1), int64_t(K), w32 + off_b, nullptr, Y, int64_t(M));
}

template <bool kTanh>
static void dense_bf16(hipStream_t s, int64_t n, int K, int M, const float* X, const uint16_t* w16,
                       int off_w, int off_b, float* Y) {
  hipLaunchKernelGGL((k_dense_bf16<kTanh>), gemm_grid(n, M), dim3(kT), 0, s, n, M, K, X, w16, off_w, off_b, Y);
}

void launch_mlp_forward(hipStream_t s, int64_t n, const float* feat, const float* w32,
                        const uint16_t* wbf16, float* h1, float* h2, float* out) {
  if (n <= 0) return;
  if (wbf16) {
    dense_bf16<true>(s, n, kPcFeat, kPcHid, feat, wbf16, kPcOffW1, kPcOffB1, h1);
    dense_bf16<true>(s, n, kPcHid, kPcHid, h1, wbf16, kPcOffW2, kPcOffB2, h2);
    dense_bf16<false>(s, n, kPcHid, kPcOut, h2, wbf16, kPcOffW3, kPcOffB3, out);
  } else {
    dense_f32<kEpiTanh>(s, n, kPcFeat, kPcHid, feat, w32, kPcOffW1, kPcOffB1, h1);
    dense_f32<kEpiTanh>(s, n, kPcHid, kPcHid, h1, w32, kPcOffW2, kPcOffB2, h2);
    dense_f32<kEpiPlain>(s, n, kPcHid, kPcOut, h2, w32, kPcOffW3, kPcOffB3, out);
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

void launch_fill_f32(hipStream_t s, float* p, int64_t n, float v) {
  if (n > 0) hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(n)), dim3(kT), 0, s, n, p, v);
}
void launch_gather_rows(hipStream_t s, int64_t n, const int64_t* rowmap, int64_t rows_cap, int probes,
                        const float* feat, const int32_t* ok, const double* w, const double* ut,
                        float* feat_out, int32_t* ok_out, double* w_out, double* ut_out, int64_t out_cap) {
  if (n > 0)
    hipLaunchKernelGGL(k_gather_rows, dim3(grid_for(n)), dim3(kT), 0, s, n, rowmap, rows_cap, probes, feat,
                       ok, w, ut, feat_out, ok_out, w_out, ut_out, out_cap);
}

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
// grad W (M x K at off_w) = dA^T X and grad b (at off_b) = dA^T 1: the reduction over the n
// rows of the batch runs down the matrix-core k dimension.
static void weight_grad(hipStream_t s, int64_t n, int M, int K, const float* dA, const float* X,
                        const float* ones, int off_w, int off_b, float* grad) {
  hipLaunchKernelGGL((k_gemm_f32<kEpiPlain>), gemm_grid(M, K), dim3(kT), 0, s, int64_t(M), K, n, dA,
                     int64_t(1), int64_t(M), X, int64_t(K), int64_t(1), nullptr, nullptr, grad + off_w,
                     int64_t(K));
  hipLaunchKernelGGL((k_gemm_f32<kEpiPlain>), gemm_grid(M, 1), dim3(kT), 0, s, int64_t(M), 1, n, dA,
                     int64_t(1), int64_t(M), ones, int64_t(1), int64_t(0), nullptr, nullptr, grad + off_b,
                     int64_t(1));
}

// dA (n x M) = (dAn (n x Mn) Wn (Mn x M at off_wn)) (1 - H^2)
static void back_hidden(hipStream_t s, int64_t n, int Mn, int M, const float* dAn, const float* w32,
                        int off_wn, const float* H, float* dA) {
  hipLaunchKernelGGL((k_gemm_f32<kEpiDtanh>), gemm_grid(n, M), dim3(kT), 0, s, n, M, int64_t(Mn), dAn,
                     int64_t(Mn), int64_t(1), w32 + off_wn, int64_t(M), int64_t(1), nullptr, H, dA,
                     int64_t(M));
}

void launch_mlp_backward(hipStream_t s, int64_t n, const float* feat, const float* w32, const float* h1,
                         const float* h2, const float* dO, const float* ones, float* dA2, float* dA1,
                         float* grad) {
  weight_grad(s, n, kPcOut, kPcHid, dO, h2, ones, kPcOffW3, kPcOffB3, grad);
  back_hidden(s, n, kPcOut, kPcHid, dO, w32, kPcOffW3, h2, dA2);
  weight_grad(s, n, kPcHid, kPcHid, dA2, h1, ones, kPcOffW2, kPcOffB2, grad);
  back_hidden(s, n, kPcHid, kPcHid, dA2, w32, kPcOffW2, h1, dA1);
  weight_grad(s, n, kPcHid, kPcFeat, dA1, feat, ones, kPcOffW1, kPcOffB1, grad);
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
  float* ones = a.alloc<float>(size_t(n));
  launch_fill_f32(nullptr, ones, n, 1.0f);
  SGPU_HIP(hipMemcpy(dF, features, size_t(kPcFeat) * n * 4, hipMemcpyHostToDevice));
  SGPU_HIP(hipMemcpy(dD, diag_blocks, size_t(36) * n * 8, hipMemcpyHostToDevice));
  SGPU_HIP(hipMemcpy(dU, u, size_t(6) * n * 8, hipMemcpyHostToDevice));
  SGPU_HIP(hipMemcpy(dW, w, size_t(6) * n * 8, hipMemcpyHostToDevice));
  SGPU_HIP(hipMemcpy(w32, weights, size_t(kPcWeights) * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_whiten_diag, dim3(grid_for(n)), dim3(kT), 0, nullptr, n, dD, dB, ok);
  launch_utilde(nullptr, n, dB, dU, dUt);
  launch_mlp_forward(nullptr, n, dF, w32, nullptr, h1, h2, out);
  launch_loss_grad(nullptr, n, ok, dUt, dW, out, dO, l, 0);
  launch_mlp_backward(nullptr, n, dF, w32, h1, h2, dO, ones, dA2, dA1, g);
  SGPU_HIP(hipGetLastError());
  std::vector<double> hl(static_cast<size_t>(n));
  SGPU_HIP(hipMemcpy(hl.data(), l, size_t(n) * 8, hipMemcpyDeviceToHost));
  double tot = 0.0;
  for (double v : hl) tot += v;
  *loss = tot;
  SGPU_HIP(hipMemcpy(grad, g, size_t(kPcWeights) * 4, hipMemcpyDeviceToHost));
}

}  // namespace sgpu
