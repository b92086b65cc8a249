// Block-Jacobi PCG on the reduced camera system (see pcg.hpp).
//
// One iteration is two launches:
//   k_pcg_update:  alpha = rz / pq; x += alpha p; r -= alpha q; z = M^-1 r;
//                  partials of r.z and r.r
//   k_pcg_matvec:  convergence test |r| <= tol |b|; beta = rz_new / rz_old;
//                  w = S z; q = w + beta q; p = z + beta p  (q = S p without re-reading p);
//                  partials of p.q
// Partial-sum layout in g.part (nv = vector workgroups, nm = matvec workgroups):
//   [0, nv)            |b|^2
//   [nv, 3 nv)         r.z   (two generations, by iteration parity)
//   [3 nv, 5 nv)       r.r   (two generations)
//   [5 nv, 5 nv + nm)  p.q
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "api_internal.hpp"
#include "pcg.hpp"

namespace sgpu {
namespace {

constexpr int kT = 256;
constexpr int kRowsPerMv = kT / 64;  // one wavefront per block row
constexpr int32_t kTransBit = 1 << 30;

// Fixed-order sum of part[0..n) computed by every thread of the workgroup (uniform call).
__device__ double wg_sum(const double* __restrict__ part, int n) {
  __shared__ double red[kT];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += kT) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kT / 2; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  const double out = red[0];
  __syncthreads();
  return out;
}

// state->done as seen by thread 0, broadcast to the workgroup, so every thread takes the
// same early exit (another workgroup of the same launch may set it meanwhile).
__device__ int wg_done(const PcgState* st) {
  __shared__ int flag;
  if (threadIdx.x == 0) flag = *reinterpret_cast<const volatile int32_t*>(&st->done);
  __syncthreads();
  const int v = flag;
  __syncthreads();
  return v;
}

// Sum of one value per thread into part[blockIdx.x] (uniform call).
__device__ void wg_partial(double v, double* __restrict__ out) {
  __shared__ double red[kT];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = kT / 2; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
  __syncthreads();
}

// M^-1 of a symmetric 6x6 block through L D L^T with the reference's PSD pivot rule
// (dense_ldlt.rs:164-256): pivots |d| <= max(1, max|d|) * 1e-12 are zero and their
// directions map to 0.
__device__ void block_inverse6(const double* __restrict__ A, double* __restrict__ out) {
  double L[36], d[6];
  double max_abs_pivot = 1.0;
  for (int j = 0; j < 6; ++j) {
    for (int i = 0; i < 6; ++i) L[6 * i + j] = 0.0;
    double djj = A[6 * j + j];
    for (int k = 0; k < j; ++k) djj -= L[6 * j + k] * L[6 * j + k] * d[k];
    const double tol = fmax(max_abs_pivot, 1.0) * 1e-12;
    if (!(fabs(djj) > tol)) {
      d[j] = 0.0;
    } else {
      d[j] = djj;
      max_abs_pivot = fmax(max_abs_pivot, fabs(djj));
    }
    L[6 * j + j] = 1.0;
    for (int i = j + 1; i < 6; ++i) {
      double s = A[6 * i + j];
      for (int k = 0; k < j; ++k) s -= L[6 * i + k] * L[6 * j + k] * d[k];
      L[6 * i + j] = d[j] == 0.0 ? 0.0 : s / d[j];
    }
  }
  for (int c = 0; c < 6; ++c) {
    double y[6];
    for (int i = 0; i < 6; ++i) {
      double s = i == c ? 1.0 : 0.0;
      for (int k = 0; k < i; ++k) s -= L[6 * i + k] * y[k];
      y[i] = s;
    }
    for (int i = 0; i < 6; ++i) y[i] = d[i] == 0.0 ? 0.0 : y[i] / d[i];
    for (int i = 5; i >= 0; --i) {
      double s = y[i];
      for (int k = i + 1; k < 6; ++k) s -= L[6 * k + i] * y[k];
      y[i] = s;
    }
    for (int i = 0; i < 6; ++i) out[6 * i + c] = y[i];
  }
}

__global__ void __launch_bounds__(kT) k_pcg_block_jacobi(DevPcg g) {
  const int b = blockIdx.x * kT + threadIdx.x;
  if (b >= g.nb) return;
  double A[36], M[36];
  const double* B = g.bsr + 36 * int64_t(g.diag_slot[b]);
  for (int e = 0; e < 36; ++e) A[e] = B[e];
  // the stored diagonal block is symmetric up to rounding; use its lower triangle
  for (int i = 0; i < 6; ++i)
    for (int j = i + 1; j < 6; ++j) A[6 * i + j] = A[6 * j + i];
  block_inverse6(A, M);
  for (int e = 0; e < 36; ++e) g.minv[36 * int64_t(b) + e] = M[e];
}

__device__ __forceinline__ void apply_minv(const double* __restrict__ M, const double r[6],
                                           double z[6]) {
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < 6; ++j) s += M[6 * i + j] * r[j];
    z[i] = s;
  }
}

__global__ void __launch_bounds__(kT) k_pcg_init(DevPcg g, const double* __restrict__ bvec) {
  const int nv = g.n_vec_blocks;
  const int b = blockIdx.x * kT + threadIdx.x;
  double rz = 0.0, bb = 0.0;
  if (b < g.nb) {
    double r[6], z[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) r[k] = bvec[6 * b + k];
    apply_minv(g.minv + 36 * int64_t(b), r, z);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      g.x[6 * b + k] = 0.0;
      g.r[6 * b + k] = r[k];
      g.z[6 * b + k] = z[k];
      g.p[6 * b + k] = z[k];
      rz += r[k] * z[k];
      bb += r[k] * r[k];
    }
  }
  wg_partial(bb, g.part);
  wg_partial(rz, g.part + nv);  // generation 0
  wg_partial(bb, g.part + 3 * nv);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    g.state->done = 0;
    g.state->iters = 0;
    g.state->bad = 0;
  }
}

// y = S v for block row `row` by one wavefront; returns the 6 sums in every lane.
__device__ __forceinline__ void row_product(const DevPcg& g, const double* __restrict__ v, int row,
                                            int lane, double y[6]) {
#pragma unroll
  for (int a = 0; a < 6; ++a) y[a] = 0.0;
  if (row < g.nb) {
    for (int e = g.mv_ptr[row] + lane; e < g.mv_ptr[row + 1]; e += 64) {
      const int slot = g.mv_ent[2 * e];
      const int cw = g.mv_ent[2 * e + 1];
      const int col = cw & (kTransBit - 1);
      const double* B = g.bsr + 36 * int64_t(slot);
      double vc[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) vc[k] = v[6 * col + k];
      if (cw & kTransBit) {
#pragma unroll
        for (int a = 0; a < 6; ++a)
#pragma unroll
          for (int k = 0; k < 6; ++k) y[a] += B[6 * k + a] * vc[k];
      } else {
#pragma unroll
        for (int a = 0; a < 6; ++a)
#pragma unroll
          for (int k = 0; k < 6; ++k) y[a] += B[6 * a + k] * vc[k];
      }
    }
  }
#pragma unroll
  for (int m = 32; m > 0; m >>= 1)
#pragma unroll
    for (int a = 0; a < 6; ++a) y[a] += __shfl_xor(y[a], m, 64);
}

// q = S p and the p.q partials (first iteration).
__global__ void __launch_bounds__(kT) k_pcg_matvec0(DevPcg g) {
  const int nv = g.n_vec_blocks;
  const double bb = wg_sum(g.part, nv);
  if (bb == 0.0) {  // b = 0: x = 0 is the solution
    if (blockIdx.x == 0 && threadIdx.x == 0) g.state->done = 1;
    return;
  }
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * kRowsPerMv + (threadIdx.x >> 6);
  double y[6];
  row_product(g, g.p, row, lane, y);
  double pq = 0.0;
  if (row < g.nb && lane == 0) {
#pragma unroll
    for (int a = 0; a < 6; ++a) {
      g.q[6 * row + a] = y[a];
      pq += g.p[6 * row + a] * y[a];
    }
  }
  wg_partial(pq, g.part + 5 * nv);
}

__global__ void __launch_bounds__(kT) k_pcg_update(DevPcg g, int it) {
  if (wg_done(g.state)) return;
  const int nv = g.n_vec_blocks, nm = g.n_mv_blocks;
  const double rz = wg_sum(g.part + nv + (it & 1) * nv, nv);
  const double pq = wg_sum(g.part + 5 * nv, nm);
  const double alpha = rz / pq;
  if (!(pq > 0.0) || !isfinite(alpha)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      g.state->bad = 1;
      g.state->done = 1;
      g.state->iters = it;
    }
    return;
  }
  const int b = blockIdx.x * kT + threadIdx.x;
  double rz2 = 0.0, rr = 0.0;
  if (b < g.nb) {
    double r[6], z[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      g.x[6 * b + k] += alpha * g.p[6 * b + k];
      r[k] = g.r[6 * b + k] - alpha * g.q[6 * b + k];
      g.r[6 * b + k] = r[k];
    }
    apply_minv(g.minv + 36 * int64_t(b), r, z);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      g.z[6 * b + k] = z[k];
      rz2 += r[k] * z[k];
      rr += r[k] * r[k];
    }
  }
  const int gen = (it + 1) & 1;
  wg_partial(rz2, g.part + nv + gen * nv);
  wg_partial(rr, g.part + 3 * nv + gen * nv);
}

__global__ void __launch_bounds__(kT) k_pcg_matvec(DevPcg g, int it, double rel_tol, int max_iter) {
  if (wg_done(g.state)) return;
  const int nv = g.n_vec_blocks;
  const int gen = (it + 1) & 1;
  const double bb = wg_sum(g.part, nv);
  const double rr = wg_sum(g.part + 3 * nv + gen * nv, nv);
  if (sqrt(rr) <= rel_tol * sqrt(bb) || it + 1 >= max_iter) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      g.state->done = 1;
      g.state->iters = it + 1;
    }
    return;
  }
  const double rz_new = wg_sum(g.part + nv + gen * nv, nv);
  const double rz_old = wg_sum(g.part + nv + (it & 1) * nv, nv);
  const double beta = rz_new / rz_old;
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * kRowsPerMv + (threadIdx.x >> 6);
  double w[6];
  row_product(g, g.z, row, lane, w);
  double pq = 0.0;
  if (row < g.nb && lane == 0) {
#pragma unroll
    for (int a = 0; a < 6; ++a) {
      const double qn = w[a] + beta * g.q[6 * row + a];
      const double pn = g.z[6 * row + a] + beta * g.p[6 * row + a];
      g.q[6 * row + a] = qn;
      g.p[6 * row + a] = pn;
      pq += pn * qn;
    }
  }
  wg_partial(pq, g.part + 5 * nv);
}

}  // namespace

void build_pcg_pattern(int32_t nb, const std::vector<int32_t>& s_ptr,
                       const std::vector<int32_t>& s_col, std::vector<int32_t>& mv_ptr,
                       std::vector<int32_t>& mv_ent, std::vector<int32_t>& diag_slot) {
  std::vector<int32_t> cnt(nb + 1, 0);
  for (int32_t r = 0; r < nb; ++r)
    for (int32_t k = s_ptr[r]; k < s_ptr[r + 1]; ++k) {
      ++cnt[r + 1];
      if (s_col[k] != r) ++cnt[s_col[k] + 1];
    }
  mv_ptr.assign(nb + 1, 0);
  for (int32_t r = 0; r < nb; ++r) mv_ptr[r + 1] = mv_ptr[r] + cnt[r + 1];
  mv_ent.assign(size_t(2) * mv_ptr[nb], 0);
  std::vector<int32_t> fill(mv_ptr.begin(), mv_ptr.end() - 1);
  diag_slot.assign(nb, -1);
  for (int32_t r = 0; r < nb; ++r)
    for (int32_t k = s_ptr[r]; k < s_ptr[r + 1]; ++k) {
      const int32_t c = s_col[k];
      int32_t at = fill[r]++;
      mv_ent[2 * at] = k;
      mv_ent[2 * at + 1] = c;
      if (c == r) {
        diag_slot[r] = k;
      } else {
        at = fill[c]++;
        mv_ent[2 * at] = k;
        mv_ent[2 * at + 1] = r | kTransBit;
      }
    }
  for (int32_t r = 0; r < nb; ++r)
    if (diag_slot[r] < 0) throw Error(SGPU_ERR_INVALID_ARGUMENT, "RCS pattern lacks a diagonal block");
}

int32_t pcg_vec_blocks(int32_t nb) { return std::max<int32_t>(1, (nb + kT - 1) / kT); }
int32_t pcg_mv_blocks(int32_t nb) { return std::max<int32_t>(1, (nb + kRowsPerMv - 1) / kRowsPerMv); }
int64_t pcg_partial_len(int32_t nb) { return 5 * int64_t(pcg_vec_blocks(nb)) + pcg_mv_blocks(nb); }

void launch_pcg_block_jacobi(hipStream_t s, const DevPcg& g) {
  hipLaunchKernelGGL(k_pcg_block_jacobi, dim3(pcg_vec_blocks(g.nb)), dim3(kT), 0, s, g);
}

int32_t run_pcg(hipStream_t s, const DevPcg& g, const double* b, double rel_tol, int32_t max_iter,
                PcgState* h_state, int32_t* bad) {
  constexpr int kCheckEvery = 8;
  const dim3 gv(g.n_vec_blocks), gm(g.n_mv_blocks), blk(kT);
  hipLaunchKernelGGL(k_pcg_init, gv, blk, 0, s, g, b);
  hipLaunchKernelGGL(k_pcg_matvec0, gm, blk, 0, s, g);
  int32_t it = 0;
  *bad = 0;
  if (max_iter < 1) max_iter = 1;
  while (it < max_iter) {
    const int32_t chunk = std::min(kCheckEvery, max_iter - it);
    for (int32_t j = 0; j < chunk; ++j, ++it) {
      hipLaunchKernelGGL(k_pcg_update, gv, blk, 0, s, g, it);
      hipLaunchKernelGGL(k_pcg_matvec, gm, blk, 0, s, g, it, rel_tol, max_iter);
    }
    SGPU_HIP(hipMemcpyAsync(h_state, g.state, sizeof(PcgState), hipMemcpyDeviceToHost, s));
    SGPU_HIP(hipStreamSynchronize(s));
    if (h_state->done) break;
  }
  *bad = h_state->bad;
  return h_state->done ? h_state->iters : it;
}

}  // namespace sgpu
