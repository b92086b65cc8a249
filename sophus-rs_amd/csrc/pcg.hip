// Block-Jacobi PCG on the reduced camera system (see pcg.hpp).
//
// Per linearization, k_pcg_expand writes S as full block rows (both triangles, each
// block oriented for its row).  A row's blocks are stored element-major: element k of
// the row's j-th block sits at 36 * row_start + k * row_len + j, so when the lanes of a
// wavefront take consecutive blocks of one row every load is coalesced.  One iteration is two launches:
//   k_pcg_update (32 camera blocks per workgroup, one thread per scalar row):
//       alpha = rz / pq; x += alpha p; r -= alpha q; z = M^-1 r; partials of r.z, r.r
//   k_pcg_matvec (4 block rows per workgroup, one wavefront per block row):
//       convergence test; beta = rz_new / rz_old; w = S z; q = w + beta q;
//       p = z + beta p (so q = S p without re-reading p); partials of p.q
// Each launch recomputes the scalars it needs from the previous launch's per-workgroup
// partials (fixed order, identical in every workgroup), so no atomics or grid barriers
// are needed and the solve is deterministic.  r.z and r.r partials are double-buffered
// by iteration parity.  Layout of g.part (nv, nm = workgroups of the two kernels):
//   |b|^2 [0, nv), r.z [nv, 3nv), r.r [3nv, 5nv), p.q [5nv, 5nv + nm).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "api_internal.hpp"
#include "pcg.hpp"

namespace sgpu {
namespace {

constexpr int kVecBlocks = 32;         // camera blocks per update workgroup
constexpr int kVecT = 6 * kVecBlocks;  // 192 threads: one per scalar row
constexpr int kMvRows = 4;             // block rows per matvec workgroup
constexpr int kMvT = 64 * kMvRows;     // 256 threads: one wavefront per block row
constexpr int kT = 256;                // per-block setup kernels
constexpr int32_t kTransBit = 1 << 30;

// Sum of part[0..n) in a fixed order, returned to every thread (uniform call).
template <int NT>
__device__ double wg_sum(const double* __restrict__ part, int n) {
  __shared__ double ws[NT / 64];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += NT) s += part[i];
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) s += __shfl_xor(s, m, 64);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) t += ws[w];
  __syncthreads();
  return t;
}

// One value per thread summed over the workgroup into out[blockIdx.x] (uniform call).
template <int NT>
__device__ void wg_partial(double v, double* __restrict__ out) {
  __shared__ double ws[NT / 64];
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) t += ws[w];
    out[blockIdx.x] = t;
  }
  __syncthreads();
}

// state->done as seen by thread 0, broadcast (another workgroup of the same launch may
// set it meanwhile; every thread of a workgroup takes the same early exit).
__device__ int wg_done(const PcgState* st) {
  __shared__ int flag;
  if (threadIdx.x == 0) flag = *reinterpret_cast<const volatile int32_t*>(&st->done);
  __syncthreads();
  const int v = flag;
  __syncthreads();
  return v;
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

// Full block rows: block e of row r is S_{r, col(e)} oriented for the row, element-major
// within the row.
__global__ void __launch_bounds__(kT) k_pcg_expand(DevPcg g, int64_t n_elems) {
  const int64_t t = int64_t(blockIdx.x) * kT + threadIdx.x;
  if (t >= n_elems) return;
  const int64_t e = t / 36;
  const int k = int(t - 36 * e), a = k / 6, b = k % 6;
  const int slot = g.mv_ent[2 * e];
  const int cw = g.mv_ent[2 * e + 1];
  const bool trans = (cw & kTransBit) != 0;
  const int row = g.frow[e];
  const int64_t start = g.mv_ptr[row];
  const int64_t len = g.mv_ptr[row + 1] - start;
  g.fbsr[36 * start + k * len + (e - start)] =
      g.bsr[36 * int64_t(slot) + (trans ? 6 * b + a : 6 * a + b)];
  if (k == 0) g.fcol[e] = cw & (kTransBit - 1);
}

// x = 0, r = b, z = M^-1 b (p is set by the first product); partials |b|^2 and r.z.
__global__ void __launch_bounds__(kVecT) k_pcg_init(DevPcg g, const double* __restrict__ bvec) {
  __shared__ double rs[kVecT];
  const int nv = g.n_vec_blocks;
  const int row = blockIdx.x * kVecT + threadIdx.x;  // scalar row
  const int blk = row / 6, i = row % 6;
  const bool on = blk < g.nb;
  const double r = on ? bvec[row] : 0.0;
  rs[threadIdx.x] = r;
  __syncthreads();
  double z = 0.0;
  if (on) {
    const double* M = g.minv + 36 * int64_t(blk) + 6 * i;
    const int base = threadIdx.x - i;
#pragma unroll
    for (int j = 0; j < 6; ++j) z += M[j] * rs[base + j];
    g.x[row] = 0.0;
    g.r[row] = r;
    g.z[row] = z;
  }
  wg_partial<kVecT>(r * r, g.part);
  wg_partial<kVecT>(r * z, g.part + nv);  // generation 0
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    g.state->done = 0;
    g.state->iters = 0;
    g.state->bad = 0;
  }
}

__global__ void __launch_bounds__(kVecT) k_pcg_update(DevPcg g, int it) {
  if (wg_done(g.state)) return;
  __shared__ double rs[kVecT];
  const int nv = g.n_vec_blocks, nm = g.n_mv_blocks;
  const double rz = wg_sum<kVecT>(g.part + nv + (it & 1) * nv, nv);
  const double pq = wg_sum<kVecT>(g.part + 5 * nv, nm);
  const double alpha = rz / pq;
  if (!(pq > 0.0) || !isfinite(alpha)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      g.state->bad = 1;
      g.state->done = 1;
      g.state->iters = it;
    }
    return;
  }
  const int row = blockIdx.x * kVecT + threadIdx.x;
  const int blk = row / 6, i = row % 6;
  const bool on = blk < g.nb;
  double r = 0.0;
  if (on) {
    g.x[row] += alpha * g.p[row];
    r = g.r[row] - alpha * g.q[row];
    g.r[row] = r;
  }
  rs[threadIdx.x] = r;
  __syncthreads();
  double z = 0.0;
  if (on) {
    const double* M = g.minv + 36 * int64_t(blk) + 6 * i;
    const int base = threadIdx.x - i;
#pragma unroll
    for (int j = 0; j < 6; ++j) z += M[j] * rs[base + j];
    g.z[row] = z;
  }
  const int gen = (it + 1) & 1;
  wg_partial<kVecT>(r * z, g.part + nv + gen * nv);
  wg_partial<kVecT>(r * r, g.part + 3 * nv + gen * nv);
}

// it < 0: the first product, q = S z and p = z.
__global__ void __launch_bounds__(kMvT)
    k_pcg_matvec(DevPcg g, int it, double rel_tol, int max_iter) {
  const int nv = g.n_vec_blocks;
  double beta = 0.0;
  if (it < 0) {
    const double bb = wg_sum<kMvT>(g.part, nv);
    if (bb == 0.0) {  // b = 0: x = 0 is the solution
      if (blockIdx.x == 0 && threadIdx.x == 0) g.state->done = 1;
      return;
    }
  } else {
    if (wg_done(g.state)) return;
    const int gen = (it + 1) & 1;
    const double bb = wg_sum<kMvT>(g.part, nv);
    const double rr = wg_sum<kMvT>(g.part + 3 * nv + gen * nv, nv);
    if (sqrt(rr) <= rel_tol * sqrt(bb) || it + 1 >= max_iter) {
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        g.state->done = 1;
        g.state->iters = it + 1;
      }
      return;
    }
    const double rz_new = wg_sum<kMvT>(g.part + nv + gen * nv, nv);
    const double rz_old = wg_sum<kMvT>(g.part + nv + (it & 1) * nv, nv);
    beta = rz_new / rz_old;
  }
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * kMvRows + (threadIdx.x >> 6);
  double y[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (row < g.nb) {
    const int start = g.mv_ptr[row], len = g.mv_ptr[row + 1] - start;
    const double* Brow = g.fbsr + 36 * int64_t(start);
    for (int j = lane; j < len; j += 64) {
      const int col = g.fcol[start + j];
      double vc[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) vc[k] = g.z[6 * col + k];
#pragma unroll
      for (int a = 0; a < 6; ++a)
#pragma unroll
        for (int k = 0; k < 6; ++k) y[a] += Brow[int64_t(6 * a + k) * len + j] * vc[k];
    }
  }
#pragma unroll
  for (int m = 32; m > 0; m >>= 1)
#pragma unroll
    for (int a = 0; a < 6; ++a) y[a] += __shfl_xor(y[a], m, 64);
  double pq = 0.0;
  if (row < g.nb && lane < 6) {
    double w = y[0];
#pragma unroll
    for (int a = 1; a < 6; ++a)
      if (lane == a) w = y[a];
    const int idx = 6 * row + lane;
    const double zn = g.z[idx];
    const double qn = it < 0 ? w : w + beta * g.q[idx];
    const double pn = it < 0 ? zn : zn + beta * g.p[idx];
    g.q[idx] = qn;
    g.p[idx] = pn;
    pq = pn * qn;
  }
  wg_partial<kMvT>(pq, g.part + 5 * nv);
}

// Copies the final iteration count / breakdown flag into the caller's device words.
__global__ void k_pcg_report(const PcgState* st, int32_t* iters_out, int32_t* bad_out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    if (iters_out) *iters_out = st->iters;
    if (bad_out && st->bad) *bad_out = 1;
  }
}

}  // namespace

void build_pcg_pattern(int32_t nb, const std::vector<int32_t>& s_ptr,
                       const std::vector<int32_t>& s_col, std::vector<int32_t>& mv_ptr,
                       std::vector<int32_t>& mv_ent, std::vector<int32_t>& diag_slot,
                       std::vector<int32_t>& mv_row) {
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
  mv_row.assign(mv_ptr[nb], 0);
  for (int32_t r = 0; r < nb; ++r)
    for (int32_t e = mv_ptr[r]; e < mv_ptr[r + 1]; ++e) mv_row[e] = r;
}

int32_t pcg_vec_blocks(int32_t nb) { return std::max<int32_t>(1, (nb + kVecBlocks - 1) / kVecBlocks); }
int32_t pcg_mv_blocks(int32_t nb) { return std::max<int32_t>(1, (nb + kMvRows - 1) / kMvRows); }
int64_t pcg_partial_len(int32_t nb) { return 5 * int64_t(pcg_vec_blocks(nb)) + pcg_mv_blocks(nb); }

void launch_pcg_setup(hipStream_t s, const DevPcg& g) {
  hipLaunchKernelGGL(k_pcg_block_jacobi, dim3((g.nb + kT - 1) / kT), dim3(kT), 0, s, g);
  const int64_t n = 36 * int64_t(g.n_full);
  if (n > 0)
    hipLaunchKernelGGL(k_pcg_expand, dim3(unsigned((n + kT - 1) / kT)), dim3(kT), 0, s, g, n);
}

void launch_pcg(hipStream_t s, const DevPcg& g, const double* b, double rel_tol, int32_t max_iter,
                int32_t* iters_out, int32_t* bad_out, PcgState* h_state) {
  constexpr int kCheckEvery = 16;
  const dim3 gv(g.n_vec_blocks), gm(g.n_mv_blocks);
  if (max_iter < 1) max_iter = 1;
  hipLaunchKernelGGL(k_pcg_init, gv, dim3(kVecT), 0, s, g, b);
  hipLaunchKernelGGL(k_pcg_matvec, gm, dim3(kMvT), 0, s, g, -1, rel_tol, max_iter);
  for (int32_t it = 0; it < max_iter;) {
    const int32_t chunk = std::min(kCheckEvery, max_iter - it);
    for (int32_t j = 0; j < chunk; ++j, ++it) {
      hipLaunchKernelGGL(k_pcg_update, gv, dim3(kVecT), 0, s, g, it);
      hipLaunchKernelGGL(k_pcg_matvec, gm, dim3(kMvT), 0, s, g, it, rel_tol, max_iter);
    }
    if (!h_state || it >= max_iter) break;
    // poll: stop launching once the device reports done
    SGPU_HIP(hipMemcpyAsync(h_state, g.state, sizeof(PcgState), hipMemcpyDeviceToHost, s));
    SGPU_HIP(hipStreamSynchronize(s));
    if (h_state->done) break;
  }
  hipLaunchKernelGGL(k_pcg_report, dim3(1), dim3(64), 0, s, g.state, iters_out, bad_out);
}

}  // namespace sgpu
