// 6x6 block LDLt with the reference's PSD pivot rule, shared by the RCS factorizations.
#pragma once

#include <hip/hip_runtime.h>

namespace sgpu {

// 1 / x from the hardware reciprocal refined by two Newton steps (a short dependent chain
// instead of the division sequence; within an ulp of the quotient).
__device__ __forceinline__ double recip_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  r = fma(r, fma(-x, r, 1.0), r);
  return r;
}

// ldlt_decompose_inplace (dense_ldlt.rs:164-256) on a 6x6 block with the PSD pivot rule.
// The off-diagonal quotients use each pivot's reciprocal (formed once); dinv, if given,
// receives those reciprocals (0 for a zero pivot).  Row j keeps w_q = l_jq d_q, so each
// inner term is one multiply-add.
//
// kScalarRule: the pivot rule of the scalar SparseLdlt instead (append_to_ldlt,
// sparse_ldlt.rs:399-445): each pivot is tested on its own, |d_jj| <= max(|d_jj|, 1) tol,
// and a kept pivot's column of L is never zeroed later.  Eliminating a 6x6 block column by
// column is the scalar factorization of those six columns, so with this rule the block
// factorization is SparseLdlt's in the same (block) order.
template <bool kScalarRule = false>
__device__ inline bool ldlt6(const double* A, double* L, double* d, double* dinv = nullptr) {
  double a[36], rd[6], w[36];
  for (int k = 0; k < 36; ++k) a[k] = A[k];
  double max_abs_pivot = 1.0;
  const double tol = 1e-12;
  bool finite = true;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    for (int k = 0; k < j; ++k) {
      double s = a[6 * j + k];
      for (int q = 0; q < k; ++q) s -= w[6 * j + q] * a[6 * k + q];
      if constexpr (kScalarRule) {
        a[6 * j + k] = d[k] == 0.0 ? 0.0 : s * rd[k];
      } else {
        const double zt = fmax(max_abs_pivot, 1.0) * tol;
        a[6 * j + k] = fabs(d[k]) <= zt ? 0.0 : s * rd[k];
      }
      w[6 * j + k] = a[6 * j + k] * d[k];
    }
    double djj = a[6 * j + j];
    for (int q = 0; q < j; ++q) djj -= w[6 * j + q] * a[6 * j + q];
    // branch-free: the pivot's tests are selects and a non-finite pivot is reported after
    // the loop (its outputs are discarded then), which keeps the dependent chain short
    finite = finite && isfinite(djj);
    const double zt = (kScalarRule ? fmax(fabs(djj), 1.0) : fmax(max_abs_pivot, 1.0)) * tol;
    const bool zero = fabs(djj) <= zt;
    const double r = recip_nr(djj);
    d[j] = zero ? 0.0 : djj;
    rd[j] = zero ? 0.0 : r;
    max_abs_pivot = zero ? max_abs_pivot : fmax(max_abs_pivot, fabs(djj));
    a[6 * j + j] = 1.0;
  }
  if (!finite) return false;
  if (dinv)
    for (int k = 0; k < 6; ++k) dinv[k] = rd[k];
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 6; ++c) L[6 * r + c] = c > r ? 0.0 : a[6 * r + c];
  return true;
}

}  // namespace sgpu
