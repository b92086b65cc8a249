// Direct block-LDLt solve of one symmetric system with 6x6 blocks on the device: the
// buffers of the multifrontal factorization are bound from an allocator, then the blocks
// are scattered into the fronts, factored and solved.  Shared by sgpu_block_ldlt_solve
// (device memory from hipMalloc) and its allocation-free variant (device memory carved
// from a static array), so both run the same launches on the same layout.
//
// An allocator provides
//   T* alloc<T>(size_t n)                      n elements of uninitialised device memory
//   T* up<T>(const T* host, size_t n)          n elements copied from the host
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "api_internal.hpp"
#include "kernels.hpp"
#include "problem.hpp"

namespace sgpu {

// Lower blocks in CSC-by-block-column (the input of sgpu_block_ldlt_solve) -> the slot
// pattern of S (rows of lower blocks, camera order) and row-major blocks in slot order.
void parse_lower_blocks(int64_t num_blocks, const int64_t* col_ptr, const int64_t* row_idx,
                        const double* values, const double* rhs, const double* x, Symbolic& S,
                        std::vector<double>& blocks);

struct BlockSolveBufs {
  DevProblem dp{};
  DevFactor df{};
  double* blocks = nullptr;  // slot-order row-major blocks
  double* rhs = nullptr;     // natural order
  double* y = nullptr;       // elimination order (forward-solve result)
  double* x = nullptr;       // natural order (solution)
  DevScalars* sc = nullptr;
};

template <typename Alloc, typename T>
T* bind_up(Alloc& a, const std::vector<T>& v) {
  return a.template up<T>(v.data(), v.size());
}

// Binds every device buffer of the factorization of S (after analyze_rcs) from `a`.
template <typename Alloc>
BlockSolveBufs bind_block_solve(const Symbolic& S, const std::vector<double>& blocks,
                                const double* rhs, Alloc& a) {
  BlockSolveBufs b;
  const int32_t nb = S.nb;
  b.dp.nb = nb;
  b.dp.s_off = bind_up(a, S.s_off);
  b.dp.s_ld = bind_up(a, S.s_ld);
  DevFactor& df = b.df;
  df.nb = nb;
  df.nsn = S.nsn;
  df.perm = bind_up(a, S.perm);
  df.sn_c0 = bind_up(a, S.sn_c0);
  df.sn_nc = bind_up(a, S.sn_nc);
  df.sn_nf = bind_up(a, S.sn_nf);
  df.sn_off = bind_up(a, S.sn_off);
  df.sn_loff = bind_up(a, S.sn_loff);
  df.sn_rows_ptr = bind_up(a, S.sn_rows_ptr);
  df.sn_rows = bind_up(a, S.sn_rows);
  df.sn_parent = bind_up(a, S.sn_parent);
  df.sn_upos_ptr = bind_up(a, S.sn_upos_ptr);
  df.sn_upos = bind_up(a, S.sn_upos);
  df.lvl_sn = bind_up(a, S.lvl_sn);
  df.launch_tiles = bind_up(a, S.launch_tiles);
  df.leaves = bind_up(a, S.leaves);
  df.sn_hand = bind_up(a, S.sn_hand);
  df.sn_toff = bind_up(a, S.sn_toff);
  df.sn_poff = bind_up(a, S.sn_poff);
  df.panel_items = bind_up(a, S.panel_items);
  df.sn_woff = bind_up(a, S.sn_woff);
  df.solve_cols = bind_up(a, S.solve_cols);
  df.sn_uoff = bind_up(a, S.sn_uoff);
  df.usolve = a.template alloc<double>(size_t(S.u_elems));
  df.front = a.template alloc<double>(size_t(S.front_elems));
  df.lbuf = a.template alloc<double>(size_t(S.l_elems));
  df.ltbuf = a.template alloc<double>(size_t(S.t_elems));
  df.tinv = a.template alloc<double>(size_t(S.p_elems));
  df.wbuf = a.template alloc<double>(size_t(S.w_elems));
  df.piv_L = a.template alloc<double>(size_t(36) * nb);
  df.piv_d = a.template alloc<double>(size_t(6) * nb);
  df.piv_M = a.template alloc<double>(size_t(36) * nb);
  df.front_elems = S.front_elems;
  df.host = &S;
  b.blocks = bind_up(a, blocks);
  b.rhs = a.template up<double>(rhs, size_t(6) * nb);
  b.y = a.template alloc<double>(size_t(6) * nb);
  b.x = a.template alloc<double>(size_t(6) * nb);
  b.sc = a.template alloc<DevScalars>(1);
  return b;
}

// Zeroes the scalars and the fronts (every call starts from zero, also when the buffers
// hold a previous call's values), scatters S, factors, solves and copies x to the host.
inline void run_block_solve(const Symbolic& S, BlockSolveBufs& b, double* x_host) {
  launch_zero_scalars(nullptr, b.sc);
  launch_zero_fronts(nullptr, b.df);
  launch_scatter_s(nullptr, b.dp, int64_t(S.s_col.size()), b.blocks, b.df);
  launch_rcs_factor(nullptr, b.df, b.sc, b.rhs, b.y);
  launch_rcs_solve(nullptr, b.df, b.y, b.x);
  SGPU_HIP(hipGetLastError());
  SGPU_HIP(hipDeviceSynchronize());
  DevScalars sc;
  SGPU_HIP(hipMemcpy(&sc, b.sc, sizeof(sc), hipMemcpyDeviceToHost));
  if (sc.bad_pivot) throw Error(SGPU_ERR_NON_FINITE, "non-finite pivot");
  SGPU_HIP(hipMemcpy(x_host, b.x, size_t(6) * S.nb * 8, hipMemcpyDeviceToHost));
}

}  // namespace sgpu
