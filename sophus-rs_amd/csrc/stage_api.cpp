// Stage-level C entry points: one kernel at a time on host arrays, for parity tests
// of a single stage against the oracle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "api_internal.hpp"
#include "block_solve.hpp"
#include "kernels.hpp"
#include "learned.hpp"
#include "pcg.hpp"
#include "problem.hpp"

using namespace sgpu;

namespace {

struct Buf {
  void* p = nullptr;
  ~Buf() {
    if (p) (void)hipFree(p);
  }
  template <typename T>
  T* get(size_t n) {
    SGPU_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    return static_cast<T*>(p);
  }
};

template <typename T>
T* upload(Buf& b, const T* host, size_t n) {
  T* d = b.get<T>(n);
  SGPU_HIP(hipMemcpy(d, host, n * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

struct DeviceBufs {
  std::vector<void*> ptrs;
  ~DeviceBufs() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  T* alloc(size_t n) {
    void* p = nullptr;
    SGPU_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    ptrs.push_back(p);
    return static_cast<T*>(p);
  }
  template <typename T>
  T* up(const T* host, size_t n) {
    T* d = alloc<T>(n);
    if (n) SGPU_HIP(hipMemcpy(d, host, n * sizeof(T), hipMemcpyHostToDevice));
    return d;
  }
  template <typename T>
  T* up(const std::vector<T>& v) {
    return up<T>(v.data(), v.size());
  }
};

void require_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) throw Error(SGPU_ERR_NO_DEVICE, "no HIP device");
}

}  // namespace

namespace sgpu {
// Parses lower blocks in CSC-by-block-column (see sgpu_block_ldlt_solve) into the slot
// pattern (rows of lower blocks) and row-major blocks in slot order.
void parse_lower_blocks(int64_t num_blocks, const int64_t* col_ptr, const int64_t* row_idx,
                               const double* values, const double* rhs, const double* x,
                               Symbolic& S, std::vector<double>& blocks) {
  if (num_blocks <= 0 || !col_ptr || !row_idx || !values || !rhs || !x)
    throw Error(SGPU_ERR_INVALID_ARGUMENT, "bad arguments");
  if (num_blocks > INT32_MAX / 64) throw Error(SGPU_ERR_UNSUPPORTED, "too many blocks");
  const int32_t nb = int32_t(num_blocks);
  const int64_t nnz = col_ptr[nb];
  if (col_ptr[0] != 0 || nnz < 0) throw Error(SGPU_ERR_INVALID_ARGUMENT, "bad col_ptr");
  std::vector<std::vector<int32_t>> rows(nb);
  for (int32_t c = 0; c < nb; ++c) {
    if (col_ptr[c + 1] < col_ptr[c]) throw Error(SGPU_ERR_INVALID_ARGUMENT, "bad col_ptr");
    rows[c].push_back(c);
    for (int64_t k = col_ptr[c]; k < col_ptr[c + 1]; ++k) {
      const int64_t r = row_idx[k];
      if (r < c || r >= nb)
        throw Error(SGPU_ERR_INDEX_OUT_OF_RANGE, "block row must satisfy col <= row < nb");
      rows[r].push_back(c);
    }
  }
  S = Symbolic();
  S.nb = nb;
  S.s_ptr.assign(nb + 1, 0);
  for (int32_t r = 0; r < nb; ++r) {
    std::sort(rows[r].begin(), rows[r].end());
    rows[r].erase(std::unique(rows[r].begin(), rows[r].end()), rows[r].end());
    S.s_ptr[r + 1] = S.s_ptr[r] + int32_t(rows[r].size());
  }
  S.s_col.resize(S.s_ptr[nb]);
  for (int32_t r = 0; r < nb; ++r) std::copy(rows[r].begin(), rows[r].end(), &S.s_col[S.s_ptr[r]]);
  blocks.assign(size_t(36) * S.s_col.size(), 0.0);
  for (int32_t c = 0; c < nb; ++c)
    for (int64_t k = col_ptr[c]; k < col_ptr[c + 1]; ++k) {
      const int32_t r = int32_t(row_idx[k]);
      auto b = S.s_col.begin() + S.s_ptr[r], e = S.s_col.begin() + S.s_ptr[r + 1];
      const int64_t slot = int64_t(std::lower_bound(b, e, c) - S.s_col.begin());
      for (int a = 0; a < 6; ++a)
        for (int bb = 0; bb < 6; ++bb)
          blocks[36 * slot + 6 * a + bb] += values[36 * k + 6 * bb + a];  // input column-major
    }
}
}  // namespace sgpu

extern "C" sgpu_status sgpu_eval_pinhole_terms(int32_t precision, int64_t n, const double* intr,
                                               const double* pose, const double* pt,
                                               const double* uv, int32_t robust, double delta,
                                               double* res, double* jpose, double* jpt,
                                               double* cost, double* weight) {
  return guarded([&] {
    if (n < 0 || (n > 0 && (!intr || !pose || !pt || !uv)))
      throw Error(SGPU_ERR_INVALID_ARGUMENT, "bad arguments");
    if (n == 0) return;
    require_device();
    Buf bi, bp, bx, bu, br, bj, bk, bc, bw;
    const double* di = upload(bi, intr, 4 * n);
    const double* dpo = upload(bp, pose, 7 * n);
    const double* dx = upload(bx, pt, 3 * n);
    const double* du = upload(bu, uv, 2 * n);
    double* dr = br.get<double>(2 * n);
    double* dj = bj.get<double>(12 * n);
    double* dk = bk.get<double>(6 * n);
    double* dc = bc.get<double>(n);
    double* dw = bw.get<double>(n);
    launch_eval_pinhole(nullptr, precision, n, di, dpo, dx, du, robust, delta, dr, dj, dk, dc, dw);
    SGPU_HIP(hipGetLastError());
    SGPU_HIP(hipDeviceSynchronize());
    if (res) SGPU_HIP(hipMemcpy(res, dr, 2 * n * 8, hipMemcpyDeviceToHost));
    if (jpose) SGPU_HIP(hipMemcpy(jpose, dj, 12 * n * 8, hipMemcpyDeviceToHost));
    if (jpt) SGPU_HIP(hipMemcpy(jpt, dk, 6 * n * 8, hipMemcpyDeviceToHost));
    if (cost) SGPU_HIP(hipMemcpy(cost, dc, n * 8, hipMemcpyDeviceToHost));
    if (weight) SGPU_HIP(hipMemcpy(weight, dw, n * 8, hipMemcpyDeviceToHost));
  });
}

extern "C" sgpu_status sgpu_eval_eucm_terms(int64_t n, const double* intr, const double* pose,
                                            const double* pt, const double* uv, int32_t robust,
                                            double delta, double* res, double* jintr,
                                            double* jpose, double* jpt, double* cost,
                                            double* weight) {
  return guarded([&] {
    if (n < 0 || (n > 0 && (!intr || !pose || !pt || !uv)))
      throw Error(SGPU_ERR_INVALID_ARGUMENT, "bad arguments");
    if (n == 0) return;
    require_device();
    Buf bi, bp, bx, bu, br, bq, bj, bk, bc, bw;
    const double* di = upload(bi, intr, 6 * n);
    const double* dpo = upload(bp, pose, 7 * n);
    const double* dx = upload(bx, pt, 3 * n);
    const double* du = upload(bu, uv, 2 * n);
    double* dr = br.get<double>(2 * n);
    double* dq = bq.get<double>(12 * n);
    double* dj = bj.get<double>(12 * n);
    double* dk = bk.get<double>(6 * n);
    double* dc = bc.get<double>(n);
    double* dw = bw.get<double>(n);
    launch_eval_eucm(nullptr, n, di, dpo, dx, du, robust, delta, dr, dq, dj, dk, dc, dw);
    SGPU_HIP(hipGetLastError());
    SGPU_HIP(hipDeviceSynchronize());
    if (res) SGPU_HIP(hipMemcpy(res, dr, 2 * n * 8, hipMemcpyDeviceToHost));
    if (jintr) SGPU_HIP(hipMemcpy(jintr, dq, 12 * n * 8, hipMemcpyDeviceToHost));
    if (jpose) SGPU_HIP(hipMemcpy(jpose, dj, 12 * n * 8, hipMemcpyDeviceToHost));
    if (jpt) SGPU_HIP(hipMemcpy(jpt, dk, 6 * n * 8, hipMemcpyDeviceToHost));
    if (cost) SGPU_HIP(hipMemcpy(cost, dc, n * 8, hipMemcpyDeviceToHost));
    if (weight) SGPU_HIP(hipMemcpy(weight, dw, n * 8, hipMemcpyDeviceToHost));
  });
}

extern "C" sgpu_status sgpu_eval_isometry3_prior_terms(int64_t n, const double* poses,
                                                       const double* means, const double* prec,
                                                       int32_t robust, double delta, double* res,
                                                       double* jac, double* cost) {
  return guarded([&] {
    if (n < 0 || (n > 0 && (!poses || !means || !prec || !res || !jac || !cost)))
      throw Error(SGPU_ERR_INVALID_ARGUMENT, "bad arguments");
    if (robust != SGPU_ROBUST_NONE && robust != SGPU_ROBUST_HUBER)
      throw Error(SGPU_ERR_INVALID_ARGUMENT, "unknown robust kernel");
    if (n == 0) return;
    require_device();
    Buf b1, b2, b3, b4, b5, b6;
    const double* d1 = upload(b1, poses, 7 * n);
    const double* d2 = upload(b2, means, 7 * n);
    const double* d3 = upload(b3, prec, 36 * n);
    double* d4 = b4.get<double>(6 * n);
    double* d5 = b5.get<double>(36 * n);
    double* d6 = b6.get<double>(n);
    launch_eval_prior(nullptr, n, d1, d2, d3, robust, delta, d4, d5, d6);
    SGPU_HIP(hipGetLastError());
    SGPU_HIP(hipDeviceSynchronize());
    SGPU_HIP(hipMemcpy(res, d4, 6 * n * 8, hipMemcpyDeviceToHost));
    SGPU_HIP(hipMemcpy(jac, d5, 36 * n * 8, hipMemcpyDeviceToHost));
    SGPU_HIP(hipMemcpy(cost, d6, n * 8, hipMemcpyDeviceToHost));
  });
}

extern "C" sgpu_status sgpu_eval_pose_graph3_terms(int64_t n, const double* m, const double* nn,
                                                   const double* mn, double* res, double* jm,
                                                   double* jn) {
  return guarded([&] {
    if (n < 0 || (n > 0 && (!m || !nn || !mn || !res || !jm || !jn)))
      throw Error(SGPU_ERR_INVALID_ARGUMENT, "bad arguments");
    if (n == 0) return;
    require_device();
    Buf b1, b2, b3, b4, b5, b6;
    const double* d1 = upload(b1, m, 7 * n);
    const double* d2 = upload(b2, nn, 7 * n);
    const double* d3 = upload(b3, mn, 7 * n);
    double* d4 = b4.get<double>(6 * n);
    double* d5 = b5.get<double>(36 * n);
    double* d6 = b6.get<double>(36 * n);
    launch_eval_pose_graph(nullptr, n, d1, d2, d3, d4, d5, d6);
    SGPU_HIP(hipGetLastError());
    SGPU_HIP(hipDeviceSynchronize());
    SGPU_HIP(hipMemcpy(res, d4, 6 * n * 8, hipMemcpyDeviceToHost));
    SGPU_HIP(hipMemcpy(jm, d5, 36 * n * 8, hipMemcpyDeviceToHost));
    SGPU_HIP(hipMemcpy(jn, d6, 36 * n * 8, hipMemcpyDeviceToHost));
  });
}

namespace {

sgpu_status direct_block_solve(int64_t num_blocks, const int64_t* col_ptr, const int64_t* row_idx,
                               const double* values, const double* rhs, double* x, bool scalar_rule) {
  return guarded([&] {
    Symbolic S;
    std::vector<double> blocks;
    parse_lower_blocks(num_blocks, col_ptr, row_idx, values, rhs, x, S, blocks);
    analyze_rcs(S);
    require_device();
    DeviceBufs bufs;
    BlockSolveBufs b = bind_block_solve(S, blocks, rhs, bufs);
    b.df.scalar_pivots = scalar_rule ? 1 : 0;
    run_block_solve(S, b, x);
  });
}

}  // namespace

extern "C" sgpu_status sgpu_block_ldlt_solve(int64_t num_blocks, const int64_t* col_ptr,
                                             const int64_t* row_idx, const double* values,
                                             const double* rhs, double* x) {
  return direct_block_solve(num_blocks, col_ptr, row_idx, values, rhs, x, false);
}

extern "C" sgpu_status sgpu_sparse_ldlt_solve(int64_t num_blocks, const int64_t* col_ptr,
                                              const int64_t* row_idx, const double* values,
                                              const double* rhs, double* x) {
  return direct_block_solve(num_blocks, col_ptr, row_idx, values, rhs, x, true);
}

sgpu_status sgpu_block_pcg_solve(int64_t num_blocks, const int64_t* col_ptr, const int64_t* row_idx,
                                 const double* values, const double* rhs, double* x,
                                 double relative_tolerance, int64_t max_iterations,
                                 int64_t* iterations) {
  return guarded([&] {
    if (!(relative_tolerance > 0.0) || max_iterations <= 0)
      throw Error(SGPU_ERR_INVALID_ARGUMENT, "PCG needs a positive tolerance and iteration cap");
    Symbolic S;
    std::vector<double> blocks;
    parse_lower_blocks(num_blocks, col_ptr, row_idx, values, rhs, x, S, blocks);
    const int32_t nb = S.nb;
    std::vector<int32_t> mv_ptr, mv_ent, diag_slot, mv_row;
    build_pcg_pattern(nb, S.s_ptr, S.s_col, mv_ptr, mv_ent, diag_slot, mv_row);
    require_device();
    DeviceBufs bufs;
    DevPcg g{};
    g.nb = nb;
    g.mv_ptr = bufs.up(mv_ptr);
    g.mv_ent = bufs.up(mv_ent);
    g.n_full = mv_ptr[nb];
    g.frow = bufs.up(mv_row);
    g.fbsr = bufs.alloc<double>(size_t(36) * g.n_full);
    g.fcol = bufs.alloc<int32_t>(size_t(g.n_full));
    g.diag_slot = bufs.up(diag_slot);
    g.bsr = bufs.up(blocks);
    g.minv = bufs.alloc<double>(size_t(36) * nb);
    g.x = bufs.alloc<double>(size_t(6) * nb);
    g.r = bufs.alloc<double>(size_t(6) * nb);
    g.z = bufs.alloc<double>(size_t(6) * nb);
    g.p = bufs.alloc<double>(size_t(6) * nb);
    g.q = bufs.alloc<double>(size_t(6) * nb);
    g.part = bufs.alloc<double>(size_t(pcg_partial_len(nb)));
    g.state = bufs.alloc<PcgState>(1);
    g.n_vec_blocks = pcg_vec_blocks(nb);
    g.n_mv_blocks = pcg_mv_blocks(nb);
    std::vector<double> b(rhs, rhs + 6 * num_blocks);
    double* d_rhs = bufs.up(b);
    launch_pcg_setup(nullptr, g);
    launch_pcg(nullptr, g, d_rhs, relative_tolerance,
               int32_t(std::min<int64_t>(max_iterations, INT32_MAX)), nullptr, nullptr, nullptr);
    SGPU_HIP(hipGetLastError());
    SGPU_HIP(hipDeviceSynchronize());
    PcgState st;
    SGPU_HIP(hipMemcpy(&st, g.state, sizeof(st), hipMemcpyDeviceToHost));
    const int32_t it = st.iters, bad = st.bad;
    if (iterations) *iterations = it;
    if (bad) throw Error(SGPU_ERR_NON_FINITE, "PCG breakdown (p.Sp <= 0)");
    SGPU_HIP(hipMemcpy(x, g.x, size_t(6) * nb * 8, hipMemcpyDeviceToHost));
  });
}

sgpu_status sgpu_precond_features(int64_t num_blocks, const int64_t* col_ptr, const int64_t* row_idx,
                                  const double* values, float* features_out) {
  return guarded([&] {
    if (!features_out) throw Error(SGPU_ERR_INVALID_ARGUMENT, "features_out required");
    Symbolic S;
    std::vector<double> blocks;
    std::vector<double> dummy(6 * std::max<int64_t>(num_blocks, 1), 0.0);
    parse_lower_blocks(num_blocks, col_ptr, row_idx, values, dummy.data(), dummy.data(), S, blocks);
    const int32_t nb = S.nb;
    std::vector<int32_t> mv_ptr, mv_ent, diag_slot, mv_row;
    build_pcg_pattern(nb, S.s_ptr, S.s_col, mv_ptr, mv_ent, diag_slot, mv_row);
    require_device();
    DeviceBufs bufs;
    DevPcg g{};
    g.nb = nb;
    g.mv_ptr = bufs.up(mv_ptr);
    g.mv_ent = bufs.up(mv_ent);
    g.n_full = mv_ptr[nb];
    g.frow = bufs.up(mv_row);
    g.fbsr = bufs.alloc<double>(size_t(36) * g.n_full);
    g.fcol = bufs.alloc<int32_t>(size_t(g.n_full));
    g.diag_slot = bufs.up(diag_slot);
    g.bsr = bufs.up(blocks);
    g.minv = bufs.alloc<double>(size_t(36) * nb);
    float* feat = bufs.alloc<float>(size_t(kPcFeat) * nb);
    launch_pcg_setup(nullptr, g);
    launch_precond_features(nullptr, g, feat);
    SGPU_HIP(hipGetLastError());
    SGPU_HIP(hipDeviceSynchronize());
    SGPU_HIP(hipMemcpy(features_out, feat, size_t(kPcFeat) * nb * 4, hipMemcpyDeviceToHost));
  });
}
