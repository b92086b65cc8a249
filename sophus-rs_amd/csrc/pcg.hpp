// Block-Jacobi preconditioned conjugate gradients on the reduced camera system.
//
// S is kept as its lower 6x6 blocks in slot order (row-major blocks, `bsr`); the
// matrix-vector product walks, per block row, the lower blocks of that row and the
// transposes of the lower blocks below the diagonal in that column (`mv_ptr`/`mv_ent`).
// Every reduction is a fixed-order sum of per-workgroup partials, recomputed by each
// workgroup that needs it, so a solve is deterministic.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace sgpu {

struct PcgState {
  int32_t done;   // 1 once converged, broken down or out of iterations
  int32_t iters;  // iterations taken
  int32_t bad;    // breakdown (p.q <= 0 or non-finite)
  int32_t pad;
};

struct DevPcg {
  int32_t nb;
  const int32_t* mv_ptr;  // nb + 1
  const int32_t* mv_ent;  // (slot, col | trans << 30) per entry
  const int32_t* diag_slot;  // nb
  double* bsr;            // nslot x 36 lower blocks, row-major
  double* minv;           // nb x 36 preconditioner blocks (symmetric)
  double* x;
  double* r;
  double* z;
  double* p;
  double* q;
  double* part;           // partial sums, see pcg.hip
  PcgState* state;
  int32_t n_vec_blocks;   // workgroups of the vector kernels
  int32_t n_mv_blocks;    // workgroups of the matvec kernels
};

// Host-side pattern of the product: for every block row, its lower slots and the
// transposed slots of its column.
void build_pcg_pattern(int32_t nb, const std::vector<int32_t>& s_ptr,
                       const std::vector<int32_t>& s_col, std::vector<int32_t>& mv_ptr,
                       std::vector<int32_t>& mv_ent, std::vector<int32_t>& diag_slot);

// Sizes of the partial-sum buffer and the launch grids.
int32_t pcg_vec_blocks(int32_t nb);
int32_t pcg_mv_blocks(int32_t nb);
int64_t pcg_partial_len(int32_t nb);

// Block-Jacobi inverse of every diagonal block (PSD pivot rule; zero pivots -> 0).
void launch_pcg_block_jacobi(hipStream_t s, const DevPcg& g);

// Solve S x = b to |r| <= rel_tol |b| or max_iter iterations (x starts at 0).  The host
// checks the device state every `check_every` iterations; returns the iterations taken
// and sets *bad on breakdown.
int32_t run_pcg(hipStream_t s, const DevPcg& g, const double* b, double rel_tol, int32_t max_iter,
                PcgState* h_state, int32_t* bad);

}  // namespace sgpu
