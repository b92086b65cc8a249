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
  int64_t n_full;         // entries of the full block rows (= mv_ptr[nb])
  double* fbsr;           // n_full x 36 full block rows, row-major blocks
  int32_t* fcol;          // n_full block columns
  const int32_t* frow;    // n_full block rows
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
  int32_t n_vec_blocks;   // workgroups of the per-block kernels
  int32_t n_mv_blocks;    // wavefront groups of the product
};

// Host-side pattern of the product: for every block row, its lower slots and the
// transposed slots of its column.
void build_pcg_pattern(int32_t nb, const std::vector<int32_t>& s_ptr,
                       const std::vector<int32_t>& s_col, std::vector<int32_t>& mv_ptr,
                       std::vector<int32_t>& mv_ent, std::vector<int32_t>& diag_slot,
                       std::vector<int32_t>& mv_row);

// Sizes of the partial-sum buffer and the launch grids.
int32_t pcg_vec_blocks(int32_t nb);
int32_t pcg_mv_blocks(int32_t nb);
int64_t pcg_partial_len(int32_t nb);

// Per linearization: block-Jacobi inverse of every diagonal block (PSD pivot rule; zero
// pivots -> 0) and the full block rows of S.
void launch_pcg_setup(hipStream_t s, const DevPcg& g);

// Solve S x = b (x starts at 0) to |r| <= rel_tol |b| or max_iter iterations: two
// launches per iteration.  With a pinned `h_state` the host polls the device every 16
// iterations and stops launching once it is done (a stream synchronization per poll);
// without, all max_iter iterations are enqueued (finished iterations exit at once).
// The iterations taken go to *iters_out and a breakdown (p.Sp <= 0 or non-finite) sets
// *bad_out = 1 (device words, may be null).
void launch_pcg(hipStream_t s, const DevPcg& g, const double* b, double rel_tol, int32_t max_iter,
                int32_t* iters_out, int32_t* bad_out, PcgState* h_state);

}  // namespace sgpu
