// Host-side problem description and symbolic analysis of the Schur BA path.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "api_internal.hpp"

namespace sgpu {

struct Family {
  std::string name;
  int type = 0, kind = 0, dof = 0, nparams = 0;
  int64_t n = 0;
  std::vector<double> members;
  std::vector<char> is_const;
  std::vector<int32_t> block;  // active block index within its kind, -1 if none
  int64_t num_active = 0;
};

struct PoseGraphFn {
  int robust = 0;
  double delta = 0;
  std::vector<int32_t> m, n;      // pose member indices, sorted by (m, n)
  std::vector<double> meas;       // 7 per term
  std::vector<uint8_t> first;     // 1 if the term opens a reduction range
};

// Validated problem: VarFamilies layout (var_families.rs:60-113), cost terms sorted
// by CostFn::sort (cost_fn.rs:354-393).
struct HostProblem {
  std::vector<Family> fams;          // sorted by name (BTreeMap order)
  std::vector<int> sorted_of_input;  // caller family index -> sorted index
  int f_pose = -1, f_point = -1, f_intr = -1;
  bool has_reproj = false;
  int robust = 0;
  double delta = 0;
  // reprojection terms sorted by (point, pose, intrinsics)
  std::vector<int32_t> obs_pose, obs_point, obs_intr;
  std::vector<double> obs_uv;
  std::vector<PoseGraphFn> pg;
  int64_t num_free_blocks = 0, num_marg_blocks = 0;
  std::vector<int32_t> cam_member;  // free block -> pose member
};

// Builds and validates; throws sgpu::Error with the reference's error variant.
HostProblem build_host_problem(const sgpu_var_family* families, int64_t num_families,
                               const sgpu_cost_fn* cost_fns, int64_t num_cost_fns);

// Symbolic structure of one problem (computed once per solve).
struct Symbolic {
  int32_t nb = 0;  // free camera blocks
  int32_t npt = 0, nobs = 0, npose = 0;
  // point-major observations
  std::vector<int32_t> pt_ptr;      // npt + 1
  std::vector<int32_t> pt_nu;       // reduced reprojection terms per point member
  std::vector<int32_t> cam_nu;      // reduced reprojection terms per free block
  std::vector<int32_t> pose_block;  // pose member -> free block or -1
  // camera-major observation lists (free cameras only)
  std::vector<int32_t> cam_ptr, cam_obs;
  // RCS S pattern in camera order: row c, columns <= c, sorted
  std::vector<int32_t> s_ptr, s_col;
  std::vector<int32_t> s_dest;  // per S slot: unified slot * 2 + transpose flag
  // ordering (perm[new] = old)
  std::vector<int32_t> perm, perm_inv;
  // L structure in the permuted order
  std::vector<int32_t> l_ptr, l_row;  // CSC, rows > column
  std::vector<int32_t> r_ptr, r_col, r_lpos;  // reach of each column: (k, L index of L_jk)
  // gather-form update lists over unified slots (0..nb-1 diagonal, nb + L index)
  std::vector<int32_t> u_ptr, u_src, u_m;
  std::vector<int32_t> lvl_ptr, lvl_col;
  int32_t max_slots_per_row = 0;
  double factor_flops = 0;
  int64_t num_schur_pairs = 0;
  // pose-graph terms: per term the unified destinations of (m,m), (n,n), (m,n)
  struct PgDest {
    int32_t mm, nn, mn;  // unified slot * 2 + transpose, -1 if absent
  };
  std::vector<std::vector<PgDest>> pg_dest;
};

Symbolic analyze(const HostProblem& P);

// Ordering, elimination tree, L structure, update lists and levels of the RCS whose
// lower pattern is S.nb, S.s_ptr, S.s_col (camera order); fills the remaining fields.
void analyze_rcs(Symbolic& S);

// Nested-dissection ordering of an undirected graph (adjacency lists, no self
// loops).  Returns perm[new] = old.
std::vector<int32_t> nested_dissection(int32_t n, const std::vector<std::vector<int32_t>>& adj,
                                       int32_t leaf_size);

}  // namespace sgpu
