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

// An equality-constraint set (EqConstraintFn): translation-norm terms on pose members, sorted
// by member; all terms share one constraint row and multiplier (eq_system.rs:33-37,207-262).
struct EqSet {
  std::vector<int32_t> member;
  std::vector<double> radius;
};

// SE(3) between factors, or (prior) Isometry3 priors: one pose per term (n = m), the prior
// mean in meas and the precision in prec.
struct PoseGraphFn {
  int robust = 0;
  double delta = 0;
  bool prior = false;
  std::vector<int32_t> m, n;      // pose member indices, sorted by (m, n)
  std::vector<double> meas;       // 7 per term
  std::vector<double> prec;       // priors: 36 per term (row-major W)
  std::vector<uint8_t> first;     // 1 if the term opens a reduction range
};

// Validated problem: VarFamilies layout (var_families.rs:60-113), cost terms sorted
// by CostFn::sort (cost_fn.rs:354-393).
struct HostProblem {
  std::vector<Family> fams;          // sorted by name (BTreeMap order)
  std::vector<int> sorted_of_input;  // caller family index -> sorted index
  int f_pose = -1, f_point = -1, f_intr = -1;
  bool has_reproj = false;
  // camera model of the reprojection terms (0 pinhole, 1 EUCM) and its parameter count
  int cam_model = 0, intr_params = 4;
  // kind of the intrinsics family; with a Free family, its one active member (a calibration
  // shared by the observations) is the free block intr_block = num_pose_blocks, after the
  // pose blocks
  int intr_kind = SGPU_VAR_CONDITIONED;
  int32_t intr_member = -1, intr_block = -1;
  int64_t num_pose_blocks = 0;
  int robust = 0;
  double delta = 0;
  // reprojection terms sorted by CostFn::sort's key: (point, pose, intrinsics), or (point,
  // intrinsics, pose) when the intrinsics family is Free
  std::vector<int32_t> obs_pose, obs_point, obs_intr;
  std::vector<double> obs_uv;
  std::vector<PoseGraphFn> pg;
  std::vector<EqSet> eqs;  // non-empty equality-constraint sets on the pose family
  bool terms_sorted = true;  // false: the reprojection terms are still in the caller's order
  int64_t num_free_blocks = 0, num_marg_blocks = 0;
  std::vector<int32_t> cam_member;  // free block -> pose member (-1: the intrinsics block)
  // free block -> its index in the reference's layout (Free families in name order, active
  // members in member order; var_families.rs:60-113)
  std::vector<int32_t> ref_block;
};

// Builds and validates; throws sgpu::Error with the reference's error variant.
HostProblem build_host_problem(const sgpu_var_family* families, int64_t num_families,
                               const sgpu_cost_fn* cost_fns, int64_t num_cost_fns,
                               const sgpu_eq_fn* eq_fns = nullptr, int64_t num_eq_fns = 0);
// build_host_problem with the term sort optional (sort_terms = false keeps the caller's order
// for a sort elsewhere, applied with apply_term_order: ord[sorted] = caller's index).
HostProblem build_host_problem_sorted(const sgpu_var_family* families, int64_t num_families,
                                      const sgpu_cost_fn* cost_fns, int64_t num_cost_fns,
                                      const sgpu_eq_fn* eq_fns, int64_t num_eq_fns,
                                      bool sort_terms);
void apply_term_order(HostProblem& P, const std::vector<int32_t>& ord);
// Most equality-constraint sets one handle takes (the KKT block is solved densely per step).
constexpr int32_t kMaxEqSets = 16;

// Symbolic structure of one problem (computed once per solve).
constexpr int32_t kOwnChunk = 256;  // camera-major observations per own-term workgroup

struct Symbolic {
  int32_t nb = 0;  // free blocks of the RCS: the poses', then the intrinsics' (if free)
  // trailing blocks coupled to (nearly) every other block: kept out of the nested
  // dissection and eliminated last
  int32_t n_dense_tail = 0;
  int32_t npt = 0, nobs = 0, npose = 0;
  // point-major observations
  std::vector<int32_t> pt_ptr;      // npt + 1
  std::vector<int32_t> pt_nu;       // reduced reprojection terms per point member
  std::vector<int32_t> cam_nu;      // reduced terms per free block (the intrinsics' included)
  // free intrinsics: per pose block, the S slot of (intrinsics, pose) or -1; the slot of
  // (intrinsics, intrinsics)
  std::vector<int32_t> intr_slot;
  int32_t intr_diag_slot = -1;
  std::vector<int32_t> pose_block;  // pose member -> free block or -1
  // camera-major observation lists (free cameras only)
  std::vector<int32_t> cam_ptr, cam_obs;
  // RCS S pattern in camera order: row c, columns <= c, sorted
  std::vector<int32_t> s_ptr, s_col;
  // per S slot, the cross pairs (obs_i, obs_j), i != j, of one active point with
  // cam(obs_i) = row and cam(obs_j) = column, ordered by point, stored as positions in
  // cam_obs; the camera's own (o, o) terms are summed over chunks of its cam_obs range
  std::vector<int64_t> pair_ptr;
  std::vector<int32_t> pairs;  // 2 per pair
  std::vector<int32_t> own_chunks;     // (camera, q0, q1) per chunk of <= kOwnChunk
  std::vector<int32_t> cam_chunk_ptr;  // nb + 1: chunks of each camera
  // ordering (perm[new] = old) and column elimination tree
  std::vector<int32_t> perm, perm_inv;
  std::vector<int32_t> etree_parent;
  int64_t nnz_l = 0;  // strictly-lower L blocks (column structure)
  int32_t max_slots_per_row = 0;
  double factor_flops = 0;
  int64_t num_schur_pairs = 0;
  // destination of each S slot in the supernodal fronts: element offset of the block's
  // (0,0) entry, and leading dimension * 2 + transpose flag
  std::vector<int64_t> s_off;
  std::vector<int32_t> s_ld;
  // supernodes: contiguous column chains of the elimination tree
  int32_t nsn = 0;
  std::vector<int32_t> sn_c0, sn_nc, sn_nf, sn_parent, sn_level;
  std::vector<int64_t> sn_off, sn_loff;       // front / L-panel element offsets
  std::vector<int32_t> sn_rows_ptr, sn_rows;  // front row blocks (cols first, then ancestors)
  std::vector<int32_t> sn_upos_ptr, sn_upos;  // update-row blocks -> position in parent front
  std::vector<int32_t> blk_sn;                // permuted block -> supernode
  int64_t front_elems = 0, l_elems = 0;
  int32_t max_front = 0;                      // largest front (blocks)
  std::vector<int32_t> lvl_ptr, lvl_sn;
  std::vector<int32_t> sn_child_ptr, sn_child;  // children of each supernode (nsn + 1, nsn - roots)
  // Schedule of the numeric factorization in rounds, one launch (or a panel-row launch and
  // a trailing-update launch) per round, each holding one panel step of every node that
  // is ready, so nodes of different tree levels share a launch and the number of rounds
  // follows the elimination tree's critical path.  The leaves' first diagonal blocks are
  // factored before the first round.  A child's last step adds its update matrix (and its
  // update rows of the right-hand side) into the parent's front; children of one parent
  // end in distinct rounds, so those adds are plain and their order is fixed.  The last
  // of them also factors the parent's first diagonal block (in its tile (0, 0), the only
  // tile whose update reaches that block), and the parent's steps follow from the next
  // round.
  struct PanelLaunch {
    int32_t tile_begin, tile_count;  // tiles in launch_tiles
    int32_t item_begin, item_count;  // split rounds: panel-row items
    bool fused;
    double flops;  // algorithmic FP64 flops of the trailing updates
    double bytes;  // algorithmic HBM bytes: lower trailing RMW, panel columns read, L written
  };
  std::vector<PanelLaunch> launches;
  // 4 per tile: (node, panel, tile_row, tile_col), a 64x64 tile of the lower trailing
  // region below the node's panel
  std::vector<int32_t> launch_tiles;
  std::vector<int32_t> panel_items;    // (node, panel, kPanelRows-row chunk) (split rounds)
  std::vector<int32_t> leaves;         // nodes without children
  std::vector<int32_t> sn_hand;        // 1: the node's last step factors its parent's first block
  std::vector<int32_t> sn_last_round;  // round of each node's last panel step (-1: none)
  std::vector<int64_t> sn_woff;       // W scratch offsets (6 nf x kPW per node)
  int64_t w_elems = 0;
  std::vector<int64_t> sn_toff;       // row-major copies of each node's L_KK (K x K)
  int64_t t_elems = 0;
  std::vector<int64_t> sn_poff;       // per node: its panels' L_PP^-T (kPW x kPW each)
  int64_t p_elems = 0;
  // backward-solve work items per level: (node, kSolveCols-column tile of K)
  std::vector<int32_t> solve_cols_ptr, solve_cols;
  std::vector<int32_t> lvl_max_nc;  // largest node (column blocks) per level
  // forward-solve update vectors: node x owns 6 * sn_nf[x] entries at sn_uoff[x]
  std::vector<int64_t> sn_uoff;
  int64_t u_elems = 0;
  // pose-graph terms: per term the S slots of (m,m), (n,n) and (max,min), -1 if absent;
  // mn_flip = 1 when the stored (max,min) block is the transpose of J_m^T J_n
  struct PgDest {
    int32_t mm, nn, mn, mn_flip;
  };
  std::vector<std::vector<PgDest>> pg_dest;
};

Symbolic analyze(const HostProblem& P);
// The steps of analyze(): point-major arrays, reduced-term counts and camera-major lists; the
// S pattern entries that do not come from shared points (diagonal, pose-graph pairs, the free
// intrinsics' row) as (row, column) with column <= row; the own-term chunks; and, once the
// pattern and the pairs are in S, the intrinsics' slots, the factor structure (analyze_rcs)
// and the pose-graph destinations.
void analyze_points(const HostProblem& P, Symbolic& S);
std::vector<std::pair<int32_t, int32_t>> pattern_extras(const HostProblem& P, const Symbolic& S);
void analyze_chunks(Symbolic& S);
void analyze_finish(const HostProblem& P, Symbolic& S);

// Ordering, elimination tree, supernodes, fronts and the launch schedule of the RCS
// whose lower pattern is S.nb, S.s_ptr, S.s_col (camera order).
void analyze_rcs(Symbolic& S);

// Nested-dissection ordering of an undirected graph (adjacency lists, no self
// loops).  Returns perm[new] = old.
std::vector<int32_t> nested_dissection(int32_t n, const std::vector<std::vector<int32_t>>& adj,
                                       int32_t leaf_size);

// Panel width (in 6x6 pivot blocks) and trailing-update tile edge (scalars) of the
// multifrontal factorization.
constexpr int32_t kPanelBlocks = 8;
constexpr int32_t kTile = 64;
// rows per panel-row workgroup; a round with at least kSplitTiles tiles is launched split
// (panel rows, then the trailing updates) instead of fused
constexpr int32_t kPanelRows = 64;
constexpr int32_t kSplitTiles = 640;
// backward-solve column tile of the rows below a node
constexpr int32_t kSolveCols = 16;

}  // namespace sgpu
