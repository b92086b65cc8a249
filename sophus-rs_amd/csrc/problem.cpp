// Problem validation, variable layout, term sorting and symbolic analysis.
#include "problem.hpp"

#include <algorithm>
#include <cstring>
#include <functional>
#include <limits>
#include <numeric>
#include <queue>

namespace sgpu {
namespace {

int var_dof(int type) {
  return type == SGPU_VAR_ISOMETRY3 || type == SGPU_VAR_EUCM ? 6 : type == SGPU_VAR_VEC3 ? 3 : 4;
}
int var_params(int type) {
  return type == SGPU_VAR_ISOMETRY3 ? 7 : type == SGPU_VAR_VEC3 ? 3 : type == SGPU_VAR_EUCM ? 6 : 4;
}
bool is_reprojection(int kind) {
  return kind == SGPU_COST_PINHOLE_REPROJECTION || kind == SGPU_COST_EUCM_REPROJECTION;
}

[[noreturn]] void raise(sgpu_status s, const std::string& m) { throw Error(s, m); }

}  // namespace

HostProblem build_host_problem(const sgpu_var_family* families, int64_t num_families,
                               const sgpu_cost_fn* cost_fns, int64_t num_cost_fns,
                               const sgpu_eq_fn* eq_fns, int64_t num_eq_fns) {
  return build_host_problem_sorted(families, num_families, cost_fns, num_cost_fns, eq_fns,
                                   num_eq_fns, true);
}

HostProblem build_host_problem_sorted(const sgpu_var_family* families, int64_t num_families,
                                      const sgpu_cost_fn* cost_fns, int64_t num_cost_fns,
                                      const sgpu_eq_fn* eq_fns, int64_t num_eq_fns,
                                      bool sort_terms) {
  HostProblem P;
  if (!families || num_families <= 0) raise(SGPU_ERR_INVALID_ARGUMENT, "no variable families");
  if (!cost_fns || num_cost_fns <= 0) raise(SGPU_ERR_INVALID_ARGUMENT, "no cost functions");
  std::vector<Family> in(num_families);
  for (int64_t i = 0; i < num_families; ++i) {
    const sgpu_var_family& s = families[i];
    if (!s.name || s.type < 0 || s.type > 3 || s.kind < 0 || s.kind > 2 || s.num_members < 0 ||
        (s.num_members > 0 && !s.members) || (s.num_const_ids > 0 && !s.const_ids))
      raise(SGPU_ERR_INVALID_ARGUMENT, "bad variable family descriptor");
    if (s.num_members > INT32_MAX / 8) raise(SGPU_ERR_UNSUPPORTED, "family too large");
    Family& f = in[i];
    f.name = s.name;
    f.type = s.type;
    f.kind = s.kind;
    f.dof = var_dof(s.type);
    f.nparams = var_params(s.type);
    f.n = s.num_members;
    f.members.assign(s.members, s.members + f.n * f.nparams);
    f.is_const.assign(f.n, 0);
    for (int64_t k = 0; k < s.num_const_ids; ++k) {
      if (s.const_ids[k] < 0 || s.const_ids[k] >= f.n)
        raise(SGPU_ERR_INDEX_OUT_OF_RANGE, "const id out of range in family '" + f.name + "'");
      f.is_const[s.const_ids[k]] = 1;
    }
  }
  std::vector<int> order(num_families);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](int a, int b) { return in[a].name < in[b].name; });
  for (size_t k = 1; k < order.size(); ++k)
    if (in[order[k]].name == in[order[k - 1]].name)
      raise(SGPU_ERR_INVALID_ARGUMENT, "duplicate family name '" + in[order[k]].name + "'");
  P.sorted_of_input.assign(num_families, 0);
  for (int k = 0; k < num_families; ++k) {
    P.sorted_of_input[order[k]] = k;
    P.fams.push_back(std::move(in[order[k]]));
  }
  // layout: active members of Free / Marginalized families get block indices
  int64_t free_scalars = 0, marg_scalars = 0;
  for (auto& f : P.fams) {
    f.block.assign(f.n, -1);
    if (f.kind == SGPU_VAR_CONDITIONED) continue;
    int32_t b = 0;
    for (int64_t i = 0; i < f.n; ++i)
      if (!f.is_const[i]) f.block[i] = b++;
    f.num_active = b;
    (f.kind == SGPU_VAR_FREE ? free_scalars : marg_scalars) += int64_t(b) * f.dof;
  }
  auto find = [&](const char* name) -> int {
    for (size_t i = 0; i < P.fams.size(); ++i)
      if (P.fams[i].name == name) return int(i);
    return -1;
  };
  // validate_solver_config (optimizer.rs:649-717)
  if (free_scalars == 0 || marg_scalars == 0)
    raise(SGPU_ERR_SCHUR_MISSING_FREE_OR_MARG,
          "Schur solver requires at least one active free and one active marginalized variable");
  for (int64_t i = 0; i < num_cost_fns; ++i) {
    const sgpu_cost_fn& c = cost_fns[i];
    const int arity = is_reprojection(c.kind) ? 3 : c.kind == SGPU_COST_ISOMETRY3_PRIOR ? 1 : 2;
    int marg = 0;
    for (int a = 0; a < arity; ++a) {
      if (!c.family_names[a]) raise(SGPU_ERR_INVALID_ARGUMENT, "missing cost family name");
      const int fi = find(c.family_names[a]);
      if (fi >= 0 && P.fams[fi].kind == SGPU_VAR_MARGINALIZED) ++marg;
    }
    if (marg > 1)
      raise(SGPU_ERR_SCHUR_COST_CONNECTS_TWO_MARG_FAMILIES,
            "cost term connects two marginalized families");
  }
  for (int64_t i = 0; i < num_cost_fns; ++i) {
    const sgpu_cost_fn& c = cost_fns[i];
    if (!is_reprojection(c.kind) && c.kind != SGPU_COST_POSE_GRAPH3 && c.kind != SGPU_COST_ISOMETRY3_PRIOR)
      raise(SGPU_ERR_INVALID_ARGUMENT, "unknown cost kind");
    if (c.robust_kernel != SGPU_ROBUST_NONE && c.robust_kernel != SGPU_ROBUST_HUBER)
      raise(SGPU_ERR_INVALID_ARGUMENT, "unknown robust kernel");
    const int arity = is_reprojection(c.kind) ? 3 : c.kind == SGPU_COST_ISOMETRY3_PRIOR ? 1 : 2;
    const int want[4][3] = {{SGPU_VAR_PINHOLE, SGPU_VAR_ISOMETRY3, SGPU_VAR_VEC3},
                            {SGPU_VAR_ISOMETRY3, SGPU_VAR_ISOMETRY3, -1},
                            {SGPU_VAR_EUCM, SGPU_VAR_ISOMETRY3, SGPU_VAR_VEC3},
                            {SGPU_VAR_ISOMETRY3, -1, -1}};
    int fi[3] = {-1, -1, -1};
    for (int a = 0; a < arity; ++a) {
      fi[a] = find(c.family_names[a]);
      if (fi[a] < 0)
        raise(SGPU_ERR_UNKNOWN_FAMILY, std::string("unknown family '") + c.family_names[a] + "'");
      if (P.fams[fi[a]].type != want[c.kind][a])
        raise(SGPU_ERR_FAMILY_TYPE, std::string("family '") + c.family_names[a] +
                                        "' has the wrong variable type");
    }
    if (c.num_terms <= 0 || !c.entity_indices || !c.measurements)
      raise(SGPU_ERR_INVALID_ARGUMENT, "empty cost function");
    if (c.num_terms > INT32_MAX / 4) raise(SGPU_ERR_UNSUPPORTED, "too many terms");
    for (int64_t t = 0; t < c.num_terms; ++t)
      for (int a = 0; a < arity; ++a) {
        const int64_t v = c.entity_indices[t * arity + a];
        if (v < 0 || v >= P.fams[fi[a]].n)
          raise(SGPU_ERR_INDEX_OUT_OF_RANGE, "entity index out of range");
      }
    // The GPU path implements the BA layout: SE(3) poses Free, points Marginalized, and
    // intrinsics Conditioned, or (EUCM) Free with one active member.
    if (is_reprojection(c.kind)) {
      if (P.has_reproj) raise(SGPU_ERR_UNSUPPORTED, "more than one reprojection cost function");
      const Family& kf = P.fams[fi[0]];
      if (kf.kind == SGPU_VAR_MARGINALIZED || P.fams[fi[1]].kind != SGPU_VAR_FREE ||
          P.fams[fi[2]].kind != SGPU_VAR_MARGINALIZED)
        raise(SGPU_ERR_UNSUPPORTED,
              "reprojection terms need Conditioned or Free intrinsics, Free poses and "
              "Marginalized points");
      if (kf.kind == SGPU_VAR_FREE && c.kind == SGPU_COST_PINHOLE_REPROJECTION && kf.num_active > 0)
        raise(SGPU_ERR_UNSUPPORTED,
              "free pinhole intrinsics are not implemented (a 4-dof block in the RCS of 6x6 "
              "blocks); free EUCM intrinsics are");
      if (kf.kind == SGPU_VAR_FREE && kf.num_active > 1)
        raise(SGPU_ERR_UNSUPPORTED,
              "more than one active member of a free intrinsics family (one shared calibration "
              "block is implemented)");
      if ((P.f_pose >= 0 && P.f_pose != fi[1]))
        raise(SGPU_ERR_UNSUPPORTED, "cost functions use different pose families");
      P.has_reproj = true;
      P.f_intr = fi[0];
      P.f_pose = fi[1];
      P.f_point = fi[2];
      P.cam_model = c.kind == SGPU_COST_EUCM_REPROJECTION ? 1 : 0;
      P.intr_params = kf.nparams;
      P.intr_kind = kf.kind;
      if (kf.kind == SGPU_VAR_FREE)
        for (int64_t i = 0; i < kf.n; ++i)
          if (kf.block[i] >= 0) P.intr_member = int32_t(i);
      P.robust = c.robust_kernel;
      P.delta = c.huber_delta;
      // CostFn::sort (compare_idx.rs:38-128): stable sort by the marginalized argument,
      // then the free ones in argument order, then the conditioned one
      const bool intr_free = kf.kind == SGPU_VAR_FREE;
      const int64_t nt = c.num_terms;
      std::vector<int32_t> ord(nt);
      std::iota(ord.begin(), ord.end(), 0);
      const int64_t* E = c.entity_indices;
      const int k1 = intr_free ? 0 : 1, k2 = intr_free ? 1 : 0;
      P.terms_sorted = sort_terms;
      if (sort_terms)
        std::stable_sort(ord.begin(), ord.end(), [&](int32_t x, int32_t y) {
          if (E[3 * x + 2] != E[3 * y + 2]) return E[3 * x + 2] < E[3 * y + 2];
          if (E[3 * x + k1] != E[3 * y + k1]) return E[3 * x + k1] < E[3 * y + k1];
          return E[3 * x + k2] < E[3 * y + k2];
        });
      P.obs_pose.resize(nt);
      P.obs_point.resize(nt);
      P.obs_intr.resize(nt);
      P.obs_uv.resize(2 * nt);
      for (int64_t t = 0; t < nt; ++t) {
        const int32_t o = ord[t];
        P.obs_intr[t] = int32_t(E[3 * o]);
        P.obs_pose[t] = int32_t(E[3 * o + 1]);
        P.obs_point[t] = int32_t(E[3 * o + 2]);
        P.obs_uv[2 * t] = c.measurements[2 * o];
        P.obs_uv[2 * t + 1] = c.measurements[2 * o + 1];
      }
    } else if (c.kind == SGPU_COST_ISOMETRY3_PRIOR) {
      // Isometry3PriorCostTerm (isometry3_prior.rs:46-88) on the pose family; terms sorted by
      // member (CostFn::sort), repeated members reduced into one term (one nu)
      if (P.fams[fi[0]].kind != SGPU_VAR_FREE)
        raise(SGPU_ERR_UNSUPPORTED, "Isometry3 priors need the Free SE(3) pose family");
      if (P.f_pose >= 0 && P.f_pose != fi[0])
        raise(SGPU_ERR_UNSUPPORTED, "cost functions use different pose families");
      P.f_pose = fi[0];
      PoseGraphFn g;
      g.prior = true;
      g.robust = c.robust_kernel;
      g.delta = c.huber_delta;
      const int64_t nt = c.num_terms;
      std::vector<int32_t> ord(nt);
      std::iota(ord.begin(), ord.end(), 0);
      const int64_t* E = c.entity_indices;
      std::stable_sort(ord.begin(), ord.end(), [&](int32_t x, int32_t y) { return E[x] < E[y]; });
      for (int64_t t = 0; t < nt; ++t) {
        const int32_t o = ord[t];
        g.m.push_back(int32_t(E[o]));
        g.n.push_back(int32_t(E[o]));
        for (int k = 0; k < 7; ++k) g.meas.push_back(c.measurements[43 * int64_t(o) + k]);
        for (int k = 0; k < 36; ++k) g.prec.push_back(c.measurements[43 * int64_t(o) + 7 + k]);
        g.first.push_back(t == 0 || g.m[t] != g.m[t - 1]);
      }
      P.pg.push_back(std::move(g));
    } else {
      if (P.fams[fi[0]].kind != SGPU_VAR_FREE || fi[0] != fi[1])
        raise(SGPU_ERR_UNSUPPORTED, "pose-graph terms need one Free SE(3) family");
      if (P.f_pose >= 0 && P.f_pose != fi[0])
        raise(SGPU_ERR_UNSUPPORTED, "cost functions use different pose families");
      P.f_pose = fi[0];
      PoseGraphFn g;
      g.robust = c.robust_kernel;
      g.delta = c.huber_delta;
      const int64_t nt = c.num_terms;
      std::vector<int32_t> ord(nt);
      std::iota(ord.begin(), ord.end(), 0);
      const int64_t* E = c.entity_indices;
      std::stable_sort(ord.begin(), ord.end(), [&](int32_t x, int32_t y) {
        if (E[2 * x] != E[2 * y]) return E[2 * x] < E[2 * y];
        return E[2 * x + 1] < E[2 * y + 1];
      });
      for (int64_t t = 0; t < nt; ++t) {
        const int32_t o = ord[t];
        if (E[2 * o] == E[2 * o + 1])
          raise(SGPU_ERR_UNSUPPORTED, "pose-graph term connects a pose to itself");
        g.m.push_back(int32_t(E[2 * o]));
        g.n.push_back(int32_t(E[2 * o + 1]));
        for (int k = 0; k < 7; ++k) g.meas.push_back(c.measurements[7 * o + k]);
        g.first.push_back(t == 0 || g.m[t] != g.m[t - 1] || g.n[t] != g.n[t - 1]);
      }
      P.pg.push_back(std::move(g));
    }
  }
  // every Free / Marginalized family must be the pose (or intrinsics) / point family of this
  // layout
  for (size_t i = 0; i < P.fams.size(); ++i) {
    const Family& f = P.fams[i];
    if (f.kind == SGPU_VAR_FREE && int(i) != P.f_pose && int(i) != P.f_intr && f.num_active > 0)
      raise(SGPU_ERR_UNSUPPORTED, "free family '" + f.name + "' is not the pose family");
    if (f.kind == SGPU_VAR_MARGINALIZED && int(i) != P.f_point && f.num_active > 0)
      raise(SGPU_ERR_UNSUPPORTED, "marginalized family '" + f.name + "' is not the point family");
  }
  if (P.f_point < 0) raise(SGPU_ERR_UNSUPPORTED, "no reprojection cost function");
  // equality-constraint sets (EqSystem::new, eq_system.rs:24-48; the Schur check of
  // validate_solver_config, optimizer.rs:649-683)
  if (num_eq_fns < 0 || (num_eq_fns > 0 && !eq_fns))
    raise(SGPU_ERR_INVALID_ARGUMENT, "bad equality-constraint sets");
  for (int64_t i = 0; i < num_eq_fns; ++i) {
    const sgpu_eq_fn& q = eq_fns[i];
    if (q.kind != SGPU_EQ_TRANSLATION_NORM || !q.family_names[0] || q.num_terms < 0 ||
        (q.num_terms > 0 && (!q.entity_indices || !q.params)))
      raise(SGPU_ERR_INVALID_ARGUMENT, "bad equality-constraint set");
    if (q.num_terms == 0) continue;  // a set with no terms has no residual rows
    const int fi = find(q.family_names[0]);
    if (fi < 0)
      raise(SGPU_ERR_UNKNOWN_FAMILY, std::string("unknown family '") + q.family_names[0] + "'");
    const Family& f = P.fams[fi];
    // validate_solver_config's Schur check comes first (optimizer.rs:664-678)
    if (f.kind == SGPU_VAR_MARGINALIZED)
      raise(SGPU_ERR_SCHUR_CONSTRAINT_ON_MARGINALIZED_FAMILY,
            "equality constraint on the marginalized family '" + f.name + "'");
    if (f.type != SGPU_VAR_ISOMETRY3)
      raise(SGPU_ERR_FAMILY_TYPE, "family '" + f.name + "' has the wrong variable type");
    if (fi != P.f_pose)
      raise(SGPU_ERR_UNSUPPORTED, "equality constraints on a family other than the pose family");
    if (q.num_terms > INT32_MAX / 8) raise(SGPU_ERR_UNSUPPORTED, "too many constraint terms");
    std::vector<int32_t> ord(q.num_terms);
    for (int64_t t = 0; t < q.num_terms; ++t) {
      if (q.entity_indices[t] < 0 || q.entity_indices[t] >= f.n)
        raise(SGPU_ERR_INDEX_OUT_OF_RANGE, "constraint entity index out of range");
      ord[t] = int32_t(t);
    }
    std::stable_sort(ord.begin(), ord.end(), [&](int32_t a, int32_t b) {
      return q.entity_indices[a] < q.entity_indices[b];
    });
    EqSet e;
    for (int32_t t : ord) {
      e.member.push_back(int32_t(q.entity_indices[t]));
      e.radius.push_back(q.params[t]);
    }
    P.eqs.push_back(std::move(e));
  }
  if (int64_t(P.eqs.size()) > kMaxEqSets)
    raise(SGPU_ERR_UNSUPPORTED, "more than " + std::to_string(kMaxEqSets) + " equality-constraint sets");
  if (!P.has_reproj) P.terms_sorted = true;
  P.num_pose_blocks = P.fams[P.f_pose].num_active;
  P.intr_block = P.intr_member >= 0 ? int32_t(P.num_pose_blocks) : -1;
  P.num_free_blocks = P.num_pose_blocks + (P.intr_member >= 0 ? 1 : 0);
  P.num_marg_blocks = P.fams[P.f_point].num_active;
  const Family& fp = P.fams[P.f_pose];
  P.cam_member.assign(P.num_free_blocks, -1);
  for (int64_t i = 0; i < fp.n; ++i)
    if (fp.block[i] >= 0) P.cam_member[fp.block[i]] = int32_t(i);
  // the reference's free-block numbering: Free families in name order (fams is sorted)
  P.ref_block.assign(P.num_free_blocks, -1);
  int32_t next = 0;
  for (size_t i = 0; i < P.fams.size(); ++i) {
    if (P.fams[i].kind != SGPU_VAR_FREE) continue;
    if (int(i) == P.f_pose) {
      for (int32_t b = 0; b < int32_t(P.num_pose_blocks); ++b) P.ref_block[b] = next++;
    } else if (int(i) == P.f_intr && P.intr_block >= 0) {
      P.ref_block[P.intr_block] = next++;
    }
  }
  return P;
}

void apply_term_order(HostProblem& P, const std::vector<int32_t>& ord) {
  const size_t n = ord.size();
  std::vector<int32_t> pose(n), point(n), intr(n);
  std::vector<double> uv(2 * n);
  for (size_t t = 0; t < n; ++t) {
    const int32_t o = ord[t];
    pose[t] = P.obs_pose[o];
    point[t] = P.obs_point[o];
    intr[t] = P.obs_intr[o];
    uv[2 * t] = P.obs_uv[2 * size_t(o)];
    uv[2 * t + 1] = P.obs_uv[2 * size_t(o) + 1];
  }
  P.obs_pose.swap(pose);
  P.obs_point.swap(point);
  P.obs_intr.swap(intr);
  P.obs_uv.swap(uv);
  P.terms_sorted = true;
}

// ---- nested dissection ----------------------------------------------------------------

std::vector<int32_t> nested_dissection(int32_t n, const std::vector<std::vector<int32_t>>& adj,
                                       int32_t leaf_size) {
  std::vector<int32_t> out;
  out.reserve(n);
  std::vector<int32_t> mark(n, -1), level(n, -1), loc(n, -1);
  int32_t stamp = 0;
  auto bfs = [&](const std::vector<int32_t>& nodes, int32_t tag, int32_t src,
                 std::vector<int32_t>& orderv) {
    // nodes carry mark == tag; returns BFS order, sets level[]
    orderv.clear();
    std::queue<int32_t> q;
    for (int32_t v : nodes) level[v] = -1;
    level[src] = 0;
    q.push(src);
    while (!q.empty()) {
      int32_t v = q.front();
      q.pop();
      orderv.push_back(v);
      for (int32_t u : adj[v])
        if (mark[u] == tag && level[u] < 0) {
          level[u] = level[v] + 1;
          q.push(u);
        }
    }
  };
  // recursive worker
  std::vector<int32_t> orderv;
  std::function<void(std::vector<int32_t>&)> rec;
  rec = [&](std::vector<int32_t>& nodes) {
    if (nodes.empty()) return;
    std::sort(nodes.begin(), nodes.end());
    if (int32_t(nodes.size()) <= leaf_size) {
      out.insert(out.end(), nodes.begin(), nodes.end());
      return;
    }
    const int32_t tag = ++stamp;
    for (int32_t v : nodes) mark[v] = tag;
    bfs(nodes, tag, nodes[0], orderv);
    if (orderv.size() < nodes.size()) {  // disconnected: split into components
      std::vector<std::vector<int32_t>> comps;
      std::vector<int32_t> rest = nodes;
      while (!rest.empty()) {
        const int32_t t2 = ++stamp;
        for (int32_t v : rest) mark[v] = t2;
        bfs(rest, t2, rest[0], orderv);
        comps.push_back(orderv);
        std::vector<int32_t> nxt;
        for (int32_t v : rest)
          if (level[v] < 0) nxt.push_back(v);
        rest.swap(nxt);
      }
      for (auto& c : comps) rec(c);
      return;
    }
    // pseudo-peripheral start: two BFS sweeps
    int32_t src = orderv.back();
    bfs(nodes, tag, src, orderv);
    src = orderv.back();
    bfs(nodes, tag, src, orderv);
    const int32_t maxlev = level[orderv.back()];
    if (maxlev < 2) {  // (nearly) a clique: no useful separator
      out.insert(out.end(), nodes.begin(), nodes.end());
      return;
    }
    const int32_t N = int32_t(nodes.size());
    std::vector<int32_t> cnt(maxlev + 1, 0);
    for (int32_t v : nodes) ++cnt[level[v]];
    std::vector<std::vector<int32_t>> by_level(maxlev + 1);
    for (int32_t v : orderv) by_level[level[v]].push_back(v);
    // Candidate splits between BFS levels k and k + 1 (edges join equal or adjacent
    // levels only) whose lower side holds 30-70% of the nodes.  The separator of a split is
    // a minimum vertex cover of the edges between the two levels, which by Konig's theorem
    // follows from a maximum matching of that bipartite graph; it is never larger than
    // either level set.  The smallest such separator wins (the first on ties).
    std::vector<int32_t> best_sep;
    int32_t best_k = -1, acc = 0;
    for (int32_t k = 0; k < maxlev; ++k) {
      acc += cnt[k];
      if (10 * acc < 3 * N || 10 * acc > 7 * N) continue;
      // bipartite graph: level-k nodes (left) -> level-(k+1) neighbours (right)
      std::vector<int32_t>& Lk = by_level[k];
      std::vector<int32_t>& Rk = by_level[k + 1];
      for (size_t i = 0; i < Rk.size(); ++i) loc[Rk[i]] = int32_t(i);
      std::vector<int32_t> eptr(Lk.size() + 1, 0), eadj;
      for (size_t i = 0; i < Lk.size(); ++i) {
        for (int32_t u : adj[Lk[i]])
          if (mark[u] == tag && level[u] == k + 1) eadj.push_back(loc[u]);
        eptr[i + 1] = int32_t(eadj.size());
      }
      // maximum matching by augmenting paths (Kuhn)
      std::vector<int32_t> match_l(Lk.size(), -1), match_r(Rk.size(), -1), seen(Rk.size(), -1);
      std::function<bool(int32_t, int32_t)> augment = [&](int32_t a, int32_t round) {
        for (int32_t e = eptr[a]; e < eptr[a + 1]; ++e) {
          const int32_t b = eadj[e];
          if (seen[b] == round) continue;
          seen[b] = round;
          if (match_r[b] < 0 || augment(match_r[b], round)) {
            match_l[a] = b;
            match_r[b] = a;
            return true;
          }
        }
        return false;
      };
      for (int32_t a = 0; a < int32_t(Lk.size()); ++a) augment(a, a);
      // Konig: alternating search from the unmatched left nodes; the cover is the left
      // nodes it does not reach that have edges, plus the right nodes it reaches
      std::vector<char> vis_l(Lk.size(), 0), vis_r(Rk.size(), 0);
      std::vector<int32_t> stack;
      for (int32_t a = 0; a < int32_t(Lk.size()); ++a)
        if (match_l[a] < 0) {
          vis_l[a] = 1;
          stack.push_back(a);
        }
      while (!stack.empty()) {
        const int32_t a = stack.back();
        stack.pop_back();
        for (int32_t e = eptr[a]; e < eptr[a + 1]; ++e) {
          const int32_t b = eadj[e];
          if (vis_r[b]) continue;
          vis_r[b] = 1;
          const int32_t a2 = match_r[b];
          if (a2 >= 0 && !vis_l[a2]) {
            vis_l[a2] = 1;
            stack.push_back(a2);
          }
        }
      }
      std::vector<int32_t> sep;
      int32_t sep_l = 0, sep_r = 0;
      for (int32_t a = 0; a < int32_t(Lk.size()); ++a)
        if (!vis_l[a] && eptr[a + 1] > eptr[a]) {
          sep.push_back(Lk[a]);
          ++sep_l;
        }
      for (int32_t b = 0; b < int32_t(Rk.size()); ++b)
        if (vis_r[b]) {
          sep.push_back(Rk[b]);
          ++sep_r;
        }
      if (acc - sep_l == 0 || N - acc - sep_r == 0) continue;
      if (best_k < 0 || sep.size() < best_sep.size()) {
        best_k = k;
        best_sep.swap(sep);
      }
    }
    std::vector<int32_t> left, right, sep;
    if (best_k < 0) {  // no candidate: the median level set
      int32_t m = 1;
      acc = 0;
      for (int32_t l = 0; l <= maxlev; ++l) {
        acc += cnt[l];
        if (2 * acc >= N) {
          m = l;
          break;
        }
      }
      m = std::max(1, std::min(m, maxlev - 1));
      for (int32_t v : nodes)
        (level[v] < m ? left : level[v] > m ? right : sep).push_back(v);
    } else {
      for (int32_t v : best_sep) loc[v] = -2;
      for (int32_t v : nodes) {
        if (loc[v] == -2)
          sep.push_back(v);
        else
          (level[v] <= best_k ? left : right).push_back(v);
      }
      for (int32_t v : best_sep) loc[v] = -1;
    }
    rec(left);
    rec(right);
    std::sort(sep.begin(), sep.end());
    out.insert(out.end(), sep.begin(), sep.end());
  };
  std::vector<int32_t> all(n);
  std::iota(all.begin(), all.end(), 0);
  rec(all);
  return out;
}

// ---- symbolic analysis -------------------------------------------------------------------

void analyze_points(const HostProblem& P, Symbolic& S) {
  if (!P.terms_sorted) throw Error(SGPU_ERR_INVALID_ARGUMENT, "terms not sorted");
  const Family& poses = P.fams[P.f_pose];
  const Family& points = P.fams[P.f_point];
  S.nb = int32_t(P.num_free_blocks);
  S.npt = int32_t(points.n);
  S.npose = int32_t(poses.n);
  S.nobs = int32_t(P.obs_point.size());
  S.pose_block = poses.block;
  // point-major CSR and nu multiplicities (one per reduced (point, pose) term)
  S.pt_ptr.assign(S.npt + 1, 0);
  for (int32_t o = 0; o < S.nobs; ++o) ++S.pt_ptr[P.obs_point[o] + 1];
  for (int32_t b = 0; b < S.npt; ++b) S.pt_ptr[b + 1] += S.pt_ptr[b];
  S.pt_nu.assign(S.npt, 0);
  S.cam_nu.assign(S.nb, 0);
  // a reduced term is a run of terms equal in every non-conditioned argument
  const bool intr_key = P.intr_kind != SGPU_VAR_CONDITIONED;
  for (int32_t o = 0; o < S.nobs; ++o) {
    const bool first = o == 0 || P.obs_point[o] != P.obs_point[o - 1] ||
                       P.obs_pose[o] != P.obs_pose[o - 1] ||
                       (intr_key && P.obs_intr[o] != P.obs_intr[o - 1]);
    if (!first) continue;
    ++S.pt_nu[P.obs_point[o]];
    const int32_t c = poses.block[P.obs_pose[o]];
    if (c >= 0) ++S.cam_nu[c];
    if (P.intr_block >= 0 && P.obs_intr[o] == P.intr_member) ++S.cam_nu[P.intr_block];
  }
  // camera-major lists
  S.cam_ptr.assign(S.nb + 1, 0);
  for (int32_t o = 0; o < S.nobs; ++o) {
    const int32_t c = poses.block[P.obs_pose[o]];
    if (c >= 0) ++S.cam_ptr[c + 1];
  }
  for (int32_t c = 0; c < S.nb; ++c) S.cam_ptr[c + 1] += S.cam_ptr[c];
  S.cam_obs.resize(S.cam_ptr[S.nb]);
  {
    std::vector<int32_t> fill(S.cam_ptr.begin(), S.cam_ptr.end() - 1);
    for (int32_t o = 0; o < S.nobs; ++o) {
      const int32_t c = poses.block[P.obs_pose[o]];
      if (c >= 0) S.cam_obs[fill[c]++] = o;
    }
  }
}

std::vector<std::pair<int32_t, int32_t>> pattern_extras(const HostProblem& P, const Symbolic& S) {
  const Family& poses = P.fams[P.f_pose];
  std::vector<std::pair<int32_t, int32_t>> out;
  for (int32_t c = 0; c < S.nb; ++c) out.emplace_back(c, c);
  for (const auto& g : P.pg)
    for (size_t t = 0; t < g.m.size(); ++t) {
      const int32_t a = poses.block[g.m[t]], b = poses.block[g.n[t]];
      if (a >= 0 && b >= 0) out.emplace_back(std::max(a, b), std::min(a, b));
    }
  // free intrinsics: the last row, coupled to every pose observed through them
  if (P.intr_block >= 0)
    for (int32_t o = 0; o < S.nobs; ++o) {
      const int32_t c = poses.block[P.obs_pose[o]];
      if (c >= 0 && P.obs_intr[o] == P.intr_member) out.emplace_back(P.intr_block, c);
    }
  return out;
}

void analyze_chunks(Symbolic& S) {
  S.cam_chunk_ptr.assign(S.nb + 1, 0);
  S.own_chunks.clear();
  for (int32_t c = 0; c < S.nb; ++c) {
    for (int32_t q = S.cam_ptr[c]; q < S.cam_ptr[c + 1]; q += kOwnChunk)
      S.own_chunks.insert(S.own_chunks.end(), {c, q, std::min(q + kOwnChunk, S.cam_ptr[c + 1])});
    S.cam_chunk_ptr[c + 1] = int32_t(S.own_chunks.size() / 3);
  }
}

Symbolic analyze(const HostProblem& P) {
  Symbolic S;
  analyze_points(P, S);
  const Family& poses = P.fams[P.f_pose];
  const Family& points = P.fams[P.f_point];
  // S pattern: diagonal, co-observed pairs of active points, pose-graph pairs, the free
  // intrinsics' row
  std::vector<std::vector<int32_t>> rows(S.nb);
  for (const auto& rc : pattern_extras(P, S)) rows[rc.first].push_back(rc.second);
  std::vector<int32_t> cams;
  for (int32_t b = 0; b < S.npt; ++b) {
    if (points.block[b] < 0) continue;
    cams.clear();
    for (int32_t o = S.pt_ptr[b]; o < S.pt_ptr[b + 1]; ++o) {
      const int32_t c = poses.block[P.obs_pose[o]];
      if (c >= 0 && (cams.empty() || cams.back() != c)) cams.push_back(c);
    }
    for (size_t i = 0; i < cams.size(); ++i)
      for (size_t j = 0; j < i; ++j)
        rows[std::max(cams[i], cams[j])].push_back(std::min(cams[i], cams[j]));
  }
  S.s_ptr.assign(S.nb + 1, 0);
  for (int32_t c = 0; c < S.nb; ++c) {
    auto& r = rows[c];
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
    S.s_ptr[c + 1] = S.s_ptr[c] + int32_t(r.size());
  }
  S.s_col.resize(S.s_ptr[S.nb]);
  for (int32_t c = 0; c < S.nb; ++c) std::copy(rows[c].begin(), rows[c].end(), &S.s_col[S.s_ptr[c]]);
  // per-slot cross pairs i != j of active points (two passes: count, then fill in point
  // order), as positions in the camera-major observation list; the (o, o) terms of each
  // camera are summed per chunk of its camera-major list
  std::vector<int32_t> cam_pos(S.nobs, -1);
  for (int32_t q = 0; q < int32_t(S.cam_obs.size()); ++q) cam_pos[S.cam_obs[q]] = q;
  analyze_chunks(S);
  {
    const int64_t nslot = int64_t(S.s_col.size());
    auto slot_of = [&](int32_t r, int32_t c) -> int64_t {
      auto b = S.s_col.begin() + S.s_ptr[r], e = S.s_col.begin() + S.s_ptr[r + 1];
      return int64_t(std::lower_bound(b, e, c) - S.s_col.begin());
    };
    S.pair_ptr.assign(nslot + 1, 0);
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<int64_t> fill;
      if (pass == 1) {
        for (int64_t k = 0; k < nslot; ++k) S.pair_ptr[k + 1] += S.pair_ptr[k];
        S.pairs.assign(2 * S.pair_ptr[nslot], 0);
        fill.assign(S.pair_ptr.begin(), S.pair_ptr.end() - 1);
      }
      for (int32_t b = 0; b < S.npt; ++b) {
        const bool active = points.block[b] >= 0;
        for (int32_t i = S.pt_ptr[b]; i < S.pt_ptr[b + 1]; ++i) {
          const int32_t ci = poses.block[P.obs_pose[i]];
          if (ci < 0) continue;
          for (int32_t j = S.pt_ptr[b]; j < S.pt_ptr[b + 1]; ++j) {
            const int32_t cj = poses.block[P.obs_pose[j]];
            if (cj < 0 || cj > ci || !active || j == i) continue;
            const int64_t k = slot_of(ci, cj);
            if (pass == 0) {
              ++S.pair_ptr[k + 1];
            } else {
              const int64_t at = fill[k]++;
              S.pairs[2 * at] = cam_pos[i];
              S.pairs[2 * at + 1] = cam_pos[j];
            }
          }
        }
      }
    }
  }
  S.num_schur_pairs = int64_t(S.pairs.size() / 2) + S.cam_ptr[S.nb];
  analyze_finish(P, S);
  return S;
}

void analyze_finish(const HostProblem& P, Symbolic& S) {
  const Family& poses = P.fams[P.f_pose];
  if (P.intr_block >= 0) {
    const int32_t r = P.intr_block;
    S.intr_slot.assign(S.nb, -1);
    for (int32_t k = S.s_ptr[r]; k < S.s_ptr[r + 1]; ++k) {
      if (S.s_col[k] == r)
        S.intr_diag_slot = k;
      else
        S.intr_slot[S.s_col[k]] = k;
    }
    S.n_dense_tail = 1;
  }
  analyze_rcs(S);
  // pose-graph destinations (S slots)
  auto s_slot = [&](int32_t r, int32_t c) -> int32_t {
    auto b = S.s_col.begin() + S.s_ptr[r], e = S.s_col.begin() + S.s_ptr[r + 1];
    auto it = std::lower_bound(b, e, c);
    return int32_t(it - S.s_col.begin());
  };
  for (const auto& g : P.pg) {
    std::vector<Symbolic::PgDest> d(g.m.size());
    for (size_t t = 0; t < g.m.size(); ++t) {
      const int32_t a = poses.block[g.m[t]], b = poses.block[g.n[t]];
      d[t].mm = a >= 0 ? s_slot(a, a) : -1;
      d[t].nn = b >= 0 && !g.prior ? s_slot(b, b) : -1;
      d[t].mn = (a >= 0 && b >= 0 && !g.prior) ? s_slot(std::max(a, b), std::min(a, b)) : -1;
      d[t].mn_flip = a < b ? 1 : 0;
    }
    S.pg_dest.push_back(std::move(d));
  }
}

void analyze_rcs(Symbolic& S) {
  S.max_slots_per_row = 0;
  for (int32_t c = 0; c < S.nb; ++c)
    S.max_slots_per_row = std::max(S.max_slots_per_row, S.s_ptr[c + 1] - S.s_ptr[c]);
  // ordering: nested dissection of the blocks before the dense tail, which follows them
  const int32_t nd = S.nb - S.n_dense_tail;
  std::vector<std::vector<int32_t>> adj(nd);
  for (int32_t c = 0; c < nd; ++c)
    for (int32_t k = S.s_ptr[c]; k < S.s_ptr[c + 1]; ++k)
      if (S.s_col[k] != c) {
        adj[c].push_back(S.s_col[k]);
        adj[S.s_col[k]].push_back(c);
      }
  S.perm = nested_dissection(nd, adj, 24);
  for (int32_t c = nd; c < S.nb; ++c) S.perm.push_back(c);
  S.perm_inv.assign(S.nb, 0);
  for (int32_t i = 0; i < S.nb; ++i) S.perm_inv[S.perm[i]] = i;
  // permuted lower pattern: acol[j] = rows i > j of A(:, j); upper[i] = columns j < i
  std::vector<std::vector<int32_t>> acol(S.nb), upper(S.nb);
  for (int32_t c = 0; c < S.nb; ++c)
    for (int32_t k = S.s_ptr[c]; k < S.s_ptr[c + 1]; ++k) {
      const int32_t pi = S.perm_inv[c], pj = S.perm_inv[S.s_col[k]];
      const int32_t i = std::max(pi, pj), j = std::min(pi, pj);
      if (i != j) {
        acol[j].push_back(i);
        upper[i].push_back(j);
      }
    }
  // elimination tree (elimination_tree.rs:20-60)
  std::vector<int32_t>& parent = S.etree_parent;
  parent.assign(S.nb, -1);
  std::vector<int32_t> ancestor(S.nb, -1);
  for (int32_t j = 0; j < S.nb; ++j) {
    std::sort(upper[j].begin(), upper[j].end());
    for (int32_t i0 : upper[j]) {
      int32_t i = i0;
      while (i != j) {
        const int32_t next = ancestor[i];
        ancestor[i] = j;
        if (next < 0) {
          parent[i] = j;
          break;
        }
        i = next;
      }
    }
  }
  // column structure of L: struct(A(:,j)) U struct(L(:,child)) minus rows <= j
  std::vector<std::vector<int32_t>> lcol(S.nb), children(S.nb);
  for (int32_t j = 0; j < S.nb; ++j)
    if (parent[j] >= 0) children[parent[j]].push_back(j);
  S.nnz_l = 0;
  for (int32_t j = 0; j < S.nb; ++j) {
    std::vector<int32_t> st = acol[j];
    for (int32_t c : children[j])
      for (int32_t i : lcol[c])
        if (i > j) st.push_back(i);
    std::sort(st.begin(), st.end());
    st.erase(std::unique(st.begin(), st.end()), st.end());
    lcol[j].swap(st);
    S.nnz_l += int64_t(lcol[j].size());
  }
  // supernodes: runs j, j+1, ... with parent(j) == j+1.  Along such a chain
  // struct(L(:,k)) \ {k+1} is contained in struct(L(:,k+1)), so the run's columns plus
  // the structure of its last column hold every column's structure; entries outside a
  // column's own structure stay exact zeros through the dense front updates.
  S.sn_c0.clear();
  S.sn_nc.clear();
  S.blk_sn.assign(S.nb, -1);
  constexpr int32_t kMaxSupernode = 96;
  for (int32_t j = 0; j < S.nb;) {
    int32_t e = j + 1;
    while (e < S.nb && parent[e - 1] == e && e - j < kMaxSupernode) ++e;
    const int32_t sn = int32_t(S.sn_c0.size());
    S.sn_c0.push_back(j);
    S.sn_nc.push_back(e - j);
    for (int32_t k = j; k < e; ++k) S.blk_sn[k] = sn;
    j = e;
  }
  S.nsn = int32_t(S.sn_c0.size());
  S.sn_nf.assign(S.nsn, 0);
  S.sn_parent.assign(S.nsn, -1);
  S.sn_rows_ptr.assign(S.nsn + 1, 0);
  S.sn_rows.clear();
  for (int32_t x = 0; x < S.nsn; ++x) {
    const int32_t c0 = S.sn_c0[x], nc = S.sn_nc[x], last = c0 + nc - 1;
    for (int32_t k = c0; k <= last; ++k) S.sn_rows.push_back(k);
    for (int32_t i : lcol[last]) S.sn_rows.push_back(i);
    S.sn_nf[x] = nc + int32_t(lcol[last].size());
    S.sn_rows_ptr[x + 1] = int32_t(S.sn_rows.size());
    if (!lcol[last].empty()) S.sn_parent[x] = S.blk_sn[lcol[last][0]];
    S.max_front = std::max(S.max_front, S.sn_nf[x]);
  }
  // positions of each node's update rows in its parent's front
  S.sn_upos_ptr.assign(S.nsn + 1, 0);
  S.sn_upos.clear();
  for (int32_t x = 0; x < S.nsn; ++x) {
    const int32_t p = S.sn_parent[x];
    if (p >= 0) {
      const int32_t* prow = &S.sn_rows[S.sn_rows_ptr[p]];
      const int32_t pn = S.sn_nf[p];
      for (int32_t q = S.sn_rows_ptr[x] + S.sn_nc[x]; q < S.sn_rows_ptr[x + 1]; ++q) {
        const int32_t r = S.sn_rows[q];
        const int32_t pos = int32_t(std::lower_bound(prow, prow + pn, r) - prow);
        if (pos >= pn || prow[pos] != r) throw Error(SGPU_ERR_INVALID_ARGUMENT, "front structure");
        S.sn_upos.push_back(pos);
      }
    }
    S.sn_upos_ptr[x + 1] = int32_t(S.sn_upos.size());
  }
  // front storage (column-major F x F) and the factor's L columns (F x K per node)
  S.sn_off.assign(S.nsn, 0);
  S.sn_loff.assign(S.nsn, 0);
  S.front_elems = 0;
  S.l_elems = 0;
  S.factor_flops = 0;
  for (int32_t x = 0; x < S.nsn; ++x) {
    const int64_t F = 6 * int64_t(S.sn_nf[x]), K = 6 * int64_t(S.sn_nc[x]);
    S.sn_off[x] = S.front_elems;
    S.sn_loff[x] = S.l_elems;
    S.front_elems += F * F;
    S.l_elems += F * K;
    // K pivots of a front of F rows: ~ sum_k (F - k)^2 multiply-adds on the lower part
    for (int64_t k = 0; k < K; ++k) S.factor_flops += double(F - k) * double(F - k);
  }
  // levels of the supernodal tree (children before parents)
  S.sn_level.assign(S.nsn, 0);
  int32_t nlev = 0;
  for (int32_t x = 0; x < S.nsn; ++x) {  // parents have larger indices
    if (S.sn_parent[x] >= 0)
      S.sn_level[S.sn_parent[x]] = std::max(S.sn_level[S.sn_parent[x]], S.sn_level[x] + 1);
    nlev = std::max(nlev, S.sn_level[x] + 1);
  }
  S.lvl_ptr.assign(nlev + 1, 0);
  for (int32_t x = 0; x < S.nsn; ++x) ++S.lvl_ptr[S.sn_level[x] + 1];
  for (int32_t l = 0; l < nlev; ++l) S.lvl_ptr[l + 1] += S.lvl_ptr[l];
  S.lvl_sn.assign(S.nsn, 0);
  {
    std::vector<int32_t> fill(S.lvl_ptr.begin(), S.lvl_ptr.end() - 1);
    for (int32_t x = 0; x < S.nsn; ++x) S.lvl_sn[fill[S.sn_level[x]]++] = x;
  }
  // per-node scratch offsets (W of a panel, the panel inverses, the row-major L_KK copy),
  // then the round schedule
  S.launches.clear();
  S.launch_tiles.clear();
  S.panel_items.clear();
  S.sn_woff.assign(S.nsn, 0);
  S.sn_toff.assign(S.nsn, 0);
  S.sn_poff.assign(S.nsn, 0);
  S.w_elems = 0;
  S.t_elems = 0;
  S.p_elems = 0;
  for (int32_t x = 0; x < S.nsn; ++x) {
    S.sn_woff[x] = S.w_elems;
    S.w_elems += int64_t(6) * S.sn_nf[x] * 6 * kPanelBlocks;
    S.sn_poff[x] = S.p_elems;
    S.p_elems += int64_t(36) * kPanelBlocks * kPanelBlocks *
                 ((S.sn_nc[x] + kPanelBlocks - 1) / kPanelBlocks);
    S.sn_toff[x] = S.t_elems;
    S.t_elems += int64_t(36) * S.sn_nc[x] * S.sn_nc[x];
  }
  // round schedule (list scheduling, every step as early as it can run): a node's panel
  // steps run one per round, a leaf's from round 0, a parent's from the round after its
  // last child's last step (which also factors the parent's first diagonal block);
  // children of one parent end in distinct rounds, taken in order of the rounds left on
  // their path to the root (which is also the order of the work within a round).  Holding
  // back steps off the critical path to even out the rounds gained nothing here: the
  // rounds with the most tiles hold the large separators, which are all near-critical.
  {
    auto npanels = [&](int32_t x) { return (S.sn_nc[x] + kPanelBlocks - 1) / kPanelBlocks; };
    // panel steps: every panel with rows below it (all but a root's last)
    auto nsteps = [&](int32_t x) { return npanels(x) - (S.sn_nf[x] > S.sn_nc[x] ? 0 : 1); };
    std::vector<int32_t> pending(S.nsn, 0);
    for (int32_t x = 0; x < S.nsn; ++x)
      if (S.sn_parent[x] >= 0) {
        if (nsteps(x) == 0) throw std::runtime_error("supernode with a parent but no update rows");
        ++pending[S.sn_parent[x]];
      }
    // rounds from a node's first step to its root's last
    std::vector<int32_t> prio(S.nsn, 0);
    for (int32_t x = S.nsn - 1; x >= 0; --x)  // parents have larger indices
      prio[x] = nsteps(x) + (S.sn_parent[x] >= 0 ? prio[S.sn_parent[x]] : 0);
    std::vector<int32_t> next(S.nsn, 0), ready(S.nsn, std::numeric_limits<int32_t>::max());
    std::vector<int32_t> sibling_end(S.nsn, -1);  // per parent: the last round a child ended in
    S.sn_last_round.assign(S.nsn, -1);
    S.sn_hand.assign(S.nsn, 0);
    S.leaves.clear();
    int64_t remaining = 0;
    for (int32_t x = 0; x < S.nsn; ++x) {
      if (pending[x] == 0) {
        S.leaves.push_back(x);
        ready[x] = 0;
      }
      remaining += nsteps(x);
    }
    auto rem = [&](int32_t x) {  // rounds left on x's path, this step included
      return nsteps(x) - next[x] + (S.sn_parent[x] >= 0 ? prio[S.sn_parent[x]] : 0);
    };
    std::vector<std::vector<std::pair<int32_t, int32_t>>> steps;  // (node, panel) per round
    std::vector<int32_t> cand;
    for (int32_t r = 0; remaining > 0; ++r) {
      cand.clear();
      for (int32_t x = 0; x < S.nsn; ++x)
        if (ready[x] <= r && next[x] < nsteps(x)) cand.push_back(x);
      std::sort(cand.begin(), cand.end(),
                [&](int32_t a, int32_t b) { return rem(a) != rem(b) ? rem(a) > rem(b) : a < b; });
      auto& st = steps.emplace_back();
      for (const int32_t x : cand) {
        const int32_t p = next[x], P = S.sn_parent[x];
        const bool last = p + 1 == nsteps(x);
        if (last && P >= 0 && sibling_end[P] == r) continue;  // a sibling ends in this round
        st.push_back({x, p});
        ++next[x];
        ready[x] = r + 1;
        --remaining;
        if (last) {
          S.sn_last_round[x] = r;
          if (P >= 0) {
            sibling_end[P] = r;
            if (--pending[P] == 0) {
              ready[P] = r + 1;
              S.sn_hand[x] = 1;
            }
          }
        }
      }
    }
    const int32_t nround = int32_t(steps.size());
    for (int32_t r = 0; r < nround; ++r) {
      const auto& st = steps[r];  // in order of priority
      Symbolic::PanelLaunch L{int32_t(S.launch_tiles.size() / 4), 0,
                              int32_t(S.panel_items.size() / 3), 0, true, 0.0, 0.0};
      auto push = [&](int32_t x, int32_t p, int32_t tr, int32_t tc) {
        S.launch_tiles.insert(S.launch_tiles.end(), {x, p, tr, tc});
        ++L.tile_count;
      };
      // the critical work first: tile (0, 0) of every step that factors a diagonal block
      // (the node's next, or its parent's first), then the other tiles
      auto chain = [&](int32_t x, int32_t p) { return p + 1 < npanels(x) || S.sn_hand[x]; };
      for (const auto& [x, p] : st)
        if (chain(x, p)) push(x, p, 0, 0);
      for (const auto& [x, p] : st) {
        const int32_t kend = 6 * std::min(S.sn_nc[x], (p + 1) * kPanelBlocks);
        const int32_t T = 6 * S.sn_nf[x] - kend;
        const int32_t nt = (T + kTile - 1) / kTile;
        const double np = double(kend - 6 * p * kPanelBlocks);
        // lower trailing update T(T+1)/2 x np multiply-adds; bytes: the lower trailing
        // region read and written, the panel columns of the rows below read, their L written
        L.flops += double(T) * double(T + 1) * np;
        L.bytes += 16.0 * double(T) * double(T + 1) / 2.0 + 16.0 * double(T) * np;
        for (int32_t ch = 0; ch * kPanelRows < T; ++ch) S.panel_items.insert(S.panel_items.end(), {x, p, ch});
        for (int32_t tr = 0; tr < nt; ++tr)
          for (int32_t tc = 0; tc <= tr; ++tc)
            if (tr > 0 || !chain(x, p)) push(x, p, tr, tc);
      }
      L.item_count = int32_t(S.panel_items.size() / 3) - L.item_begin;
      L.fused = L.tile_count < kSplitTiles;
      if (L.fused) {  // the fused launch forms its panel rows itself
        S.panel_items.resize(size_t(3) * L.item_begin);
        L.item_count = 0;
      }
      S.launches.push_back(L);
    }
  }
  // solve work items
  S.solve_cols_ptr.assign(nlev + 1, 0);
  S.solve_cols.clear();
  S.lvl_max_nc.assign(nlev, 0);
  for (int32_t l = 0; l < nlev; ++l) {
    for (int32_t q = S.lvl_ptr[l]; q < S.lvl_ptr[l + 1]; ++q) {
      const int32_t x = S.lvl_sn[q];
      S.lvl_max_nc[l] = std::max(S.lvl_max_nc[l], S.sn_nc[x]);
      const int32_t m = 6 * (S.sn_nf[x] - S.sn_nc[x]), K = 6 * S.sn_nc[x];
      if (m > 0)
        for (int32_t t = 0; t * kSolveCols < K; ++t) S.solve_cols.insert(S.solve_cols.end(), {x, t});
    }
    S.solve_cols_ptr[l + 1] = int32_t(S.solve_cols.size() / 2);
  }
  {
    S.sn_uoff.assign(S.nsn, 0);
    S.u_elems = 0;
    for (int32_t x = 0; x < S.nsn; ++x) {
      S.sn_uoff[x] = S.u_elems;
      S.u_elems += 6 * int64_t(S.sn_nf[x]);
    }
  }
  // S slot -> front element offset of the block's (0, 0) entry
  std::vector<int32_t> pos_in_front(S.nb, -1);  // position of block j within its own node
  for (int32_t x = 0; x < S.nsn; ++x)
    for (int32_t k = 0; k < S.sn_nc[x]; ++k) pos_in_front[S.sn_c0[x] + k] = k;
  S.s_off.resize(S.s_col.size());
  S.s_ld.resize(S.s_col.size());
  for (int32_t c = 0; c < S.nb; ++c)
    for (int32_t k = S.s_ptr[c]; k < S.s_ptr[c + 1]; ++k) {
      const int32_t pi = S.perm_inv[c], pj = S.perm_inv[S.s_col[k]];
      const int32_t row = std::max(pi, pj), col = std::min(pi, pj);
      const int32_t x = S.blk_sn[col];
      const int32_t* rows = &S.sn_rows[S.sn_rows_ptr[x]];
      const int32_t rp = int32_t(std::lower_bound(rows, rows + S.sn_nf[x], row) - rows);
      const int64_t F = 6 * int64_t(S.sn_nf[x]);
      S.s_off[k] = S.sn_off[x] + 6 * int64_t(pos_in_front[col]) * F + 6 * int64_t(rp);
      S.s_ld[k] = int32_t(F) * 2 + (pi < pj ? 1 : 0);
    }
  // children of each supernode, in node order (the standalone forward solve gathers them)
  S.sn_child_ptr.assign(S.nsn + 1, 0);
  for (int32_t x = 0; x < S.nsn; ++x)
    if (S.sn_parent[x] >= 0) ++S.sn_child_ptr[S.sn_parent[x] + 1];
  for (int32_t x = 0; x < S.nsn; ++x) S.sn_child_ptr[x + 1] += S.sn_child_ptr[x];
  S.sn_child.assign(S.sn_child_ptr[S.nsn], 0);
  {
    std::vector<int32_t> fill(S.sn_child_ptr.begin(), S.sn_child_ptr.end() - 1);
    for (int32_t x = 0; x < S.nsn; ++x)
      if (S.sn_parent[x] >= 0) S.sn_child[fill[S.sn_parent[x]]++] = x;
  }
}

}  // namespace sgpu
