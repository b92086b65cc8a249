// Host-only check of the factorization's round schedule (Symbolic::launches) on generated
// bundle-adjustment problems: prints "ok <rounds> <nodes>" per problem, or the first broken
// invariant and exits 1.  Built and run by test_schedule.py (g++, no GPU).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "problem.hpp"

extern "C" sgpu_status sgpu_generate_ba(const sgpu_ba_spec*, sgpu_ba_sizes*, double*, double*, double*,
                                        double*, double*, int64_t*, double*);
namespace sgpu {
void clear_error() {}
sgpu_status set_error(sgpu_status s, const std::string&) { return s; }
}  // namespace sgpu

#define CHECK(cond, ...)                            \
  do {                                              \
    if (!(cond)) {                                  \
      std::printf("FAIL %s: ", #cond);              \
      std::printf(__VA_ARGS__);                     \
      std::printf("\n");                            \
      std::exit(1);                                 \
    }                                               \
  } while (0)

static void check(int layout, int ncam, long npt, uint64_t seed) {
  sgpu_ba_spec s{};
  s.layout = layout;
  s.num_cameras = ncam;
  s.num_points = npt;
  s.obs_per_point_ring = 0;
  s.min_obs = 2;
  s.max_obs = layout == 1 ? 30 : 6;
  s.geometric_p = 0.135;
  s.pixel_noise = 0.5;
  s.rot_noise = 0.05;
  s.trans_noise = 0.1;
  s.point_noise = 0.1;
  s.step_length = 0.3;
  s.yaw_jitter = 5 * M_PI / 180;
  s.seed = seed;
  sgpu_ba_sizes z{};
  CHECK(sgpu_generate_ba(&s, &z, 0, 0, 0, 0, 0, 0, 0) == SGPU_OK, "sizes");
  std::vector<double> K(4), pg(7 * z.num_cameras), pi(7 * z.num_cameras), xg(3 * z.num_points),
      xi(3 * z.num_points), uv(2 * z.num_obs);
  std::vector<int64_t> idx(3 * z.num_obs);
  CHECK(sgpu_generate_ba(&s, &z, K.data(), pg.data(), pi.data(), xg.data(), xi.data(), idx.data(),
                         uv.data()) == SGPU_OK,
        "generate");
  int64_t c0 = 0;
  sgpu_var_family f[3] = {{"cams", 2, 1, 1, K.data(), 0, nullptr},
                          {"poses", 0, 0, z.num_cameras, pi.data(), 1, &c0},
                          {"points", 1, 2, z.num_points, xi.data(), 0, nullptr}};
  sgpu_cost_fn c{0, 1, 1.0, {"cams", "poses", "points"}, z.num_obs, idx.data(), uv.data()};
  const auto P = sgpu::build_host_problem(f, 3, &c, 1);
  const sgpu::Symbolic S = sgpu::analyze(P);
  const int nsn = S.nsn, kP = sgpu::kPanelBlocks, kT = sgpu::kTile;
  auto npanels = [&](int x) { return (S.sn_nc[x] + kP - 1) / kP; };
  auto nsteps = [&](int x) { return npanels(x) - (S.sn_nf[x] > S.sn_nc[x] ? 0 : 1); };

  // every tile of every round, and the round each (node, panel) step runs in
  std::map<std::pair<int, int>, int> step_round;
  std::map<std::pair<int, int>, std::set<std::pair<int, int>>> step_tiles;
  for (size_t r = 0; r < S.launches.size(); ++r) {
    const auto& L = S.launches[r];
    bool in_chain_prefix = true;
    for (int i = 0; i < L.tile_count; ++i) {
      const int* t = &S.launch_tiles[4 * size_t(L.tile_begin + i)];
      const int x = t[0], p = t[1], tr = t[2], tc = t[3];
      CHECK(x >= 0 && x < nsn && p >= 0 && p < nsteps(x), "tile (%d, %d) outside the node's steps", x, p);
      auto key = std::make_pair(x, p);
      auto it = step_round.find(key);
      CHECK(it == step_round.end() || it->second == int(r), "step (%d, %d) in two rounds", x, p);
      step_round[key] = int(r);
      CHECK(step_tiles[key].insert(std::make_pair(tr, tc)).second, "tile (%d, %d) of step (%d, %d) twice", tr,
            tc, x, p);
      // tiles (0, 0) that factor a diagonal block (the node's next, or its parent's first)
      // come first in the round
      const bool chain = tr == 0 && tc == 0 && (p + 1 < npanels(x) || S.sn_hand[x]);
      if (!chain) in_chain_prefix = false;
      CHECK(!chain || in_chain_prefix, "chain tile of (%d, %d) after other tiles in round %zu", x, p, r);
    }
    // the panel-row items of a split round: every row chunk of every step in the round
    if (!L.fused) {
      std::set<std::tuple<int, int, int>> items;
      for (int i = 0; i < L.item_count; ++i) {
        const int* it = &S.panel_items[3 * size_t(L.item_begin + i)];
        CHECK(items.insert(std::make_tuple(it[0], it[1], it[2])).second, "panel item twice");
      }
      for (const auto& kv : step_round)
        if (kv.second == int(r)) {
          const int x = kv.first.first, p = kv.first.second;
          const int T = 6 * S.sn_nf[x] - 6 * std::min(S.sn_nc[x], (p + 1) * kP);
          for (int ch = 0; ch * sgpu::kPanelRows < T; ++ch)
            CHECK(items.count(std::make_tuple(x, p, ch)), "missing panel item (%d, %d, %d)", x, p, ch);
        }
    }
  }
  // each step's tiles: exactly the lower triangle of 64x64 tiles of its trailing region
  for (int x = 0; x < nsn; ++x)
    for (int p = 0; p < nsteps(x); ++p) {
      auto key = std::make_pair(x, p);
      CHECK(step_round.count(key), "step (%d, %d) never runs", x, p);
      const int T = 6 * S.sn_nf[x] - 6 * std::min(S.sn_nc[x], (p + 1) * kP);
      const int nt = (T + kT - 1) / kT;
      CHECK(int(step_tiles[key].size()) == nt * (nt + 1) / 2, "step (%d, %d) has %zu tiles, want %d", x, p,
            step_tiles[key].size(), nt * (nt + 1) / 2);
      for (const auto& t : step_tiles[key]) CHECK(t.first < nt && t.second <= t.first, "tile outside");
      if (p > 0) CHECK(step_round[key] > step_round[std::make_pair(x, p - 1)], "steps of %d out of order", x);
    }
  // parents after children; siblings end in distinct rounds; one hand-off child per parent,
  // the one that ends last; leaves are exactly the nodes without children
  std::vector<std::vector<int>> kids(nsn);
  for (int x = 0; x < nsn; ++x)
    if (S.sn_parent[x] >= 0) {
      CHECK(S.sn_parent[x] > x, "parent index below child");
      kids[S.sn_parent[x]].push_back(x);
    }
  // a child's update rows keep their order in the parent (strictly increasing positions), so
  // the parent's first kPW rows get the update only from the child's first update rows
  // (its tile (0, 0)), which the hand-off relies on
  for (int x = 0; x < nsn; ++x) {
    const int m = S.sn_nf[x] - S.sn_nc[x];
    CHECK(S.sn_upos_ptr[x + 1] - S.sn_upos_ptr[x] == m, "node %d: %d update positions for %d rows", x,
          S.sn_upos_ptr[x + 1] - S.sn_upos_ptr[x], m);
    for (int b = 1; b < m; ++b)
      CHECK(S.sn_upos[S.sn_upos_ptr[x] + b] > S.sn_upos[S.sn_upos_ptr[x] + b - 1], "node %d: update rows reorder", x);
    if (m > 0 && S.sn_parent[x] >= 0)
      CHECK(S.sn_upos[S.sn_upos_ptr[x] + m - 1] < S.sn_nf[S.sn_parent[x]], "node %d: position past the parent", x);
  }
  std::set<int> leaves(S.leaves.begin(), S.leaves.end());
  for (int x = 0; x < nsn; ++x) {
    CHECK(leaves.count(x) == (kids[x].empty() ? 1u : 0u), "leaf list wrong at %d", x);
    if (kids[x].empty()) {
      CHECK(S.sn_hand[x] == 0 || S.sn_parent[x] >= 0, "hand flag on a root leaf");
      continue;
    }
    int last = -1, last_kid = -1, hands = 0;
    std::set<int> ends;
    for (const int c : kids[x]) {
      CHECK(nsteps(c) > 0, "child %d without a step", c);
      const int e = step_round[std::make_pair(c, nsteps(c) - 1)];
      CHECK(ends.insert(e).second, "two children of %d end in round %d", x, e);
      if (e > last) last = e, last_kid = c;
      hands += S.sn_hand[c];
    }
    CHECK(hands == 1 && S.sn_hand[last_kid] == 1, "hand-off of %d not on its last child", x);
    if (nsteps(x) > 0)
      CHECK(step_round[std::make_pair(x, 0)] > last, "node %d steps before its children end", x);
  }
  std::printf("ok %zu %d\n", S.launches.size(), nsn);
}

int main() {
  check(1, 30, 3000, 1);      // small walk
  check(1, 120, 12000, 2);    // medium walk
  check(1, 400, 40000, 3);    // larger walk: many supernodes, split rounds
  check(1, 2000, 120000, 0);  // the benchmark's camera graph with fewer points
  check(0, 24, 2000, 4);      // ring of cameras
  check(1, 6, 600, 5);        // a few cameras: one supernode
  return 0;
}
