"""Pins the oracle's normal-equation assembly, Schur step, merit and LM control.

- assembly (cost_system.rs:74-376, evaluated_term.rs:286-327): brute-force dense
  J^T J with unweighted Jacobians, r~ = w r in the gradient, nu * I added once per
  reduced term to each non-conditioned argument's diagonal block, free families first
  (by name), then marginalized ones (var_families.rs:60-113, var_family.rs:156-237);
- the Schur step equals numpy's dense solve of that system (schur_ldlt.rs:82-812);
- the merit is the sum of rho(|r|) (optimizer_helpers.rs:309-321);
- LM control (optimizer_step.rs:366-443, optimizer.rs:463-480): nu update rules and
  termination reasons read off the oracle's per-step trace;
- the reference's BA convergence gate (ba_problem.rs:518-547);
- validation order and error variants (optimizer.rs:649-717) on the oracle and on the
  library's host-side analysis.
"""
import numpy as np
import pytest


def _ring(sg):
    return sg.generate_ba(sg.config_spec(0))


def _families(p, pose_name="poses", point_name="points", poses=None, points=None):
    return [("cams", 2, 1, p.intrinsics, []),
            (pose_name, 0, 0, p.poses_init if poses is None else poses, [0]),
            (point_name, 1, 2, p.points_init if points is None else points, [])]


def _costs(p, idx=None, uv=None, delta=1.0, pose_name="poses", point_name="points"):
    return [(0, ["cams", pose_name, point_name], p.obs_idx if idx is None else idx,
             p.obs_uv if uv is None else uv, delta)]


def _brute_force(orc, p, idx, uv, nu, delta, points_first=False):
    """Dense damped normal equations from per-term Jacobians (pose 0 is constant)."""
    nc, npt = p.poses_init.shape[0], p.points_init.shape[0]
    nfree = 6 * (nc - 1)
    n = nfree + 3 * npt
    o = orc.eval_pinhole_terms(np.repeat(p.intrinsics, len(idx), 0), p.poses_init[idx[:, 1]],
                               p.points_init[idx[:, 2]], uv, delta)
    H = np.zeros((n, n))
    g = np.zeros(n)
    for t in range(len(idx)):
        c, b = idx[t, 1], idx[t, 2]
        J = np.zeros((2, n))
        if c > 0:
            J[:, 6 * (c - 1):6 * c] = o["jac_pose"][t]
        J[:, nfree + 3 * b:nfree + 3 * b + 3] = o["jac_point"][t]
        H += J.T @ J
        g -= J.T @ (o["weight"][t] * o["residual"][t])
    # nu * I once per reduced term (terms sharing pose and point are summed first)
    for c, b in {(int(c), int(b)) for _, c, b in idx}:
        if c > 0:
            H[6 * (c - 1):6 * c, 6 * (c - 1):6 * c] += nu * np.eye(6)
        H[nfree + 3 * b:nfree + 3 * b + 3, nfree + 3 * b:nfree + 3 * b + 3] += nu * np.eye(3)
    return H, g, float(np.sum(o["cost"]))


@pytest.mark.parametrize("delta", [None, 1.0])
def test_assembly_matches_brute_force(sg, orc, delta):
    p = _ring(sg)
    nu = 0.7
    H, g = orc.linearize_dense(_families(p), _costs(p, delta=delta), nu)
    Hb, gb, _ = _brute_force(orc, p, p.obs_idx, p.obs_uv, nu, delta)
    assert H.shape == Hb.shape == (54 + 600, 54 + 600)
    assert np.max(np.abs(H - Hb)) <= 1e-9 * np.max(np.abs(Hb))
    assert np.max(np.abs(g - gb)) <= 1e-9 * np.max(np.abs(gb))


def test_assembly_reduces_duplicate_terms_before_damping(sg, orc):
    """Repeated (pose, point) observations are summed into one term: nu is added once."""
    p = _ring(sg)
    rng = np.random.default_rng(0)
    dup = rng.choice(len(p.obs_idx), 40, replace=False)
    idx = np.concatenate([p.obs_idx, p.obs_idx[dup]])
    uv = np.concatenate([p.obs_uv, p.obs_uv[dup] + rng.normal(0, 2.0, (40, 2))])
    nu = 2.5
    H, g = orc.linearize_dense(_families(p), _costs(p, idx, uv), nu)
    Hb, gb, _ = _brute_force(orc, p, idx, uv, nu, 1.0)
    assert np.max(np.abs(H - Hb)) <= 1e-9 * np.max(np.abs(Hb))
    assert np.max(np.abs(g - gb)) <= 1e-9 * np.max(np.abs(gb))


def test_layout_orders_free_families_by_name(sg, orc):
    """Family names decide the order within a kind; free families always precede
    marginalized ones, whatever the names."""
    p = _ring(sg)
    H1, g1 = orc.linearize_dense(_families(p), _costs(p), 1.0)
    H2, g2 = orc.linearize_dense(_families(p, "zz_poses", "aa_points"),
                                 _costs(p, pose_name="zz_poses", point_name="aa_points"), 1.0)
    assert np.array_equal(H1, H2) and np.array_equal(g1, g2)


def test_schur_step_equals_dense_solve(sg, orc):
    p = _ring(sg)
    nu = 1.0
    H, g = orc.linearize_dense(_families(p), _costs(p), nu)
    dx = orc.schur_step(_families(p), _costs(p), nu, H.shape[0])
    ref = np.linalg.solve(H, g)
    assert np.max(np.abs(dx - ref)) <= 1e-9 * np.max(np.abs(ref))


def test_merit_is_sum_of_robust_costs(sg, orc):
    p = _ring(sg)
    _, _, cost = _brute_force(orc, p, p.obs_idx, p.obs_uv, 0.0, 1.0)
    m = orc.merit(_families(p), _costs(p))
    assert abs(m - cost) <= 1e-12 * cost


def test_lm_trace_follows_damping_rules(sg, orc):
    """accept: rho > 0 and F_new < F, nu *= max(1/3, 1 - (2 rho - 1)^3);
    reject: nu *= 2, F unchanged (optimizer.rs:463-480)."""
    p = _ring(sg)
    sol, _, trace = orc.optimize_nlls(_families(p), _costs(p),
                                      orc.default_params(num_iterations=20, initial_lm_damping=1.0))
    # each record holds the merit and nu after its step (StepInfo, optimizer_types.rs:67-83)
    nu, merit = 1.0, sol["initial_cost"]
    n_acc = 0
    for rec in trace:
        if rec["accepted"]:
            n_acc += 1
            assert rec["gain_ratio"] > 0 and rec["new_merit"] < merit
            assert rec["merit"] == rec["new_merit"]
            f = max(1.0 / 3.0, 1.0 - (2.0 * rec["gain_ratio"] - 1.0) ** 3)
            assert abs(rec["lm_damping"] - nu * f) <= 1e-15 * rec["lm_damping"]
        else:
            assert rec["merit"] == merit
            assert rec["lm_damping"] == 2.0 * nu
        nu, merit = rec["lm_damping"], rec["merit"]
    assert n_acc >= 5
    assert sol["final_cost"] < 0.01 * sol["initial_cost"]
    assert sol["num_steps"] == len(trace)


def test_lm_stops_at_max_iterations(sg, orc):
    p = _ring(sg)
    sol, _, trace = orc.optimize_nlls(_families(p), _costs(p),
                                      orc.default_params(num_iterations=3, initial_lm_damping=1.0))
    assert sol["num_steps"] == 3 and len(trace) == 3
    assert sol["termination"] == 1  # MaxIterations


def test_lm_gradient_tolerance_stops_at_optimum(sg, orc):
    """At the ground truth without noise the gradient is zero: GradientTolerance at once."""
    spec = sg.config_spec(0)
    spec.pixel_noise = 0.0
    p = sg.generate_ba(spec)
    fams = _families(p, poses=p.poses_gt, points=p.points_gt)
    sol, _, trace = orc.optimize_nlls(fams, _costs(p), orc.default_params(num_iterations=10))
    assert sol["termination"] == 3 and sol["num_steps"] <= 1


def test_ba_converges_gate(sg, orc):
    """ba_problem.rs:518-547: pose RMS < 0.01 and point RMS < 0.1 after the solve."""
    p = _ring(sg)
    sol, vars_, _ = orc.optimize_nlls(_families(p), _costs(p),
                                      orc.default_params(num_iterations=20, initial_lm_damping=1.0))
    pose_rms = np.sqrt(np.mean((vars_["poses"][:, 4:] - p.poses_gt[:, 4:]) ** 2))
    pt_rms = np.sqrt(np.mean((vars_["points"] - p.points_gt) ** 2))
    assert pose_rms < 0.01 and pt_rms < 0.1


# ---- validation: the same status from the oracle and from the library's host analysis ----

def _lib_status(sg, fams, costs):
    vb = sg.VarBuilder()
    for name, typ, kind, members, const_ids in fams:
        vb.add_family(name, sg.VarFamily(sg.VarKind(kind), sg.VarType(typ), members, const_ids))
    cfs = [sg.CostFn(sg.CostKind(k), names, idx, meas,
                     sg.RobustKernel.huber(d) if d is not None else None)
           for k, names, idx, meas, d in costs]
    try:
        sg.analyze(vb.build(), cfs)
        return 0
    except sg.NllsError as e:
        return e.status


def _oracle_status(orc, fams, costs):
    r = orc.optimize_nlls(fams, costs, orc.default_params(num_iterations=1))
    return r[0] if isinstance(r[0], int) else 0


def _cases(p):
    fams = _families(p)
    costs = _costs(p)
    bad_idx = p.obs_idx.copy()
    bad_idx[5, 2] = p.points_init.shape[0]
    return {
        "ok": (fams, costs, 0),
        "no_marg": ([fams[0], fams[1], ("points", 1, 1, p.points_init, [])], costs, 2),
        "no_free": ([fams[0], ("poses", 0, 1, p.poses_init, []), fams[2]], costs, 2),
        "two_marg": ([fams[0], ("anchor", 0, 0, p.poses_init[:2], []), ("poses", 0, 2, p.poses_init, []),
                      fams[2]],
                     [(1, ["poses", "poses"], np.array([[0, 1]]), p.poses_init[:1], None)] + costs, 3),
        "unknown_family": (fams, _costs(p, point_name="pointz"), 4),
        "family_type": (fams, [(0, ["poses", "poses", "points"], p.obs_idx, p.obs_uv, 1.0)], 5),
        "index_out_of_range": (fams, _costs(p, idx=bad_idx), 6),
    }


@pytest.mark.parametrize("case", ["ok", "no_marg", "no_free", "two_marg", "unknown_family",
                                  "family_type", "index_out_of_range"])
def test_validation_errors_match_reference_variants(sg, orc, case):
    p = _ring(sg)
    fams, costs, want = _cases(p)[case]
    assert _oracle_status(orc, fams, costs) == want
    assert _lib_status(sg, fams, costs) == want


def test_threaded_oracle_is_bitwise_sequential(sg, orc):
    """The oracle's threaded loops (row-owned accumulation, term-ordered sums) give bitwise
    the sequential result: a 60-camera walk (3,000 points, several chunks of reduction
    ranges and points), five LM iterations at 1 and 4 threads."""
    p = sg.generate_ba(sg.BaSpec(layout=1, num_cameras=60, num_points=3000, geometric_p=0.15, seed=7))
    runs = []
    for threads in (1, 4):
        prev = orc.set_threads(threads)
        try:
            runs.append(orc.optimize_nlls(orc.ba_families(p), orc.ba_costs(p),
                                          orc.default_params(num_iterations=5, initial_lm_damping=1.0)))
        finally:
            orc.set_threads(prev)
    (s1, v1, t1), (s4, v4, t4) = runs
    assert [(r["merit"], r["lm_damping"], r["accepted"]) for r in t1] == \
        [(r["merit"], r["lm_damping"], r["accepted"]) for r in t4]
    for name in ("poses", "points"):
        assert np.array_equal(v1[name], v4[name])
    assert s1["final_cost"] == s4["final_cost"]
