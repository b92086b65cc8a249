"""GPU parity of the equality-constrained Schur LM (range-space KKT) against the FP64 oracle.

The reference's BaScaleConstraintProblem (crates/sophus_opt/src/example_problems/
ba_scale_constraint.rs): the BaProblem with only pose 0 fixed and the scale fixed by
|t(pose 1)| = r.  The oracle is pinned on the CPU (tests/test_oracle_eq_constraints.py).
Bars: both sides in the GPU's elimination order; every step's accept decision equal, merit
1e-9 relative, damping 1e-6, the multipliers 1e-6 relative, final parameters 1e-7 (the
camera relative to its size).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(1e-300, float(np.max(np.abs(b)))))


def _eucm_case(sg, nc, npt, seed, sets):
    p = sg.generate_ba_eucm(nc, npt, seed=seed)
    v = (sg.VarBuilder()
         .add_family("cams", sg.VarFamily(sg.VarKind.Free, sg.VarType.Eucm, p.intrinsics_init))
         .add_family("poses", sg.VarFamily(sg.VarKind.Free, sg.VarType.Isometry3, p.poses_init, [0]))
         .add_family("points", sg.VarFamily(sg.VarKind.Marginalized, sg.VarType.Vec3, p.points_init))
         .build())
    costs = p.cost_fns()
    radius = lambda m: float(np.linalg.norm(p.poses_gt[m, 4:]))
    eqs = [sg.EqConstraintFn.translation_norm("poses", ms, [radius(m) for m in ms]) for ms in sets]
    ofams = [("cams", 3, 0, p.intrinsics_init, []), ("poses", 0, 0, p.poses_init, [0]),
             ("points", 1, 2, p.points_init, [])]
    ocosts = [(2, ["cams", "poses", "points"], p.obs_idx, p.obs_uv, None)]
    oeqs = [(0, "poses", ms, [radius(m) for m in ms]) for ms in sets]
    return p, v, costs, eqs, ofams, ocosts, oeqs


def _compare(sg, orc, v, costs, eqs, ofams, ocosts, oeqs, iterations, names, solver=None):
    solver = sg.Solver.SchurBlockSparseLdlt if solver is None else solver
    params = sg.OptParams(num_iterations=iterations, initial_lm_damping=1.0, solver=solver)
    opt = sg.Optimizer(v, costs, params, eqs)
    trace = []
    while True:
        info = opt.step()
        trace.append(info)
        if info.termination is not None:
            break
    out = {n: opt.members(n) for n in names}
    lam, res = opt.eq_state()
    merit = opt.merit()
    opt.close()
    orc.set_rcs_order(sg.analyze_rcs_order(v, costs))
    try:
        sol, ovars, otrace = orc.optimize_nlls(ofams, ocosts, orc.default_params(num_iterations=iterations,
                                                                                 initial_lm_damping=1.0,
                                                                                 solver=int(solver)),
                                               eqs=oeqs)
    finally:
        orc.set_rcs_order(None)
    assert len(trace) == len(otrace)
    worst = 0.0
    for g, o in zip(trace, otrace):
        assert g.accepted == o["accepted"]
        worst = max(worst, abs(g.merit - o["merit"]) / o["merit"])
        assert abs(g.lm_damping - o["lm_damping"]) <= 1e-6 * o["lm_damping"]
    assert worst <= 1e-9, worst
    assert abs(merit - sol["final_cost"]) <= 1e-9 * sol["final_cost"]
    dl = _rel(lam, sol["lambda"])
    dparams = {n: (_rel(out[n], ovars[n]) if n == "cams" else float(np.max(np.abs(out[n] - ovars[n]))))
               for n in names}
    print(f"merit {worst:.3g}, lambda {dl:.3g} ({lam}), params {dparams}, residuals {res}")
    assert dl < 1e-6
    assert all(d < 1e-7 for d in dparams.values()), dparams
    return out, res


@pytest.mark.parametrize("solver", ["SchurBlockSparseLdlt", "SchurSparseLdlt"])
def test_scale_constraint_problem_matches_oracle(sg, orc, solver):
    """BaScaleConstraintProblem::new(4, 20), Schur path, 20 iterations from nu = 1, with the
    reference's gates (ba_scale_constraint.rs:257-284); both Schur solvers (the reference's
    supports_eq_constraints admits both, lib.rs:386-388)."""
    p, v, costs, eqs, ofams, ocosts, oeqs = _eucm_case(sg, 4, 20, 0, [[1]])
    out, res = _compare(sg, orc, v, costs, eqs, ofams, ocosts, oeqs, 20, ("cams", "poses", "points"),
                        solver=getattr(sg.Solver, solver))
    r = float(np.linalg.norm(p.poses_gt[1, 4:]))
    assert abs(np.linalg.norm(out["poses"][1, 4:]) - r) < 1e-4 and abs(res[0]) < 1e-4
    pose_rms = np.sqrt(np.mean(np.sum((out["poses"][:, 4:] - p.poses_gt[:, 4:]) ** 2, 1)))
    pt_rms = np.sqrt(np.mean(np.sum((out["points"] - p.points_gt) ** 2, 1)))
    assert pose_rms < 0.1 and pt_rms < 0.1, (pose_rms, pt_rms)


def test_two_constraint_sets_one_with_two_terms_match_oracle(sg, orc):
    """40-camera ring: |t(pose 1)| = r1 as one set, and |t(pose 7)| + |t(pose 9)| = r7 + r9 as
    another (its two terms share one row and one multiplier, as in the reference's EqSystem)."""
    _, v, costs, eqs, ofams, ocosts, oeqs = _eucm_case(sg, 40, 2000, 1, [[1], [7, 9]])
    _, res = _compare(sg, orc, v, costs, eqs, ofams, ocosts, oeqs, 10, ("cams", "poses", "points"))
    assert np.max(np.abs(res)) < 1e-6


def test_pinhole_walk_with_scale_constraint_matches_oracle(sg, orc):
    """A 60-camera pinhole walk with camera 0 fixed and the scale fixed by |t(pose 5)|."""
    spec = sg.BaSpec(layout=1, num_cameras=60, num_points=3000, geometric_p=0.15, seed=7)
    p = sg.generate_ba(spec)
    r = float(np.linalg.norm(p.poses_gt[5, 4:]))
    v = p.variables()
    costs = p.cost_fns(1.0)
    eqs = [sg.EqConstraintFn.translation_norm("poses", [5], [r])]
    _, res = _compare(sg, orc, v, costs, eqs, orc.ba_families(p), orc.ba_costs(p), [(0, "poses", [5], [r])],
                      10, ("poses", "points"))
    assert abs(res[0]) < 1e-4  # the reference's constraint_satisfied bar (ba_scale_constraint.rs:215-219)
