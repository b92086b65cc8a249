"""Equality-constrained Schur LM on the CPU: the oracle's range-space KKT solve and constrained
step, and the library's host-side validation of constraint sets (no GPU).

The oracle restates the reference's EqSystem (crates/sophus_opt/src/nlls/linear_system/
eq_system.rs), the range-space solve of SchurFactor (crates/sophus_solver/src/ldlt/
schur_ldlt.rs:450-480,674-730,762-772) and the constrained step of step_nlls_inner
(crates/sophus_opt/src/nlls/optimizer_step.rs:366-443), on the TranslationNormConstraint of
crates/sophus_opt/src/example_problems/ba_scale_constraint.rs.  It is pinned by:
- the constraint Jacobian from duals against its closed form [0 | t^T / |t|] (the rotation
  part vanishes because t^T hat(t) = 0);
- the Schur-eliminated range-space step against numpy's solve of the whole dense KKT system
  [[H, G^T], [G, 0]] (no Schur complement, no range-space algebra): a sign slip in G^T
  lambda, -c or dlambda fails it;
- the reference's own gates (ba_scale_constraint.rs:257-284): after 20 iterations of the
  Schur path the scale constraint holds to 1e-4 and pose and point RMS are below 0.1.
"""
import numpy as np
import pytest


def _problem(sg, nc=4, npt=20, seed=0):
    p = sg.generate_ba_eucm(nc, npt, seed=seed)
    r = float(np.linalg.norm(p.poses_gt[1, 4:]))
    fams = [("cams", 3, 0, p.intrinsics_init, []), ("poses", 0, 0, p.poses_init, [0]),
            ("points", 1, 2, p.points_init, [])]
    costs = [(2, ["cams", "poses", "points"], p.obs_idx, p.obs_uv, None)]
    return p, r, fams, costs


def test_translation_norm_jacobian_closed_form(sg, orc):
    p, r, fams, costs = _problem(sg, 6, 40, seed=3)
    nfree = 6 + 6 * 5
    dim = nfree + 3 * p.points_init.shape[0]
    eqs = [(0, "poses", [1], [r]), (0, "poses", [3, 4], [1.0, 2.0])]
    G, rc, _, _ = orc.kkt_step(fams, costs, eqs, 1.0, [0.0, 0.0], nfree, dim)
    t = p.poses_init[:, 4:]
    # set 0: pose 1 (free block 0 of the poses: scalars 6..11 after the camera's 6)
    ref0 = np.zeros(nfree)
    ref0[6 + 3:6 + 6] = t[1] / np.linalg.norm(t[1])
    ref1 = np.zeros(nfree)
    ref1[6 + 12 + 3:6 + 18] = t[3] / np.linalg.norm(t[3])  # pose m at 6 + 6 (m - 1)
    ref1[6 + 18 + 3:6 + 24] = t[4] / np.linalg.norm(t[4])
    assert np.max(np.abs(G[0] - ref0)) < 1e-15 and np.max(np.abs(G[1] - ref1)) < 1e-15
    # one constraint row per set: the terms' residuals add
    assert abs(rc[0] + (np.linalg.norm(t[1]) - r)) < 1e-14
    assert abs(rc[1] + (np.linalg.norm(t[3]) - 1.0 + np.linalg.norm(t[4]) - 2.0)) < 1e-14


@pytest.mark.parametrize("lam", [[0.0], [0.37]])
def test_range_space_step_solves_the_dense_kkt_system(sg, orc, lam):
    p, r, fams, costs = _problem(sg, 6, 40, seed=5)
    nfree = 6 + 6 * 5
    H, g0 = orc.linearize_dense(fams, costs, 0.5)
    dim = H.shape[0]
    eqs = [(0, "poses", [1], [r])]
    G, rc, g, dx = orc.kkt_step(fams, costs, eqs, 0.5, lam, nfree, dim)
    assert np.allclose(g[:nfree], g0[:nfree] + G.T @ np.asarray(lam), rtol=0, atol=1e-9 * np.abs(g0).max())
    Gf = np.zeros((1, dim))
    Gf[:, :nfree] = G
    K = np.block([[H, Gf.T], [Gf, np.zeros((1, 1))]])
    ref = np.linalg.solve(K, np.concatenate([g, rc]))
    err = np.max(np.abs(dx - ref)) / np.max(np.abs(ref))
    assert err < 1e-8, err


def test_scale_constraint_problem_converges_on_the_oracle(sg, orc):
    """BaScaleConstraintProblem::new(4, 20) on the Schur path (ba_scale_constraint.rs:257-284)."""
    p, r, fams, costs = _problem(sg)
    sol, mem, tr = orc.optimize_nlls(fams, costs, orc.default_params(num_iterations=20, initial_lm_damping=1.0),
                                     eqs=[(0, "poses", [1], [r])])
    assert abs(np.linalg.norm(mem["poses"][1, 4:]) - r) < 1e-4
    pose_rms = np.sqrt(np.mean(np.sum((mem["poses"][:, 4:] - p.poses_gt[:, 4:]) ** 2, 1)))
    pt_rms = np.sqrt(np.mean(np.sum((mem["points"] - p.points_gt) ** 2, 1)))
    assert pose_rms < 0.1 and pt_rms < 0.1, (pose_rms, pt_rms)
    # the multiplier moved off zero and the merit is the cost alone
    assert sol["lambda"][0] != 0.0
    assert sol["final_cost"] < 1e-3 * sol["initial_cost"]


def test_eq_constraint_sets_are_validated(sg):
    p = sg.generate_ba_eucm(6, 100, seed=1)
    r = float(np.linalg.norm(p.poses_gt[1, 4:]))
    v = (sg.VarBuilder()
         .add_family("cams", sg.VarFamily(sg.VarKind.Free, sg.VarType.Eucm, p.intrinsics_init))
         .add_family("poses", sg.VarFamily(sg.VarKind.Free, sg.VarType.Isometry3, p.poses_init, [0]))
         .add_family("points", sg.VarFamily(sg.VarKind.Marginalized, sg.VarType.Vec3, p.points_init))
         .build())
    c = p.cost_fns()
    norm = [sg.EqConstraintFn.translation_norm("poses", [1], [r])]
    # the reference's Schur check: a constraint on a marginalized family
    with pytest.raises(sg.SchurConstraintOnMarginalizedFamily):
        sg.Optimizer(v, c, sg.OptParams(), [sg.EqConstraintFn.translation_norm("points", [1], [r])])
    with pytest.raises(sg.UnknownFamily):
        sg.Optimizer(v, c, sg.OptParams(), [sg.EqConstraintFn.translation_norm("nope", [1], [r])])
    with pytest.raises(sg.FamilyTypeError):
        sg.Optimizer(v, c, sg.OptParams(), [sg.EqConstraintFn.translation_norm("cams", [0], [r])])
    with pytest.raises(sg.IndexOutOfRange):
        sg.Optimizer(v, c, sg.OptParams(), [sg.EqConstraintFn.translation_norm("poses", [99], [r])])
    # the range-space solve needs the direct factorization in FP64 (rejected before any device call)
    with pytest.raises(sg.Unsupported):
        sg.Optimizer(v, c, sg.OptParams(solver=sg.Solver.SchurPcgJacobi), norm)
    with pytest.raises(sg.Unsupported):
        sg.Optimizer(v, c, sg.OptParams(precision=sg.Precision.Mixed), norm)
