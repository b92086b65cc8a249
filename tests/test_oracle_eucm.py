"""The oracle's EUCM camera and the reference's BaProblem on the CPU, and the library's host
side of the free-intrinsics layout (validation, ordering) — no GPU.

The oracle's EUCM (oracle_lie.hpp eucm_*, ba_oracle.cpp eval_eucm) restates
crates/sophus_sensor/src/distortions/enhanced_unified.rs and the cost term of
crates/sophus_opt/src/example_problems/ba_problem.rs:45-110.  It is pinned by:
- forward-mode dual numbers of the generic residual (every Jacobian column: a wrong sign,
  a dropped term or a transposed operand in the analytic derivatives fails);
- closed forms: alpha = 0 or beta = 0 is the pinhole (affine) camera of the pinhole cost
  term, whose oracle is pinned by the reference's fixtures; a hand-evaluated pixel;
- undistort(distort(x)) = x, the round trip the reference's camera tests use;
- the reference's own BA gate (ba_problem.rs:518-547): 20 LM iterations from the perturbed
  start give pose RMS < 0.01 and point RMS < 0.1 on the 4-camera / 20-point problem.
"""
import numpy as np
import pytest


def _terms(orc, n, seed, alpha=None, beta=None):
    rng = np.random.default_rng(seed)
    intr = np.tile([300.0, 310.0, 320.0, 240.0, 0.5, 1.0], (n, 1))
    intr[:, 4] = rng.uniform(0.1, 0.9, n) if alpha is None else alpha
    intr[:, 5] = rng.uniform(0.3, 1.7, n) if beta is None else beta
    poses = np.stack([orc.se3_exp(np.concatenate([rng.normal(0, 0.3, 3), rng.normal(0, 1.0, 3)]))
                      for _ in range(n)])
    pts = []
    for i in range(n):
        xc = np.array([rng.uniform(-2, 2), rng.uniform(-2, 2), rng.uniform(1, 6)])
        pts.append(orc.so3_matrix(poses[i][:4]) @ xc + poses[i][4:])
    uv = rng.uniform([0, 0], [640, 480], (n, 2))
    return intr, poses, np.array(pts), uv


def test_eucm_analytic_jacobians_match_dual_numbers(orc):
    intr, poses, pts, uv = _terms(orc, 300, 1)
    e = orc.eval_eucm_terms(intr, poses, pts, uv)
    jc, jp, jx = orc.eucm_dual_jacobians(intr, poses, pts, uv)
    for k, ref in (("jac_intr", jc), ("jac_pose", jp), ("jac_point", jx)):
        err = np.max(np.abs(e[k] - ref)) / np.max(np.abs(ref))
        assert err < 1e-12, (k, err)


def test_eucm_reduces_to_pinhole(orc):
    """alpha = 0 (norm = 1) and beta = 0 (rho = 1) are both the pinhole camera."""
    for alpha, beta in ((0.0, 0.8), (0.6, 0.0)):
        intr, poses, pts, uv = _terms(orc, 100, 2, alpha=alpha, beta=beta)
        e = orc.eval_eucm_terms(intr, poses, pts, uv)
        p = orc.eval_pinhole_terms(intr[:, :4], poses, pts, uv)
        for k in ("residual", "jac_pose", "jac_point", "cost"):
            err = np.max(np.abs(e[k] - p[k])) / max(1.0, np.max(np.abs(p[k])))
            assert err < 1e-12, (alpha, beta, k, err)
        # d/d(fx, fy, cx, cy) of the pinhole: -[[x, 0, 1, 0], [0, y, 0, 1]]
        xc = np.stack([orc.se3_mul(orc.se3_inverse(poses[i]), np.r_[1.0, 0, 0, 0, pts[i]])[4:]
                       for i in range(len(pts))])
        x, y = xc[:, 0] / xc[:, 2], xc[:, 1] / xc[:, 2]
        ref = -np.stack([np.stack([x, 0 * x, 1 + 0 * x, 0 * x], 1), np.stack([0 * y, y, 0 * y, 1 + 0 * y], 1)], 1)
        assert np.max(np.abs(e["jac_intr"][:, :, :4] - ref)) < 1e-12


def test_eucm_distort_closed_form_and_round_trip(orc):
    p = np.array([[300.0, 300.0, 320.0, 240.0, 0.5, 1.0]])
    # x = (0.3, -0.4): r^2 = 0.25, rho = sqrt(1.25), norm = 0.5 sqrt(1.25) + 0.5
    norm = 0.5 * np.sqrt(1.25) + 0.5
    uv = orc.eucm_distort(p, np.array([[0.3, -0.4]]))
    assert np.allclose(uv, [[300 * 0.3 / norm + 320, 300 * -0.4 / norm + 240]], rtol=0, atol=1e-12)
    rng = np.random.default_rng(3)
    n = 500
    intr = np.tile(p, (n, 1))
    intr[:, 4] = rng.uniform(0.0, 0.9, n)
    intr[:, 5] = rng.uniform(0.2, 2.0, n)
    xy = rng.uniform(-1.0, 1.0, (n, 2))
    back = orc.eucm_undistort(intr, orc.eucm_distort(intr, xy))
    assert np.max(np.abs(back - xy)) < 1e-12


def _oracle_fams(p, cams_kind=0, intr=None):
    k = p.intrinsics_init if intr is None else intr
    return [("cams", 3, cams_kind, k, []), ("poses", 0, 0, p.poses_init, [0, 1]),
            ("points", 1, 2, p.points_init, [])]


def _oracle_costs(p, delta=None):
    return [(2, ["cams", "poses", "points"], p.obs_idx, p.obs_uv, delta)]


def test_ba_problem_converges_on_the_oracle(sg, orc):
    """BaProblem::new(4, 20) with the reference's gate (ba_problem.rs:518-547), on the oracle."""
    p = sg.generate_ba_eucm(4, 20, seed=0)
    sol, mem, _ = orc.optimize_nlls(_oracle_fams(p), _oracle_costs(p),
                                    orc.default_params(num_iterations=20, initial_lm_damping=1.0))
    assert sol["final_cost"] < 1e-3 * sol["initial_cost"]
    pose_rms = np.sqrt(np.mean(np.sum((mem["poses"][:, 4:] - p.poses_gt[:, 4:]) ** 2, 1)))
    pt_rms = np.sqrt(np.mean(np.sum((mem["points"] - p.points_gt) ** 2, 1)))
    assert pose_rms < 0.01 and pt_rms < 0.1, (pose_rms, pt_rms)


def test_eucm_generator_invariants(sg):
    p = sg.generate_ba_eucm(30, 1500, seed=5)
    assert np.array_equal(p.intrinsics_gt[0], [300, 300, 320, 240, 0.5, 1.0])
    assert np.array_equal(p.intrinsics_init[0], [297, 297, 321, 241, 0.495, 1.0])
    assert np.array_equal(p.poses_init[:2], p.poses_gt[:2])
    assert not np.allclose(p.poses_init[2:], p.poses_gt[2:])
    counts = np.bincount(p.obs_idx[:, 2], minlength=p.points_gt.shape[0])
    assert counts.min() >= 2
    assert np.all(p.obs_idx[:, 0] == 0)
    assert np.all((p.obs_uv > -0.1) & (p.obs_uv < [640.1, 480.1]))
    q = sg.generate_ba_eucm(30, 1500, seed=5)
    assert np.array_equal(p.obs_uv, q.obs_uv) and np.array_equal(p.points_init, q.points_init)


def test_free_intrinsics_block_is_eliminated_last(sg):
    p = sg.generate_ba_eucm(40, 2000, seed=0)
    v, c = p.variables(), p.cost_fns()
    order = sg.analyze_rcs_order(v, c)
    st = sg.analyze(v, c)
    assert st["num_free_blocks"] == 38 + 1 == len(order)
    # the reference's layout numbers "cams" (name order before "poses") as free block 0
    assert order[-1] == 0 and sorted(order) == list(range(39))
    # conditioned intrinsics: no block
    assert sg.analyze(p.variables(sg.VarKind.Conditioned), c)["num_free_blocks"] == 38


def test_free_intrinsics_layouts_not_implemented_are_rejected(sg):
    p = sg.generate_ba_eucm(6, 100, seed=1)
    c = p.cost_fns()
    # two active members of the free intrinsics family
    two = (sg.VarBuilder()
           .add_family("cams", sg.VarFamily(sg.VarKind.Free, sg.VarType.Eucm, np.vstack([p.intrinsics_init] * 2)))
           .add_family("poses", sg.VarFamily(sg.VarKind.Free, sg.VarType.Isometry3, p.poses_init, [0, 1]))
           .add_family("points", sg.VarFamily(sg.VarKind.Marginalized, sg.VarType.Vec3, p.points_init))
           .build())
    with pytest.raises(sg.Unsupported):
        sg.analyze(two, c)
    # free pinhole intrinsics
    pin = (sg.VarBuilder()
           .add_family("cams", sg.VarFamily(sg.VarKind.Free, sg.VarType.Pinhole, p.intrinsics_init[:, :4]))
           .add_family("poses", sg.VarFamily(sg.VarKind.Free, sg.VarType.Isometry3, p.poses_init, [0, 1]))
           .add_family("points", sg.VarFamily(sg.VarKind.Marginalized, sg.VarType.Vec3, p.points_init))
           .build())
    with pytest.raises(sg.Unsupported):
        sg.analyze(pin, [sg.CostFn.pinhole_reprojection(["cams", "poses", "points"], p.obs_idx, p.obs_uv)])
    # the mixed precision with free intrinsics (rejected before any device call)
    with pytest.raises(sg.Unsupported):
        sg.Optimizer(p.variables(), c, sg.OptParams(precision=sg.Precision.Mixed))
    # EUCM terms on a pinhole family
    with pytest.raises(sg.FamilyTypeError):
        sg.analyze(pin, c)
