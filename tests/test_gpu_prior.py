"""Isometry3PriorCostTerm on the GPU against the oracle (isometry3_prior.rs:46-88), through the
C ABI: the term kernel (residual, dual Jacobian, precision- and Huber-weighted cost) and LM runs
with priors beside reprojection terms.  The oracle is pinned on the CPU
(tests/test_oracle_prior.py).  Bars: the term 1e-10 (Jacobian entries relative to 1); LM step by
step in the GPU's elimination order: the same accept sequence, merit 1e-9, damping 1e-9,
parameters 1e-7; FP32 assembly: final cost 1e-6.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _poses(orc, n, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    return np.stack([orc.se3_exp(rng.normal(0, scale, 6)) for _ in range(n)])


def test_prior_kernel_matches_oracle(sg, orc):
    rng = np.random.default_rng(7)
    n = 64
    E = _poses(orc, n, 8)
    # residual rotations from 2 rad down to 1e-8 (both sides of every small-angle branch) and 0
    scales = np.concatenate([np.geomspace(2.0, 1e-8, n - 4), [0, 0, 1e-3, 1e-6]])
    T = np.stack([orc.se3_mul(orc.se3_exp(rng.normal(0, 1, 6) * s), E[i]) for i, s in enumerate(scales)])
    A = rng.normal(size=(n, 6, 6))
    W = np.einsum("nij,nkj->nik", A, A) + np.eye(6)
    for delta in (None, 0.05):
        g = sg.eval_isometry3_prior_terms(T, E, W, huber_delta=delta)
        r, J, c = orc.eval_isometry3_prior_terms(T, E, W, huber_delta=delta)
        assert np.max(np.abs(g["residual"] - r)) < 1e-12
        assert np.max(np.abs(g["jacobian"] - J)) < 1e-10, np.max(np.abs(g["jacobian"] - J))
        # relative, plus what the residuals' own rounding (1e-16 absolute, large relative to a
        # residual of 1e-8) moves the cost: |dc| <= 2 sqrt(c |W|) |dr|
        wmax = np.linalg.norm(W, ord=2, axis=(1, 2))
        assert np.all(np.abs(g["cost"] - c) <= 1e-10 * c + 1e-14 * np.sqrt(c * wmax))


def _families(sg, p, const=()):
    return (sg.VarBuilder()
            .add_family("cams", sg.VarFamily(sg.VarKind.Conditioned, sg.VarType.Pinhole, p.intrinsics))
            .add_family("poses", sg.VarFamily(sg.VarKind.Free, sg.VarType.Isometry3, p.poses_init, list(const)))
            .add_family("points", sg.VarFamily(sg.VarKind.Marginalized, sg.VarType.Vec3, p.points_init))
            .build())


def _run_both(sg, orc, p, const, members, means, precs, iterations, prior_delta=None, precision=None):
    v = _families(sg, p, const)
    rk = sg.RobustKernel.huber(prior_delta) if prior_delta is not None else None
    costs = p.cost_fns(1.0) + [sg.CostFn.isometry3_prior(["poses"], members, means, precs, rk)]
    params = sg.OptParams(num_iterations=iterations, initial_lm_damping=1.0)
    if precision is not None:
        params.precision = precision
    opt = sg.Optimizer(v, costs, params)
    trace = []
    while True:
        info = opt.step()
        trace.append(info)
        if info.termination is not None:
            break
    out = {n: opt.members(n) for n in ("poses", "points")}
    merit = opt.merit()
    opt.close()
    prior = np.hstack([np.asarray(means).reshape(-1, 7), np.asarray(precs).reshape(-1, 36)])
    ofams = [("cams", 2, 1, p.intrinsics, []), ("points", 1, 2, p.points_init, []),
             ("poses", 0, 0, p.poses_init, list(const))]
    ocosts = orc.ba_costs(p) + [(3, ["poses"], np.asarray(members).reshape(-1, 1), prior, prior_delta)]
    orc.set_rcs_order(sg.analyze_rcs_order(v, costs))
    try:
        sol, ovars, otrace = orc.optimize_nlls(ofams, ocosts,
                                               orc.default_params(num_iterations=iterations, initial_lm_damping=1.0))
    finally:
        orc.set_rcs_order(None)
    return trace, out, merit, sol, ovars, otrace


def _compare(trace, out, sol, ovars, otrace):
    assert len(trace) == len(otrace)
    worst = 0.0
    for g, o in zip(trace, otrace):
        assert g.accepted == o["accepted"]
        worst = max(worst, abs(g.merit - o["merit"]) / o["merit"])
        assert abs(g.lm_damping - o["lm_damping"]) <= 1e-9 * o["lm_damping"]
    assert worst <= 1e-9, worst
    assert int(trace[-1].termination) == sol["termination"]
    dp = float(np.max(np.abs(out["poses"] - ovars["poses"])))
    dx = float(np.max(np.abs(out["points"] - ovars["points"])))
    print(f"merit {worst:.3g}, poses {dp:.3g}, points {dx:.3g}")
    assert dp < 1e-7 and dx < 1e-7


def test_cam_calib_style_priors_match_oracle(sg, orc):
    """cam_calib.rs:280-343 on the 10-camera ring: no pose constant, priors of precision 1e4 I
    on poses 0 and 1 at their true values fix the gauge; 10 LM iterations; the reference's gate
    on pose 2's translation."""
    p = sg.generate_ba(sg.config_spec(0))
    precs = np.repeat((1e4 * np.eye(6))[None], 2, axis=0)
    trace, out, merit, sol, ovars, otrace = _run_both(sg, orc, p, (), [0, 1], p.poses_gt[:2], precs, 10)
    _compare(trace, out, sol, ovars, otrace)
    assert np.max(np.abs(out["poses"][2, 4:] - p.poses_gt[2, 4:])) < 0.1


def test_walk_with_weighted_huber_priors_matches_oracle(sg, orc):
    """A 60-camera walk (pose 0 constant) with Huber priors of random SPD precisions on seven
    poses, one of them twice (two terms reduced into one: one damping), and one on the constant
    pose (cost only)."""
    spec = sg.BaSpec(layout=1, num_cameras=60, num_points=3000, geometric_p=0.15, seed=9)
    p = sg.generate_ba(spec)
    rng = np.random.default_rng(10)
    members = [0, 4, 11, 11, 23, 37, 50, 59]
    means = np.stack([orc.se3_mul(orc.se3_exp(rng.normal(0, 0.02, 6)), p.poses_gt[m]) for m in members])
    A = rng.normal(size=(len(members), 6, 6))
    precs = 50.0 * (np.einsum("nij,nkj->nik", A, A) + np.eye(6))
    trace, out, merit, sol, ovars, otrace = _run_both(sg, orc, p, (0,), members, means, precs, 10,
                                                      prior_delta=0.5)
    _compare(trace, out, sol, ovars, otrace)


def test_priors_with_fp32_assembly_final_cost(sg, orc):
    p = sg.generate_ba(sg.config_spec(0))
    precs = np.repeat((1e4 * np.eye(6))[None], 2, axis=0)
    trace, out, merit, sol, ovars, otrace = _run_both(sg, orc, p, (), [0, 1], p.poses_gt[:2], precs, 20,
                                                      precision=sg.Precision.Mixed)
    assert abs(merit - sol["final_cost"]) <= 1e-6 * sol["final_cost"]


def test_priors_with_pcg_final_cost(sg, orc):
    """Priors with the iterative Schur solver: the prior's 6x6 block lands on the pose's
    diagonal slot of S, which the Jacobi preconditioner inverts with it.  At a PCG tolerance
    of 1e-12 the run accepts the steps the direct oracle accepts and ends at its cost (1e-8)."""
    p = sg.generate_ba(sg.config_spec(0))
    precs = np.repeat((1e4 * np.eye(6))[None], 2, axis=0)
    v = _families(sg, p)
    costs = p.cost_fns(1.0) + [sg.CostFn.isometry3_prior(["poses"], [0, 1], p.poses_gt[:2], precs, None)]
    params = sg.OptParams(num_iterations=10, initial_lm_damping=1.0, solver=sg.Solver.SchurPcgJacobi,
                          pcg_relative_tolerance=1e-12, pcg_max_iterations=1000)
    opt = sg.Optimizer(v, costs, params)
    trace = []
    while True:
        info = opt.step()
        trace.append(info)
        if info.termination is not None:
            break
    merit = opt.merit()
    opt.close()
    assert all(t.pcg_iterations > 0 for t in trace)
    prior = np.hstack([p.poses_gt[:2], precs.reshape(-1, 36)])
    ofams = [("cams", 2, 1, p.intrinsics, []), ("points", 1, 2, p.points_init, []), ("poses", 0, 0, p.poses_init, [])]
    ocosts = orc.ba_costs(p) + [(3, ["poses"], np.array([[0], [1]]), prior, None)]
    sol, _, otrace = orc.optimize_nlls(ofams, ocosts, orc.default_params(num_iterations=10, initial_lm_damping=1.0))
    assert [t.accepted for t in trace] == [o["accepted"] for o in otrace]
    assert abs(merit - sol["final_cost"]) <= 1e-8 * sol["final_cost"]
