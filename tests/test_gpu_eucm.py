"""GPU parity of the EUCM camera and free shared intrinsics (the reference's BaProblem,
crates/sophus_opt/src/example_problems/ba_problem.rs) against the FP64 oracle.

The oracle is pinned on the CPU (tests/test_oracle_eucm.py).  Bars follow the pinhole
tests: kernels 1e-10 relative; LM runs in a shared elimination order: every step's accept
decision equal, merit 1e-9 relative, LM damping 1e-6, final parameters 1e-7 (intrinsics
relative to their size); inverse blocks 1e-8 relative to the block's largest entry.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(1e-300, float(np.max(np.abs(b)))))


def _oracle_fams(p, cams_kind=0, poses=None, points=None, intr=None):
    if intr is None:
        intr = p.intrinsics_gt if cams_kind == 1 else p.intrinsics_init
    return [("cams", 3, cams_kind, intr, []),
            ("poses", 0, 0, p.poses_init if poses is None else poses, [0, 1]),
            ("points", 1, 2, p.points_init if points is None else points, [])]


def _oracle_costs(p, delta=None):
    return [(2, ["cams", "poses", "points"], p.obs_idx, p.obs_uv, delta)]


def _run_gpu(sg, p, params, cams_kind, delta=None):
    opt = sg.Optimizer(p.variables(cams_kind), p.cost_fns(delta), params)
    trace = []
    while True:
        info = opt.step()
        trace.append(info)
        if info.termination is not None:
            break
    out = {n: opt.members(n) for n in ("cams", "poses", "points")}
    merit = opt.merit()
    opt.close()
    return merit, out, trace


def _compare_runs(sg, orc, p, cams_kind, iterations, delta=None, param_bar=1e-7):
    params = sg.OptParams(num_iterations=iterations, initial_lm_damping=1.0)
    merit, out, trace = _run_gpu(sg, p, params, cams_kind, delta)
    orc.set_rcs_order(sg.analyze_rcs_order(p.variables(cams_kind), p.cost_fns(delta)))
    try:
        sol, ovars, otrace = orc.optimize_nlls(
            _oracle_fams(p, 1 if cams_kind == sg.VarKind.Conditioned else 0), _oracle_costs(p, delta),
            orc.default_params(num_iterations=iterations, initial_lm_damping=1.0))
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
    dk = _rel(out["cams"], ovars["cams"])
    dp = np.max(np.abs(out["poses"] - ovars["poses"]))
    dx = np.max(np.abs(out["points"] - ovars["points"]))
    print(f"merit {worst:.3g}, intrinsics {dk:.3g}, poses {dp:.3g}, points {dx:.3g}")
    assert dk < param_bar and dp < param_bar and dx < param_bar
    return out


def test_eucm_kernel_matches_oracle(sg, orc):
    rng = np.random.default_rng(11)
    n = 512
    intr = np.tile([300.0, 310.0, 320.0, 240.0, 0.5, 1.0], (n, 1))
    intr[:, 4] = rng.uniform(0.1, 0.9, n)
    intr[:, 5] = rng.uniform(0.3, 1.7, n)
    poses = np.stack([orc.se3_exp(np.concatenate([rng.normal(0, 0.3, 3), rng.normal(0, 1, 3)]))
                      for _ in range(n)])
    pts = np.stack([orc.so3_matrix(poses[i][:4]) @ np.array([rng.uniform(-2, 2), rng.uniform(-2, 2),
                                                              rng.uniform(1, 6)]) + poses[i][4:]
                    for i in range(n)])
    uv = rng.uniform([0, 0], [640, 480], (n, 2))
    for delta in (None, 1.0, 50.0):
        rk = sg.RobustKernel.huber(delta) if delta is not None else None
        g = sg.eval_eucm_terms(intr, poses, pts, uv, rk)
        o = orc.eval_eucm_terms(intr, poses, pts, uv, delta)
        for k in ("residual", "jac_intr", "jac_pose", "jac_point", "cost", "weight"):
            assert _rel(g[k], o[k]) < 1e-10, (k, delta, _rel(g[k], o[k]))


def test_ba_problem_free_intrinsics_matches_oracle(sg, orc):
    """BaProblem::new(4, 20), optimize_marg_points (shared EUCM camera Free, poses 0 and 1
    constant, points marginalized), 20 iterations from nu = 1, and the reference's gate."""
    p = sg.generate_ba_eucm(4, 20, seed=0)
    out = _compare_runs(sg, orc, p, sg.VarKind.Free, 20)
    V = sg.VarFamilies({"poses": sg.VarFamily(sg.VarKind.Free, sg.VarType.Isometry3, out["poses"]),
                        "points": sg.VarFamily(sg.VarKind.Marginalized, sg.VarType.Vec3, out["points"])})
    pose_rms, pt_rms = p.gt_errors(V)
    assert pose_rms < 0.01 and pt_rms < 0.1, (pose_rms, pt_rms)


@pytest.mark.parametrize("cams_kind", ["free", "conditioned"])
def test_eucm_ring_40cams_matches_oracle(sg, orc, cams_kind):
    """40 cameras / 2000 candidate points (about 34k observations, the visibility limit
    active), Huber 1 px; the intrinsics Free (the block couples to every camera) or
    Conditioned at the ground truth."""
    p = sg.generate_ba_eucm(40, 2000, seed=0)
    kind = sg.VarKind.Free if cams_kind == "free" else sg.VarKind.Conditioned
    _compare_runs(sg, orc, p, kind, 12, delta=1.0)


def test_eucm_free_intrinsics_pcg_follows_direct(sg):
    p = sg.generate_ba_eucm(40, 2000, seed=2)
    direct = _run_gpu(sg, p, sg.OptParams(num_iterations=10, initial_lm_damping=1.0), sg.VarKind.Free)
    pcg = _run_gpu(sg, p, sg.OptParams(num_iterations=10, initial_lm_damping=1.0,
                                       solver=sg.Solver.SchurPcgJacobi, pcg_relative_tolerance=1e-12,
                                       pcg_max_iterations=2000), sg.VarKind.Free)
    for g, o in zip(pcg[2], direct[2]):
        assert g.accepted == o.accepted
        assert abs(g.merit - o.merit) <= 1e-6 * o.merit
    assert _rel(pcg[1]["cams"], direct[1]["cams"]) < 1e-6


def test_eucm_conditioned_mixed_precision_final_cost(sg, orc):
    p = sg.generate_ba_eucm(40, 2000, seed=0)
    merit, _, trace = _run_gpu(sg, p, sg.OptParams(num_iterations=15, initial_lm_damping=1.0,
                                                   precision=sg.Precision.Mixed),
                               sg.VarKind.Conditioned, 1.0)
    sol, _, _ = orc.optimize_nlls(_oracle_fams(p, 1), _oracle_costs(p, 1.0),
                                  orc.default_params(num_iterations=15, initial_lm_damping=1.0))
    assert abs(merit - sol["final_cost"]) <= 1e-6 * sol["final_cost"], (merit, sol["final_cost"])


def test_eucm_free_intrinsics_is_bitwise_deterministic(sg):
    p = sg.generate_ba_eucm(40, 2000, seed=3)
    a = _run_gpu(sg, p, sg.OptParams(num_iterations=6, initial_lm_damping=1.0), sg.VarKind.Free)
    b = _run_gpu(sg, p, sg.OptParams(num_iterations=6, initial_lm_damping=1.0), sg.VarKind.Free)
    for k in ("cams", "poses", "points"):
        assert np.array_equal(a[1][k], b[1][k]), k


def test_eucm_free_intrinsics_covariance_blocks_match_dense_inverse(sg, orc):
    """Poses 0 and 1 fixed fix the gauge and the scale: the undamped final Hessian has full
    rank, so every covariance block, the intrinsics' included, is a block of numpy's dense
    inverse of the oracle's assembled H.  (Its condition number is about 1e10 in parameter
    units but 1e5 after diagonal scaling; the relative-pivot test of covariance_block is
    scale-invariant and accepts it.)"""
    p = sg.generate_ba_eucm(8, 60, seed=4)
    opt = sg.Optimizer(p.variables(), p.cost_fns(), sg.OptParams(num_iterations=8, initial_lm_damping=1.0))
    while opt.step().termination is None:
        pass
    intr, poses, points = opt.members("cams"), opt.members("poses"), opt.members("points")
    H, _ = orc.linearize_dense(_oracle_fams(p, 0, poses, points, intr), _oracle_costs(p), 0.0)
    # the parameters' units differ by orders of magnitude (pixels of focal length against
    # radians): invert the equilibrated matrix, whose condition number is about 1e5
    dsc = 1.0 / np.sqrt(np.diag(H))
    Hs = H * dsc[:, None] * dsc[None, :]
    assert np.linalg.cond(Hs) < 1e7
    Hinv = np.linalg.inv(Hs) * dsc[:, None] * dsc[None, :]
    nfree = 6 + 6 * (p.poses_init.shape[0] - 2)

    def index(name, member):
        if name == "cams":
            return 0, 6
        if name == "poses":
            return 6 + 6 * (member - 2), 6
        return nfree + 3 * member, 3

    queries = [("cams", 0, "cams", 0), ("cams", 0, "poses", 4), ("poses", 5, "cams", 0),
               ("cams", 0, "points", 7), ("points", 11, "cams", 0), ("poses", 3, "poses", 6),
               ("poses", 2, "points", 9), ("points", 5, "points", 5), ("points", 5, "points", 30)]
    for rf, rm, cf, cm in queries:
        blk = opt.covariance_block(rf, rm, cf, cm)
        (r0, dr), (c0, dc) = index(rf, rm), index(cf, cm)
        ref = Hinv[r0:r0 + dr, c0:c0 + dc]
        rel = np.max(np.abs(blk - ref)) / np.max(np.abs(ref))
        print(f"{rf}[{rm}] x {cf}[{cm}]: rel {rel:.3g}")
        assert rel < 1e-8, (rf, rm, cf, cm, rel)
    opt.close()


@pytest.mark.parametrize("solver", ["SchurBlockSparseLdlt", "SchurSparseLdlt"])
def test_ba_problem_inverse_blocks_match_dense_pseudo_inverse(sg, orc, solver):
    """The reference's schur_inverse_block_matches_dense_pseudo_inverse
    (crates/sophus_opt/src/nlls/damped_hessian.rs:610-700): BaProblem::new(4, 20) at its
    initial variables, the damped Hessian at nu = 1 (skip_final_hessian keeps that nu), and
    every free x free block (the shared camera and the free poses), every marginal diagonal
    block and every free x marginal block against the dense pseudo-inverse of H + nu N.
    The reference's bar is 1e-6 absolute; here 1e-9 relative to the block's largest entry.
    Both Schur solvers, as the reference's run_all_checks (damped_hessian.rs:558-590)."""
    p = sg.generate_ba_eucm(4, 20, seed=0)
    opt = sg.Optimizer(p.variables(), p.cost_fns(),
                       sg.OptParams(num_iterations=1, initial_lm_damping=1.0, skip_final_hessian=1,
                                    solver=getattr(sg.Solver, solver)))
    H, _ = orc.linearize_dense(_oracle_fams(p), _oracle_costs(p), 1.0)
    Hinv = np.linalg.pinv(H, rcond=1e-14)
    free = [("cams", 0)] + [("poses", m) for m in range(2, p.poses_init.shape[0])]
    marg = [("points", m) for m in range(p.points_init.shape[0])]
    start = {}
    off = 0
    for name, m in free:
        start[(name, m)] = (off, 6)
        off += 6
    for name, m in marg:
        start[(name, m)] = (off, 3)
        off += 3
    assert off == H.shape[0]
    pairs = [(a, b) for a in free for b in free] + [(a, a) for a in marg] + \
            [(a, b) for a in free for b in marg] + [(b, a) for a in free for b in marg[::5]]
    worst = 0.0
    for (rf, rm), (cf, cm) in pairs:
        blk = opt.inverse_block(rf, rm, cf, cm)
        (r0, dr), (c0, dc) = start[(rf, rm)], start[(cf, cm)]
        ref = Hinv[r0:r0 + dr, c0:c0 + dc]
        assert np.max(np.abs(blk - ref)) < 1e-6
        worst = max(worst, np.max(np.abs(blk - ref)) / np.max(np.abs(ref)))
    print(f"{len(pairs)} blocks, worst relative {worst:.3g}")
    assert worst < 1e-9
    opt.close()
