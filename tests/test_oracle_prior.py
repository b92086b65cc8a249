"""The oracle's Isometry3PriorCostTerm on the CPU (isometry3_prior.rs:46-88).

The oracle restates the term as the reference evaluates it: r = log(T E^-1) and the Jacobian by
forward-mode duals of log(exp(x) T E^-1) at x = 0, assembled by make with a precision
(evaluated_term.rs:286-327: r~ = w(|r|) r, gradient J^T W r~, Hessian J^T W J, cost
r~^T W r~).  It is pinned by:
- central differences of the residual under T <- exp(x) T, also where the residual's
  rotation is tiny (where the analytic J_l^-1 of the library's Lie code takes its series
  branch) and where it is large;
- the closed form at T = E: r = 0 and J = I (log(exp(x)) = x);
- the reference's gate of its camera-calibration problem with priors (cam_calib.rs:280-343):
  the BA ring with no pose held constant and priors of precision 1e4 I on poses 0 and 1 at
  their true values converges, and pose 2's translation ends within 0.1 of its true value.
"""
import numpy as np
import pytest


def _random_poses(sg, orc, n, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    return np.stack([orc.se3_exp(rng.normal(0, scale, 6)) for _ in range(n)])


def test_jacobian_matches_central_differences(sg, orc):
    rng = np.random.default_rng(1)
    E = _random_poses(sg, orc, 12, 2)
    T = np.stack([orc.se3_mul(orc.se3_exp(rng.normal(0, s, 6)), E[i])
                  for i, s in enumerate([1.0, 0.5, 0.1, 1e-2, 1e-3, 3e-4, 1e-4, 1e-5, 2.0, 0.7, 1e-3, 0.05])])
    W = np.repeat(np.eye(6)[None], 12, axis=0)
    r, J, _ = orc.eval_isometry3_prior_terms(T, E, W)
    h = 1e-6
    for i in range(12):
        Jn = np.zeros((6, 6))
        for k in range(6):
            e = np.zeros(6)
            e[k] = h
            rp, _, _ = orc.eval_isometry3_prior_terms(orc.se3_mul(orc.se3_exp(e), T[i])[None], E[i][None], W[:1])
            rm, _, _ = orc.eval_isometry3_prior_terms(orc.se3_mul(orc.se3_exp(-e), T[i])[None], E[i][None], W[:1])
            Jn[:, k] = (rp[0] - rm[0]) / (2 * h)
        assert np.max(np.abs(J[i] - Jn)) < 1e-7, (i, np.max(np.abs(J[i] - Jn)))


def test_identity_residual_and_jacobian(sg, orc):
    E = _random_poses(sg, orc, 5, 3)
    W = np.repeat(np.eye(6)[None], 5, axis=0)
    r, J, c = orc.eval_isometry3_prior_terms(E, E, W)
    assert np.max(np.abs(r)) < 1e-12 and np.max(np.abs(c)) < 1e-20
    assert np.max(np.abs(J - np.eye(6))) < 1e-12


def test_precision_and_huber_weight_the_cost(sg, orc):
    rng = np.random.default_rng(4)
    E = _random_poses(sg, orc, 6, 5)
    T = np.stack([orc.se3_mul(orc.se3_exp(rng.normal(0, 0.3, 6)), e) for e in E])
    A = rng.normal(size=(6, 6, 6))
    W = np.einsum("nij,nkj->nik", A, A) + np.eye(6)
    r, _, c = orc.eval_isometry3_prior_terms(T, E, W, huber_delta=0.2)
    for i in range(6):
        n = np.linalg.norm(r[i])
        w = np.sqrt(2 * 0.2 * n - 0.2 ** 2) / n if n > 0.2 else 1.0
        rw = w * r[i]
        assert abs(c[i] - rw @ W[i] @ rw) <= 1e-12 * abs(c[i])


def test_cam_calib_style_priors_fix_the_gauge(sg, orc):
    """cam_calib.rs:280-343 on the 10-camera ring: every pose Free, priors of precision 1e4 I
    on poses 0 and 1 at their true values, 10 LM iterations."""
    p = sg.generate_ba(sg.config_spec(0))
    fams = [("cams", 2, 1, p.intrinsics, []), ("points", 1, 2, p.points_init, []),
            ("poses", 0, 0, p.poses_init, [])]
    prior = np.hstack([p.poses_gt[:2], np.repeat((1e4 * np.eye(6)).reshape(1, 36), 2, axis=0)])
    costs = orc.ba_costs(p) + [(3, ["poses"], np.array([[0], [1]]), prior, None)]
    sol, mem, tr = orc.optimize_nlls(fams, costs, orc.default_params(num_iterations=10, initial_lm_damping=1.0))
    assert sol["final_cost"] < 0.1 * sol["initial_cost"]
    assert np.max(np.abs(mem["poses"][2, 4:] - p.poses_gt[2, 4:])) < 0.1  # the reference's gate
    # the priors hold poses 0 and 1 near their means (precision 1e4 against pixel residuals)
    assert np.max(np.abs(mem["poses"][:2, 4:] - p.poses_gt[:2, 4:])) < 0.01
