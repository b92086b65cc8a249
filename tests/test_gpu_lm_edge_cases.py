"""The HIP path against the FP64 oracle on the reference's LM edge cases, through the C ABI:
duplicated (pose, point) terms (summed into one term before damping, cost_fn.rs:354-393,
evaluated_term.rs:52-57, cost_system.rs:229-235), a run that ends in MaxConsecutiveRejects
(optimizer.rs:463-485), a NaN observation (optimizer_step.rs:394-408) and an LM run whose reduced camera system has zero pivots
(dense_ldlt.rs:164-256: a zero pivot solves to 0).

Bars: the same accept/reject sequence and termination, merit 1e-9 relative at every step,
parameters 1e-7.
"""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _gpu_trace(sg, problem, params):
    opt = sg.Optimizer(problem.variables(), problem.cost_fns(1.0), params)
    trace = []
    while True:
        info = opt.step()
        trace.append(info)
        if info.termination is not None:
            break
    out = {n: opt.members(n) for n in ("poses", "points")}
    opt.close()
    return out, trace


def _oracle_trace(orc, problem, **kw):
    sol, ovars, otrace = orc.optimize_nlls(orc.ba_families(problem), orc.ba_costs(problem),
                                           orc.default_params(**kw))
    return sol, ovars, otrace


def _compare(sg, trace, out, sol, ovars, otrace, cost_rel=1e-9, param_abs=1e-7):
    assert len(trace) == len(otrace), (len(trace), len(otrace))
    for g, o in zip(trace, otrace):
        assert g.accepted == o["accepted"]
        if np.isfinite(o["merit"]):
            assert abs(g.merit - o["merit"]) <= cost_rel * abs(o["merit"]), (g.merit, o["merit"])
        else:
            assert not np.isfinite(g.merit)
        assert abs(g.lm_damping - o["lm_damping"]) <= 1e-9 * o["lm_damping"]
    assert int(trace[-1].termination) == sol["termination"]
    assert np.max(np.abs(out["poses"] - ovars["poses"])) < param_abs
    assert np.max(np.abs(out["points"] - ovars["points"])) < param_abs


def test_duplicated_observations_match_oracle(sg, orc):
    """40 observations repeated with different pixels: each (pose, point) pair is one reduced
    term, so nu is added once per pair and the Schur diagonal gets the repeated
    observations' same-camera cross products."""
    p = sg.generate_ba(sg.config_spec(0))
    rng = np.random.default_rng(0)
    dup = rng.choice(len(p.obs_idx), 40, replace=False)
    q = dataclasses.replace(p, obs_idx=np.concatenate([p.obs_idx, p.obs_idx[dup]]),
                            obs_uv=np.concatenate([p.obs_uv, p.obs_uv[dup] + rng.normal(0, 2.0, (40, 2))]))
    params = sg.OptParams(num_iterations=20, initial_lm_damping=1.0)
    out, trace = _gpu_trace(sg, q, params)
    sol, ovars, otrace = _oracle_trace(orc, q, num_iterations=20, initial_lm_damping=1.0)
    _compare(sg, trace, out, sol, ovars, otrace)


@pytest.mark.parametrize("max_rejects", [20, 3])
def test_consecutive_rejects_terminate_like_oracle(sg, orc, max_rejects):
    """A Marginalized point that no term touches: its H_mm is exactly zero (the damping is per
    term, cost_system.rs:229-235), the 3x3 inverse fails (SingularMargBlock,
    schur_ldlt.rs:101), every solve is a reject that doubles nu, and the run ends in
    MaxConsecutiveRejects after max_consecutive_rejects steps with the variables unchanged."""
    p = sg.generate_ba(sg.config_spec(0))
    extra = np.array([[0.1, 0.2, 0.3]])
    q = dataclasses.replace(p, points_init=np.vstack([p.points_init, extra]),
                            points_gt=np.vstack([p.points_gt, extra]))
    params = sg.OptParams(num_iterations=50, initial_lm_damping=1.0, max_consecutive_rejects=max_rejects)
    out, trace = _gpu_trace(sg, q, params)
    sol, ovars, otrace = _oracle_trace(orc, q, num_iterations=50, initial_lm_damping=1.0,
                                       max_consecutive_rejects=max_rejects)
    assert sol["termination"] == int(sg.TerminationReason.MaxConsecutiveRejects)
    assert len(otrace) == max_rejects and not any(o["accepted"] for o in otrace)
    _compare(sg, trace, out, sol, ovars, otrace)
    assert trace[-1].lm_damping == 2.0 ** max_rejects
    assert np.array_equal(out["poses"], q.poses_init) and np.array_equal(out["points"], q.points_init)


def test_nan_observation_ends_like_the_reference(sg, orc):
    """A NaN pixel leaves H finite but makes the merit and the gradient rows it touches NaN;
    the solve spreads NaN over every dx entry.  The reference folds max |dx| with f64::max,
    which skips NaN (optimizer_step.rs:397), so max_step is 0 < parameter_tolerance: the run
    ends at its first step with ParameterTolerance, nothing accepted (NaN < NaN is false)
    and the variables unchanged.  The oracle and the GPU path fold the same way."""
    p = sg.generate_ba(sg.config_spec(0))
    uv = p.obs_uv.copy()
    uv[17, 0] = np.nan
    q = dataclasses.replace(p, obs_uv=uv)
    params = sg.OptParams(num_iterations=50, initial_lm_damping=1.0)
    out, trace = _gpu_trace(sg, q, params)
    sol, ovars, otrace = _oracle_trace(orc, q, num_iterations=50, initial_lm_damping=1.0)
    assert sol["termination"] == int(sg.TerminationReason.ParameterTolerance)
    assert len(otrace) == 1 and not otrace[0]["accepted"]
    _compare(sg, trace, out, sol, ovars, otrace)
    assert np.array_equal(out["poses"], q.poses_init) and np.array_equal(out["points"], q.points_init)
    assert trace[0].max_abs_delta == 0.0 and np.isnan(trace[0].new_merit)


def test_camera_without_observations_gives_zero_pivots(sg, orc):
    """A free pose that no term touches: its block of the reduced camera system is exactly 0
    (the damping is per term), all six pivots fall under the PSD rule and solve to 0, and
    the rest of the trajectory is the oracle's."""
    p = sg.generate_ba(sg.config_spec(0))
    q = dataclasses.replace(p, poses_init=np.vstack([p.poses_init, p.poses_init[3:4]]),
                            poses_gt=np.vstack([p.poses_gt, p.poses_gt[3:4]]))
    params = sg.OptParams(num_iterations=20, initial_lm_damping=1.0)
    opt = sg.Optimizer(q.variables(), q.cost_fns(1.0), params)
    trace = []
    zero_pivots = []
    while True:
        info = opt.step()
        trace.append(info)
        zero_pivots.append(int(np.sum(opt.rcs_pivots() == 0.0)))
        if info.termination is not None:
            break
    out = {n: opt.members(n) for n in ("poses", "points")}
    opt.close()
    assert all(z == 6 for z in zero_pivots), zero_pivots
    sol, ovars, otrace = _oracle_trace(orc, q, num_iterations=20, initial_lm_damping=1.0)
    _compare(sg, trace, out, sol, ovars, otrace)
    assert np.array_equal(out["poses"][10], q.poses_init[10])
