"""SchurSparseLdlt on the GPU against the oracle's SparseLdlt (sparse_ldlt.rs), through the C ABI.

The GPU runs the same supernodal factorization as for SchurBlockSparseLdlt with the scalar
pivot rule of SparseLdlt (|d_jj| <= max(|d_jj|, 1) 1e-12 -> 0, decided per pivot) and its
solve's division by a zero pivot (the component is NaN).  The oracle is pinned on the CPU
(tests/test_oracle_sparse_ldlt.py).  Bars: solves 1e-10 relative (1e-3 where the kept pivot
makes the condition number 3e12); the LM run of the 10-camera ring (BaProblem-style
workload of the reference's CPU case, "Schur complement plus sparse LDLt, 20 LM
iterations") and a 60-camera walk step by step: the same accept sequence, merit 1e-9,
parameters 1e-7.
"""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _spd_blocks(nb, band, seed):
    rng = np.random.default_rng(seed)
    n = 6 * nb
    A = np.zeros((n, n))
    for i in range(nb):
        for j in range(max(0, i - band), i + 1):
            B = rng.normal(size=(6, 6))
            A[6 * i:6 * i + 6, 6 * j:6 * j + 6] = B
            A[6 * j:6 * j + 6, 6 * i:6 * i + 6] = B.T
    return 0.5 * (A + A.T) + np.eye(n) * (np.abs(A).sum(axis=1).max() + 1.0)


def _ldl_blocks(d, nb, seed):
    rng = np.random.default_rng(seed)
    A = np.zeros((6 * nb, 6 * nb))
    for i in range(nb):
        L = np.tril(rng.normal(0, 0.1, (6, 6)), -1) + np.eye(6)
        A[6 * i:6 * i + 6, 6 * i:6 * i + 6] = L @ np.diag(d) @ L.T
    return 0.5 * (A + A.T)


@pytest.mark.parametrize("nb,band", [(1, 0), (7, 2), (60, 5), (300, 12)])
def test_spd_solve_matches_oracle_and_numpy(sg, orc, nb, band):
    A = _spd_blocks(nb, band, nb)
    b = np.random.default_rng(nb + 3).normal(size=6 * nb)
    x = sg.sparse_ldlt_solve(A, b)
    assert _rel(x, np.linalg.solve(A, b)) < 1e-10
    assert _rel(x, orc.sparse_ldlt_solve_dense(A, [6] * nb, b)) < 1e-10


def test_kept_small_pivot_matches_oracle(sg, orc):
    """A pivot of 3e-7 after one of 1e6 in every block: kept by SparseLdlt's rule (the
    block rule would drop it), so the solve is the exact one."""
    A = _ldl_blocks(np.array([1e6, 3e-7, 2.0, 5.0, 8.0, 7.0]), 5, 8)
    b = np.random.default_rng(9).normal(size=30)
    x = sg.sparse_ldlt_solve(A, b)
    xo = orc.sparse_ldlt_solve_dense(A, [6] * 5, b)
    assert np.all(np.isfinite(x)) and _rel(x, xo) < 1e-3
    xb = sg.block_ldlt_solve(A, b)
    assert np.max(np.abs(xb)) < 1e-2 * np.max(np.abs(xo))  # the block rule drops it


def test_zero_pivots_divide_like_the_reference(sg, orc):
    """sparse_ldlt.rs:690-760's [[eps, 1], [1, -eps]] (eps = 1e-12, both pivots zero) in the
    top of a 6x6 block with an identity below it, and a block with a pivot of 1e-13: the
    zero pivots' components are not finite, as the reference's division by zero makes them,
    and the rest of the solution is the oracle's."""
    A = np.zeros((12, 12))
    A[:6, :6] = np.eye(6)
    A[:2, :2] = [[1e-12, 1.0], [1.0, -1e-12]]
    A[6:, 6:] = np.diag([4.0, 1e-13, 2.0, 5.0, 8.0, 7.0])
    b = np.arange(1.0, 13.0)
    x = sg.sparse_ldlt_solve(A, b)
    xo = orc.sparse_ldlt_solve_dense(A, [6, 6], b)
    bad = ~np.isfinite(xo)
    assert list(np.flatnonzero(bad)) == [0, 1, 7]
    assert np.array_equal(~np.isfinite(x), bad)
    assert np.allclose(x[~bad], xo[~bad], rtol=1e-14, atol=0)
    assert np.all(np.isfinite(sg.block_ldlt_solve(A, b)))


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


def _compare(trace, out, sol, ovars, otrace):
    assert len(trace) == len(otrace)
    worst = 0.0
    for g, o in zip(trace, otrace):
        assert g.accepted == o["accepted"]
        worst = max(worst, abs(g.merit - o["merit"]) / o["merit"])
        assert abs(g.lm_damping - o["lm_damping"]) <= 1e-9 * o["lm_damping"]
    assert worst <= 1e-9, worst
    assert int(trace[-1].termination) == sol["termination"]
    for n in ("poses", "points"):
        fin = np.isfinite(ovars[n])
        assert np.array_equal(np.isfinite(out[n]), fin)
        assert np.max(np.abs(out[n][fin] - ovars[n][fin])) < 1e-7
    return worst


def test_ring_schur_sparse_ldlt_matches_oracle(sg, orc):
    """The reference's CPU case: 10 cameras, 200 points, Huber 1 px, camera 0 fixed,
    Schur complement plus sparse LDLt, 20 LM iterations."""
    p = sg.generate_ba(sg.config_spec(0))
    out, trace = _gpu_trace(sg, p, sg.OptParams(num_iterations=20, initial_lm_damping=1.0,
                                                solver=sg.Solver.SchurSparseLdlt))
    orc.set_rcs_order(sg.analyze_rcs_order(p.variables(), p.cost_fns(1.0)))
    try:
        sol, ovars, otrace = orc.optimize_nlls(orc.ba_families(p), orc.ba_costs(p),
                                               orc.default_params(num_iterations=20, initial_lm_damping=1.0,
                                                                  solver=int(sg.Solver.SchurSparseLdlt)))
    finally:
        orc.set_rcs_order(None)
    print(f"SchurSparseLdlt ring: merit {_compare(trace, out, sol, ovars, otrace):.3g}")


def test_unobserved_camera_under_the_scalar_rule_matches_oracle(sg, orc):
    """A free pose no term touches: its six pivots are zero under either rule.  SparseLdlt's
    solve divides 0 by them, the step's delta.delta is NaN, the gain ratio falls to the
    reference's fallback (1 when the merit decreased) and the step is taken with that pose
    NaN (optimizer_helpers.rs:159-173); the block rule solves those components to 0
    instead.  The GPU follows the oracle either way.

    With the gain ratio at 1, nu falls by 3 every step (3e-10 after 20), and the ring's free
    scale (only camera 0 is held) is then barely damped: after 20 steps the oracle's own FMA
    twin moves the poses by 1e-6, after 12 by 8e-11.  The run stops at 12."""
    p = sg.generate_ba(sg.config_spec(0))
    q = dataclasses.replace(p, poses_init=np.vstack([p.poses_init, p.poses_init[3:4]]),
                            poses_gt=np.vstack([p.poses_gt, p.poses_gt[3:4]]))
    out, trace = _gpu_trace(sg, q, sg.OptParams(num_iterations=12, initial_lm_damping=1.0,
                                                solver=sg.Solver.SchurSparseLdlt))
    orc.set_rcs_order(sg.analyze_rcs_order(q.variables(), q.cost_fns(1.0)))
    try:
        sol, ovars, otrace = orc.optimize_nlls(orc.ba_families(q), orc.ba_costs(q),
                                               orc.default_params(num_iterations=12, initial_lm_damping=1.0,
                                                                  solver=int(sg.Solver.SchurSparseLdlt)))
    finally:
        orc.set_rcs_order(None)
    _compare(trace, out, sol, ovars, otrace)
    assert np.all(np.isnan(out["poses"][10])) and np.all(np.isfinite(out["poses"][:10]))
    assert np.isnan(trace[0].gain_ratio) or trace[0].gain_ratio == 1.0


def test_walk_schur_sparse_ldlt_matches_oracle(sg, orc):
    """A 60-camera random walk (3,000 points, camera 0 constant) under SchurSparseLdlt, ten LM
    iterations step by step in the GPU's elimination order: the walk's camera graph has long
    chains of separators, so the supernodal factorization takes many rounds with the scalar
    pivot rule in every diagonal block."""
    spec = sg.BaSpec(layout=1, num_cameras=60, num_points=3000, geometric_p=0.15, seed=5)
    p = sg.generate_ba(spec)
    out, trace = _gpu_trace(sg, p, sg.OptParams(num_iterations=10, initial_lm_damping=1.0,
                                                solver=sg.Solver.SchurSparseLdlt))
    orc.set_rcs_order(sg.analyze_rcs_order(p.variables(), p.cost_fns(1.0)))
    try:
        sol, ovars, otrace = orc.optimize_nlls(orc.ba_families(p), orc.ba_costs(p),
                                               orc.default_params(num_iterations=10, initial_lm_damping=1.0,
                                                                  solver=int(sg.Solver.SchurSparseLdlt)))
    finally:
        orc.set_rcs_order(None)
    print(f"SchurSparseLdlt walk: merit {_compare(trace, out, sol, ovars, otrace):.3g}")
