"""GPU tests of the batched small-window solver: every window's LM run against the FP64
oracle run on that window alone (same accept/reject count, final cost 1e-9 relative in
FP64, 1e-6 with FP32 Schur products), graph replay against plain launches, reset, and
windows of every size in the workload's range."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _windows(sg, n, seed_base=7):
    return [sg.generate_ba(sg.window_spec(i, seed_base)) for i in range(n)]


def _oracle_window(orc, w, iters):
    sol, vars_, trace = orc.optimize_nlls(orc.ba_families(w), orc.ba_costs(w),
                                          orc.default_params(num_iterations=iters, initial_lm_damping=1.0))
    return sol, vars_


def test_batch_fp64_matches_oracle_per_window(sg, orc):
    ws = _windows(sg, 10)
    # include the extremes of the workload's size range
    ws.append(sg.generate_ba(sg.BaSpec(layout=0, num_cameras=5, num_points=100, obs_per_point_ring=0,
                                       min_obs=2, max_obs=8, seed=1)))
    ws.append(sg.generate_ba(sg.BaSpec(layout=0, num_cameras=30, num_points=2000, obs_per_point_ring=0,
                                       min_obs=2, max_obs=8, seed=2)))
    iters = 20
    b = sg.Batch(ws, sg.OptParams(num_iterations=iters, initial_lm_damping=1.0))
    print("[batch] autotuner", b.info())
    b.run(iters)
    res = b.results()
    for p, w in enumerate(ws):
        sol, vars_ = _oracle_window(orc, w, iters)
        rel = abs(res["final_cost"][p] - sol["final_cost"]) / sol["final_cost"]
        assert rel <= 1e-9, (p, rel)
        assert res["steps"][p] == sol["num_steps"] and res["termination"][p] == sol["termination"]
        sl = b.window_slice(p)
        assert np.max(np.abs(res["poses"][sl["poses"]] - vars_["poses"])) < 1e-7
        assert np.max(np.abs(res["points"][sl["points"]] - vars_["points"])) < 1e-7
    b.close()


def test_batch_mixed_precision_final_cost(sg, orc):
    ws = _windows(sg, 6, seed_base=9)
    b = sg.Batch(ws, sg.OptParams(num_iterations=20, initial_lm_damping=1.0, precision=sg.Precision.Mixed))
    b.run(20)
    res = b.results()
    for p, w in enumerate(ws):
        sol, _ = _oracle_window(orc, w, 20)
        assert abs(res["final_cost"][p] - sol["final_cost"]) <= 1e-6 * sol["final_cost"]
    b.close()


def test_batch_graph_replay_and_reset_reproduce(sg):
    ws = _windows(sg, 24, seed_base=11)
    b = sg.Batch(ws, sg.OptParams(num_iterations=15, initial_lm_damping=1.0))
    b.run(15, graph=False)
    r1 = b.results()
    b.reset()
    b.run(15, graph=True)
    r2 = b.results()
    b.reset()
    b.run(15, graph=True)  # replay of the same graph
    r3 = b.results()
    for r in (r2, r3):
        assert np.array_equal(r["steps"], r1["steps"])
        assert np.allclose(r["final_cost"], r1["final_cost"], rtol=1e-10, atol=0)
    b.close()


def test_persistent_launch_matches_per_iteration_launches(sg, orc):
    """One persistent launch (a queue of windows over the resident workgroups, each window's
    whole LM run in one workgroup) gives the per-iteration launches' results, also when
    replayed from a graph and when split into two runs; the device stamps order each
    window's start and end inside the run's wall time."""
    ws = _windows(sg, 40, seed_base=13)
    ws.append(sg.generate_ba(sg.BaSpec(layout=0, num_cameras=30, num_points=2000, obs_per_point_ring=0,
                                       min_obs=2, max_obs=8, seed=3)))
    iters = 20
    b = sg.Batch(ws, sg.OptParams(num_iterations=iters, initial_lm_damping=1.0))
    info = b.info()
    print("[batch] persistent grid", info["persistent_grid"], "threads", info["threads"])
    assert 1 <= info["persistent_grid"] <= len(ws)
    b.run(iters, graph=False)
    ref = b.results()
    runs = []
    for graph in (False, True, True):
        b.reset()
        secs = b.run_persistent(iters, graph=graph)
        runs.append((b.results(), b.latencies(), secs))
    b.reset()
    b.run_persistent(5, graph=False)
    b.run_persistent(iters - 5, graph=False)
    runs.append((b.results(), None, None))
    for r, lat, secs in runs:
        assert np.array_equal(r["steps"], ref["steps"])
        assert np.array_equal(r["termination"], ref["termination"])
        assert np.allclose(r["final_cost"], ref["final_cost"], rtol=1e-10, atol=0)
        assert np.allclose(r["poses"], ref["poses"], rtol=0, atol=1e-9)
        if lat is not None:
            assert np.all(lat["start_ms"] >= 0.0) and np.all(lat["end_ms"] >= lat["start_ms"])
            assert lat["end_ms"].max() <= 1e3 * secs
    for p in (0, len(ws) - 1):  # the per-window oracle, through the persistent path
        sol, _ = _oracle_window(orc, ws[p], iters)
        r = runs[0][0]
        assert abs(r["final_cost"][p] - sol["final_cost"]) <= 1e-9 * sol["final_cost"]
        assert r["steps"][p] == sol["num_steps"] and r["termination"][p] == sol["termination"]
    b.close()
