"""Distinct optimizer handles used from different host threads at once (the reference's
optimizers are Send + Sync, cost_term.rs:35): each handle owns its stream, device arena and
LM state, and the last-error message is per thread.  Two threads step two handles on
different problems concurrently; each run is bitwise the same run done alone (the FP64 direct
path has no atomics in its sums).  An error raised on one thread leaves the other's
message alone.
"""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(sg, p, iterations):
    opt = sg.Optimizer(p.variables(), p.cost_fns(1.0),
                       sg.OptParams(num_iterations=iterations, initial_lm_damping=1.0))
    merits = []
    while True:
        info = opt.step()
        merits.append(info.merit)
        if info.termination is not None:
            break
    out = {n: opt.members(n) for n in ("poses", "points")}
    opt.close()
    return merits, out


def test_two_handles_on_two_threads_match_sequential_runs(sg):
    problems = [sg.generate_ba(sg.BaSpec(layout=1, num_cameras=80, num_points=4000, geometric_p=0.15, seed=s))
                for s in (21, 22)]
    alone = [_run(sg, p, 12) for p in problems]
    results = [None, None]
    errors = []

    def work(k):
        try:
            for _ in range(3):  # several rounds so the threads' launches interleave
                results[k] = _run(sg, problems[k], 12)
        except Exception as e:  # reported on the main thread
            errors.append(e)

    threads = [threading.Thread(target=work, args=(k,)) for k in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors
    for (m0, o0), (m1, o1) in zip(alone, results):
        assert m0 == m1
        for n in ("poses", "points"):
            assert np.array_equal(o0[n], o1[n])


def test_last_error_is_per_thread(sg):
    p = sg.generate_ba(sg.config_spec(0))
    with pytest.raises(Exception, match="pcg_forcing"):
        sg.Optimizer(p.variables(), p.cost_fns(1.0), sg.OptParams(pcg_forcing=2.0))
    mine = sg.library().sgpu_last_error().decode()
    assert "pcg_forcing" in mine
    seen = {}

    def other():
        try:
            sg.Optimizer(p.variables(), p.cost_fns(1.0), sg.OptParams(precision=7))
        except Exception as e:
            seen["raised"] = str(e)
        seen["message"] = sg.library().sgpu_last_error().decode()

    t = threading.Thread(target=other)
    t.start()
    t.join(timeout=60)
    assert "precision" in seen.get("raised", "") and "precision" in seen["message"]
    assert sg.library().sgpu_last_error().decode() == mine
