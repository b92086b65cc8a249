"""Train the learned block-Jacobi correction on the GPU and write its weights.

    python sophus-rs_amd/train_precond.py [--out sophus-rs_amd/precond_weights.f32]

Snapshots come from LM runs of the solver itself with the plain block-Jacobi PCG on seeded
random-walk BA problems: the 500-camera workload's generator and density (200 points per
camera, geometric track lengths with p = 0.1) at 200 cameras, 40 LM iterations each, about
10k reduced camera systems. Two further problems are held out for the validation loss.
After training, the gain is measured on held-out walks the network never saw (seeds past
the training ones): total PCG iterations over LM runs with the learned correction against
plain block-Jacobi, at the same CG tolerance. The weights are written as raw
little-endian float32 (SGPU_PRECOND_WEIGHTS values) and a JSON summary beside them.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))

import sophus_opt_gpu as sg  # noqa: E402

DEFAULT = sg.PrecondTrainParams(num_problems=250, num_cameras=200, num_points=40000, geometric_p=0.1,
                                lm_iterations=40, epochs=30, probes=4, learning_rate=1e-3, seed=2024,
                                batch_snapshots=16, holdout_problems=2)


def pcg_iterations(problem, solver, weights=None, iterations=20, tol=1e-8):
    """Total CG iterations and the final cost of an LM run with the given PCG solver."""
    params = sg.OptParams(num_iterations=iterations, initial_lm_damping=1.0, solver=solver,
                          pcg_relative_tolerance=tol, pcg_max_iterations=3000, precond_weights=weights)
    opt = sg.Optimizer(problem.variables(), problem.cost_fns(1.0), params)
    total, steps = 0, 0
    while True:
        info = opt.step()
        total += info.pcg_iterations
        steps += 1
        if info.termination is not None:
            break
    merit = opt.merit()
    opt.close()
    return total, steps, merit


def held_out_gain(tp, weights, count=4, iterations=20):
    """Per held-out walk: total PCG iterations learned vs block-Jacobi (same tolerance)."""
    rows = []
    first = tp.seed + tp.num_problems + tp.holdout_problems + 1000
    for k in range(count):
        spec = sg.BaSpec(layout=1, num_cameras=tp.num_cameras, num_points=tp.num_points,
                         geometric_p=tp.geometric_p, seed=first + k)
        p = sg.generate_ba(spec)
        it_j, st_j, m_j = pcg_iterations(p, sg.Solver.SchurPcgJacobi, iterations=iterations)
        it_l, st_l, m_l = pcg_iterations(p, sg.Solver.SchurPcgLearned, weights, iterations=iterations)
        rows.append({"seed": first + k, "jacobi": it_j, "learned": it_l, "ratio": it_l / it_j,
                     "lm_steps": [st_j, st_l], "final_cost_rel_diff": abs(m_l - m_j) / m_j})
    ratios = np.array([r["ratio"] for r in rows])
    return {"walks": rows, "mean_ratio": float(ratios.mean()),
            "total_ratio": float(sum(r["learned"] for r in rows) / sum(r["jacobi"] for r in rows)),
            "min_ratio": float(ratios.min()), "max_ratio": float(ratios.max())}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(HERE / "precond_weights.f32"))
    ap.add_argument("--epochs", type=int, default=DEFAULT.epochs)
    ap.add_argument("--problems", type=int, default=DEFAULT.num_problems)
    args = ap.parse_args()
    tp = DEFAULT
    tp.epochs = args.epochs
    tp.num_problems = args.problems
    t = time.perf_counter()
    w, hist, val, nsnap = sg.train_preconditioner(tp)
    dt = time.perf_counter() - t
    print(json.dumps({"snapshots": nsnap, "seconds": round(dt, 2), "train_loss": [hist[0], hist[-1]],
                      "val_loss": [val[0], val[-1]]}), flush=True)
    t = time.perf_counter()
    gain = held_out_gain(tp, w)
    print(json.dumps(gain), flush=True)
    out = Path(args.out)
    w.astype("<f4").tofile(out)
    summary = {"params": tp.__dict__, "snapshots": nsnap, "train_seconds": dt,
               "loss_history": hist.tolist(), "val_loss_history": val.tolist(),
               "held_out_pcg": gain, "held_out_seconds": time.perf_counter() - t}
    out.with_suffix(".json").write_text(json.dumps(summary, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
