"""Train the learned block-Jacobi correction on the GPU and write its weights.

    python sophus-rs_amd/train_precond.py [--out sophus-rs_amd/precond_weights.f32]

Snapshots come from LM runs of the solver itself on seeded random-walk BA problems
(the 500-camera workload's generator at a smaller size); the weights are written as
raw little-endian float32 (SGPU_PRECOND_WEIGHTS values) and a JSON summary beside them.
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

DEFAULT = sg.PrecondTrainParams(num_problems=24, num_cameras=120, num_points=6000, geometric_p=0.12,
                                lm_iterations=10, epochs=20, probes=4, learning_rate=1e-3, seed=2024)


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
    w, hist, nsnap = sg.train_preconditioner(tp)
    dt = time.perf_counter() - t
    out = Path(args.out)
    w.astype("<f4").tofile(out)
    summary = {"params": tp.__dict__, "snapshots": nsnap, "seconds": dt, "loss_history": hist.tolist()}
    out.with_suffix(".json").write_text(json.dumps(summary, indent=1))
    print(json.dumps({"snapshots": nsnap, "seconds": round(dt, 2), "loss_first": hist[0], "loss_last": hist[-1]}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
