"""Benchmark: Schur-complement LM bundle adjustment on one MI355X.

Workload: the 2000-camera / 800k-point random-walk BA (5,947,769 observations, pinhole,
Huber 1 px, camera 0 fixed, nu0 = 1), inputs resident in HBM before the timed region.

- value: observations per second of the plain FP64 GPU path (Schur + direct block LDLt),
  N_o / (seconds per LM iteration) over exactly --steps timed iterations after --warmup
  (tolerances disabled so every iteration does the full work).
- roofline: the dominant kernel, one round of panel steps of the RCS factorization (one
  k_mf_step launch holding a panel step of every ready supernode; rounds with >= 640
  tiles run as k_mf_panel + k_mf_update): rank-48 trailing updates of HBM-resident
  fronts, 6 flop/byte, so HBM-bound; algorithmic bytes per round from the symbolic factor
  / the round's average time from HIP events around its launches on the optimizer's
  stream, against HBM peak.
- time_to_tol: F* = the FP64 direct path's final cost (function tolerance 1e-10, at most
  100 iterations); each path's first iteration within 1e-6 of F*, its time and
  obs/s = N_o x iterations / time: FP64 direct, block-Jacobi PCG, learned PCG, mixed.
- cpu_baseline: one LM iteration of the FP64 oracle on this host, one core.

Prints ONE JSON line (stdout).  Progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "sophus-rs_amd"))

METRIC = "BA time-to-tol @2k cams/6M obs: learned+FP32 vs plain FP64 GPU; obs/s"
HBM_PEAK_GBS = 8000.0
TOL_REL = 1e-6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(problem, label: str) -> dict:
    """One LM iteration of the FP64 oracle (single-threaded C++ restatement of the
    reference path) on the same problem, on this host's cores."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle

    params = oracle.default_params(num_iterations=1, initial_lm_damping=1.0)
    oracle.optimize_nlls(oracle.ba_families(problem), oracle.ba_costs(problem), params)
    step_s = oracle.last_step_seconds()
    return {
        "value": problem.num_obs / step_s,
        "unit": "obs/s",
        "cores": 1,
        "kind": "port",
        "sample": f"one LM iteration (linearize, Schur, RCS LDLt, update, merit) of the FP64 oracle "
                  f"on the full {label} problem: {step_s:.2f} s",
    }


def run_path(sg, problem, max_iter=100, **kw):
    """Run LM from the initial values; per-iteration (cumulative s, merit, pcg its)."""
    params = sg.OptParams(num_iterations=max_iter, initial_lm_damping=1.0, **kw)
    opt = sg.Optimizer(problem.variables(), problem.cost_fns(1.0), params)
    hist = []
    t0 = time.perf_counter()
    while True:
        info = opt.step()
        hist.append((time.perf_counter() - t0, info.merit, info.pcg_iterations))
        if info.termination is not None:
            break
    opt.close()
    return hist


def tol_summary(hist, f_star, n_obs):
    hit = next((i for i, (_, m, _) in enumerate(hist) if (m - f_star) / f_star <= TOL_REL), None)
    best = min(m for _, m, _ in hist)
    out = {"iterations_run": len(hist), "seconds_run": hist[-1][0], "final_cost": hist[-1][1],
           "best_rel_to_f_star": (best - f_star) / f_star,
           "pcg_iterations": int(sum(h[2] for h in hist))}
    if hit is None:
        out.update({"reached": False, "iterations": None, "seconds": None, "obs_per_s": None})
    else:
        t = hist[hit][0]
        out.update({"reached": True, "iterations": hit + 1, "seconds": t,
                    "obs_per_s": n_obs * (hit + 1) / t})
    return out


def batch_leg(sg, n_windows: int, iters: int = 20) -> dict:
    """The batched workload: n_windows independent ring windows (5-30 cameras, 100-2000
    points, 2-8 cameras per point, seeded by window id), 20 LM iterations each, FP64 and
    FP32-assembly.  Throughput from a graph replay of all iterations; per-window latency
    = the time at the end of the iteration in which the window terminated (one launch
    per iteration, synchronized)."""
    t = time.perf_counter()
    windows = [sg.generate_ba(sg.window_spec(i)) for i in range(n_windows)]
    n_obs = int(sum(w.num_obs for w in windows))
    gen_s = time.perf_counter() - t
    out = {"windows": n_windows, "observations": n_obs, "lm_iterations": iters, "generate_s": gen_s}
    for name, prec in (("f64", sg.Precision.F64), ("fp32_assembly", sg.Precision.Mixed)):
        t = time.perf_counter()
        b = sg.Batch(windows, sg.OptParams(num_iterations=iters, initial_lm_damping=1.0, precision=prec))
        create_s = time.perf_counter() - t
        secs = b.run(iters, graph=True)
        res = b.results()
        # latency: synchronized single iterations from a reset state
        b.reset()
        lat = np.full(n_windows, np.nan)
        t0 = time.perf_counter()
        for it in range(iters):
            b.run(1, graph=False)
            now = time.perf_counter() - t0
            term = b.termination()
            lat[np.isnan(lat) & (term != 0)] = now
        info = b.info()
        b.close()
        out[name] = {"seconds": secs, "problems_per_s": n_windows / secs,
                     "obs_iterations_per_s": float(np.sum(res["steps"].astype(np.float64) *
                                                          np.diff(b.obs_ptr))) / secs,
                     "latency_p50_ms": float(np.nanpercentile(lat, 50) * 1e3),
                     "latency_p99_ms": float(np.nanpercentile(lat, 99) * 1e3),
                     "mean_steps": float(res["steps"].mean()),
                     "final_over_initial_cost_median": float(np.median(res["final_cost"] / res["initial_cost"])),
                     "create_s": create_s, "autotuned_threads": info["threads"],
                     "autotune_ms": info["tune_ms"]}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=2, help="workload index (0, 1, 2)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--no-tol", action="store_true", help="skip the time-to-tolerance runs")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-window leg")
    ap.add_argument("--batch-windows", type=int, default=4096)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}): this build runs the "
              f"Schur LM path on a single MI355X; the multi-GPU design is documented, not built",
              file=sys.stderr)
        return 2
    if int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs as a single process on one GPU", file=sys.stderr)
        return 2
    import sophus_opt_gpu as sg

    sg.library()
    spec = sg.config_spec(args.config)
    t = time.perf_counter()
    problem = sg.generate_ba(spec)
    n_obs = problem.num_obs
    label = f"{spec.num_cameras}-camera/{spec.num_points}-point"
    log(f"[bench] generated {spec.num_cameras} cams, {spec.num_points} pts, {n_obs} obs "
        f"in {time.perf_counter() - t:.1f}s")

    # ---- per-iteration throughput of the plain FP64 path -----------------------------------
    params = sg.OptParams(num_iterations=1 << 30, initial_lm_damping=1.0, function_tolerance=0.0,
                          gradient_tolerance=0.0, parameter_tolerance=0.0,
                          max_consecutive_rejects=1 << 30)
    t = time.perf_counter()
    opt = sg.Optimizer(problem.variables(), problem.cost_fns(1.0), params)
    st = opt.rcs_stats()
    log(f"[bench] optimizer created (upload + symbolic analysis, PCIe-inclusive) in "
        f"{time.perf_counter() - t:.2f}s: {st}")
    for _ in range(args.warmup):
        info = opt.step()
        log(f"[bench] warmup merit {info.merit:.6e} lin {info.ms_linearize:.2f} fac {info.ms_factor:.2f} "
            f"sol {info.ms_solve:.2f} cost {info.ms_cost:.2f} total {info.ms_total:.2f} ms")
    opt.set_kernel_timing(1)
    opt.reset_kernel_times()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        info = opt.step()
    t1 = time.perf_counter()
    step_s = (t1 - t0) / args.steps
    value = n_obs / step_s
    ktimes = opt.kernel_times()
    log(f"[bench] timed merit {info.merit:.6e}; {step_s * 1e3:.2f} ms/step, {value:.3e} obs/s; {ktimes}")
    # per-launch timing of the dominant kernel on two further iterations (not in `value`)
    opt.set_kernel_timing(2)
    opt.reset_kernel_times()
    for _ in range(2):
        opt.step()
    kt2 = opt.kernel_times()
    opt.close()
    n_launch, ms_launch = kt2["mf_step"]
    avg_s = ms_launch / n_launch / 1e3
    bytes_launch = st["mf_step_bytes"] / st["mf_step_launches"]
    flops_launch = st["mf_step_flops"] / st["mf_step_launches"]
    achieved = bytes_launch / avg_s / 1e9
    # HBM bytes per round from the committed rocprofv3 --pmc summary of this workload
    # (FETCH_SIZE and WRITE_SIZE runs, scaled by a 256 MiB copy's readings)
    traffic = None
    tfile = ROOT / "profiles" / "pmc_traffic_ba_walk_2000cams.json"
    if args.config == 2 and tfile.exists():
        kern = json.loads(tfile.read_text()).get("round")
        traffic = kern["hbm_bytes_per_round"] if kern else None
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": "round of panel steps (k_mf_step; k_mf_panel + k_mf_update when split)",
            "bytes_per_launch": bytes_launch, "flops_per_launch": flops_launch,
            "tflops": flops_launch / avg_s / 1e12, "avg_launch_us": avg_s * 1e6,
            "launches_per_factorization": st["mf_step_launches"]}
    log(f"[bench] roofline {roof}")
    kernels = {k: {"launches": c, "avg_ms": (ms / c if c else None)} for k, (c, ms) in ktimes.items()
               if c}
    kernels["mf_step"] = {"launches": n_launch, "avg_ms": ms_launch / n_launch}

    # ---- time to tolerance per path -----------------------------------------------------------
    tol = None
    if not args.no_tol:
        wpath = ROOT / "sophus-rs_amd" / "precond_weights.f32"
        weights = np.fromfile(wpath, dtype="<f4") if wpath.exists() else None
        paths = {
            "f64_ldlt": dict(),
            "f64_pcg_jacobi": dict(solver=sg.Solver.SchurPcgJacobi, pcg_relative_tolerance=1e-8,
                                   pcg_max_iterations=3000),
            "f64_pcg_learned": dict(solver=sg.Solver.SchurPcgLearned, pcg_relative_tolerance=1e-8,
                                    pcg_max_iterations=3000, precond_weights=weights),
            "mixed_ldlt": dict(precision=sg.Precision.Mixed),
        }
        if weights is None:
            del paths["f64_pcg_learned"]
        hists = {}
        for name, kw in paths.items():
            hists[name] = run_path(sg, problem, **kw)
            log(f"[bench] {name}: {len(hists[name])} its, {hists[name][-1][0]:.2f}s, "
                f"final {hists[name][-1][1]:.9e}")
        f_star = hists["f64_ldlt"][-1][1]
        tol = {"f_star": f_star, "rtol": TOL_REL,
               "paths": {k: tol_summary(h, f_star, n_obs) for k, h in hists.items()}}
        base = tol["paths"]["f64_ldlt"]
        for k, v in tol["paths"].items():
            v["speedup_vs_f64_ldlt"] = (base["seconds"] / v["seconds"]) if v["reached"] else None
        tol["note"] = ("FP32 S (mixed) with PCG is not run: FP32 rounding makes the reduced camera "
                       "system indefinite in its near-gauge directions and CG breaks down. The "
                       "endpoint after 100 iterations depends on rounding (the reference's "
                       "Huber weighting makes LM crawl with nu doubling), so reaching F* from "
                       "another path is not a property of its speed alone.")
        log(f"[bench] time-to-tol {json.dumps(tol)}")

    # ---- batched small windows: problems/s and per-window latency -----------------------------
    batch = None
    if not args.no_batch:
        batch = batch_leg(sg, args.batch_windows)
        log(f"[bench] batch {json.dumps(batch)}")

    cpu = None
    if not args.no_cpu:
        t = time.perf_counter()
        cpu = cpu_baseline(problem, label)
        log(f"[bench] cpu baseline {cpu} ({time.perf_counter() - t:.1f}s)")

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "obs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"ba_walk_{spec.num_cameras}cams_{spec.num_points}pts",
                   "num_cameras": spec.num_cameras, "num_points": spec.num_points, "num_obs": n_obs,
                   "value_path": "f64 Schur + direct block LDLt, per LM iteration",
                   "huber_delta_px": 1.0, "initial_lm_damping": 1.0, "seed": spec.seed},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels": kernels,
        "time_to_tol": tol,
        "batch": batch,
        "rcs": st,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
