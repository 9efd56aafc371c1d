"""LEM timestep throughput on one MI355X.

Workload (the plain-GPU LEM configuration the metric is quoted on):
4096 x 4096 cells at dx = 30 m, FP64 elevation state; a tilted plane (0.01)
plus seeded power-law noise (beta = 2, 50 m std) plus 200 pits, smoothed
until the steepest centred slope is <= 0.75 Sc so that the Roering
hillslope starts sub-critical; Roering hillslope D = 0.005, Sc = 0.8 by
Newton (FP64 residual, FP32 BiCGSTAB) to a 1e-8 relative residual; stream
power K log-uniform in [1e-6, 1e-4] (seeded), m = 0.5, n = 1.5; U = 5e-4
m/yr, dt = 500 yr; boundary "nnon" (south edge fixed).

One step = one run_lem_step on the resident state.  value = valid cells x
timed steps / wall seconds, inputs resident in HBM before the clock starts.
The roofline object is for the kernel group with the largest share of the
timed region, measured with HIP events on the step's stream; the CPU
baseline is the oracle's single-threaded FP64 step on a 512 x 512 crop of
the same recipe.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size N]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "LEM cell-steps/s at 4096², learned+batched vs plain GPU at 1e-8 residual"


def workload(size, seed=2024):
    from lsdtt3_amd.synthetic import lem_surface
    rng = np.random.default_rng(seed)
    K = float(10.0 ** rng.uniform(-6.0, -4.0))
    z = lem_surface(size, size, dx=30.0, seed=seed, max_slope=0.75 * 0.8)
    return z, K, lem_surface.last_passes


def lem_kwargs(K):
    return dict(timestep_years=500.0, erodibility=K, m_exponent=0.5, n_exponent=1.5, uplift_rate=5e-4,
                hillslope_diffusivity=0.005, nonlinear_hillslopes=True, critical_hillslope_slope=0.8,
                boundary_conditions="nnon", newton_rel_tol=1e-8)


def cpu_baseline(K, budget_s=15.0, size=512):
    """oracle FP64 LEM step (single thread) on a crop of the same recipe"""
    from oracle import oracle as orc
    z, _, _ = workload(size)
    mask = np.ones(z.shape, np.uint8)
    p = orc.lem_params(500.0, K, 0.5, 1.5, 5e-4, 0.005, Sc=0.8, nonlinear=True, bc="nnon", rel_tol=1e-8)
    bf = orc.boundary_flats(mask, "nnon")
    steps, t0 = 0, time.perf_counter()
    while True:
        z, st, rc = orc.lem_step(z, mask, 30.0, p, bf)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 200:
            break
    return {"value": size * size * steps / el, "unit": "cell-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle FP64 LEM step (orc_lem_step, GMRES Newton) on a {size}x{size} crop of the "
                      f"same recipe, {steps} steps in {el:.1f} s, 1 thread"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: this build has no multi-GPU path (asked for {args.gpus}); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import lsdtt3_amd
    if not lsdtt3_amd.gpu_available():
        print("no GPU visible to liblsdtt3_gpu.so", file=sys.stderr)
        sys.exit(1)
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is plumbing only
        sync = lambda: None  # noqa: E731

    z, K, passes = workload(args.size)
    lem = lsdtt3_amd.LandscapeModel(z, 30.0, **lem_kwargs(K))
    for _ in range(args.warmup):
        lem.step()
    sync()
    lem.profile(True)
    newton = krylov = 0
    worst_r = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = lem.step()
        newton += st["newton_iterations"]
        krylov += st["krylov_iterations"]
        if st["hillslope_initial_residual"] > 0:
            worst_r = max(worst_r, st["hillslope_residual"] / st["hillslope_initial_residual"])
        if not st["hillslope_converged"]:
            raise RuntimeError("hillslope Newton did not converge")
    sync()
    dt = time.perf_counter() - t0
    prof = lem.profile_totals()
    lem.profile(False)

    n_valid = int(st["n_valid"])
    value = n_valid * args.steps / dt
    total_ms = sum(v[0] for v in prof.values())
    dom = max(prof.items(), key=lambda kv: kv[1][0])
    dname, (dms, dlaunch, dbytes) = dom
    avg_s = dms / 1e3 / max(dlaunch, 1)
    achieved = dbytes / max(dlaunch, 1) / avg_s / 1e9 if avg_s > 0 else 0.0
    roofline = {"bound": "hbm", "kernel": dname, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "avg_launch_us": round(avg_s * 1e6, 2), "launches": dlaunch,
                "share_of_timed_kernels": round(dms / total_ms, 3) if total_ms else None}
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "cell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * dt / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"plain GPU LEM {args.size}x{args.size} dx=30 FP64 state, Roering "
                               f"Newton (D=0.005 Sc=0.8, r<=1e-8) + stream power n=1.5 m=0.5 "
                               f"K={K:.3e}, U=5e-4 dt=500 bc=nnon",
                   "n_valid": n_valid, "smoothing_passes": passes,
                   "newton_iters_per_step": round(newton / args.steps, 2),
                   "krylov_iters_per_step": round(krylov / args.steps, 2),
                   "worst_relative_residual": worst_r,
                   "count_passes_last_step": int(st["count_passes"]),
                   "stream_power_passes_last_step": int(st["stream_power_passes"]),
                   "count_tile_runs_last_step": int(st["count_tile_runs"]),
                   "stream_power_tile_runs_last_step": int(st["stream_power_tile_runs"])},
        "roofline": roofline,
        "kernels_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in sorted(prof.items(), key=lambda kv: -kv[1][0])},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(K, args.cpu_budget)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
