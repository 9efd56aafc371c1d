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

Profile summaries (text, for profiles/) from rocprofv3 CSV output:
  python bench.py --summarize-trace DIR OUT   (--kernel-trace --stats run)
  python bench.py --summarize-pmc FETCH_DIR WRITE_DIR OUT
                                              (--pmc FETCH_SIZE / WRITE_SIZE runs)
The PMC summary scales each counter by what it reads for the step's plain
copy kernel (k_copy: n doubles read and written) and records the HBM bytes
per dispatch of each kernel; the bench line's roofline.traffic comes from it.
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


# timer group -> the kernel whose dispatches it times
GROUP_KERNEL = {"stream_power": "k_tile_sp", "upstream_counts": "k_tile_counts", "tile_depths": "k_tile_order",
                "hillslope_residual": "k_trial", "hillslope_prepare": "k_prepare", "receivers": "k_receivers"}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "lem_step_4096_hbm_traffic.txt")


def _short(name):
    n = name.replace("lsdtt_gpu::(anonymous namespace)::", "").replace("lsdtt_gpu::", "")
    n = n.replace("void ", "")
    return n.split("(")[0].split("<")[0]


def _csv(path):
    import csv
    with open(path) as f:
        return list(csv.DictReader(f))


def _find(d, suffix):
    for f in sorted(os.listdir(d)):
        if f.endswith(suffix):
            return os.path.join(d, f)
    raise FileNotFoundError(f"no *{suffix} in {d}")


def summarize_trace(d, out):
    rows = _csv(_find(d, "kernel_stats.csv"))
    lines = ["# rocprofv3 kernel statistics, LEM step at 4096^2 (bench.py workload)", "#",
             "# made by: cd /tmp && rocprofv3 --kernel-trace --stats --output-format csv -d <dir> -o run -- "
             "python3 bench.py --steps 3 --warmup 1 --no-cpu-baseline",
             "# summarized by: python bench.py --summarize-trace <dir> <this file>",
             "# totals cover the 4 LEM steps of the run (1 warm-up + 3 timed) plus the context set-up", "#",
             f"{'kernel':34s} {'calls':>6s} {'total_ms':>10s} {'avg_us':>10s} {'min_us':>9s} {'max_us':>9s} {'pct':>6s}"]
    agg = {}
    for r in rows:
        k = _short(r["Name"])
        a = agg.setdefault(k, [0, 0.0, float("inf"), 0.0])
        a[0] += int(r["Calls"])
        a[1] += float(r["TotalDurationNs"])
        a[2] = min(a[2], float(r["MinNs"]))
        a[3] = max(a[3], float(r["MaxNs"]))
    tot = sum(a[1] for a in agg.values())
    for k, (c, t, mn, mx) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        lines.append(f"{k[:34]:34s} {c:6d} {t / 1e6:10.3f} {t / c / 1e3:10.2f} {mn / 1e3:9.2f} {mx / 1e3:9.2f} "
                     f"{100 * t / tot:6.2f}")
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")


def summarize_pmc(fetch_dir, write_dir, out, n_cells=4096 * 4096):
    per = {}
    for key, d in (("fetch", fetch_dir), ("write", write_dir)):
        for r in _csv(_find(d, "counter_collection.csv")):
            k = _short(r["Kernel_Name"])
            a = per.setdefault(k, {"fetch": [0, 0.0], "write": [0, 0.0]})[key]
            a[0] += 1
            a[1] += float(r["Counter_Value"])
    copy_kib = 8.0 * n_cells / 1024.0
    cf = per["k_copy"]["fetch"]
    cw = per["k_copy"]["write"]
    ff = copy_kib / (cf[1] / cf[0])
    fw = copy_kib / (cw[1] / cw[0])
    lines = ["# HBM traffic per dispatch from rocprofv3 PMC counters, LEM step at 4096^2 (bench.py workload)", "#",
             "# made by: cd /tmp && rocprofv3 --pmc FETCH_SIZE --output-format csv -d <fetch_dir> -o run -- "
             "python3 bench.py --steps 2 --warmup 1 --no-cpu-baseline",
             "#          cd /tmp && rocprofv3 --pmc WRITE_SIZE --output-format csv -d <write_dir> -o run -- "
             "python3 bench.py --steps 2 --warmup 1 --no-cpu-baseline",
             "# summarized by: python bench.py --summarize-pmc <fetch_dir> <write_dir> <this file>",
             f"# calibration: k_copy reads and writes {copy_kib:.0f} KiB per dispatch; FETCH_SIZE reads "
             f"{cf[1] / cf[0]:.1f}, WRITE_SIZE {cw[1] / cw[0]:.1f} -> scale factors fetch {ff:.4f}, write {fw:.4f}",
             "#",
             f"{'kernel':34s} {'disp':>5s} {'fetch_KiB':>12s} {'write_KiB':>12s} {'HBM_MB/disp':>12s}"]
    for k, v in sorted(per.items(), key=lambda kv: -(kv[1]["fetch"][1] * ff + kv[1]["write"][1] * fw)):
        nf, sf = v["fetch"]
        nw, sw = v["write"]
        mf = sf / nf if nf else 0.0
        mw = sw / nw if nw else 0.0
        mb = (mf * ff + mw * fw) * 1024 / 1e6
        lines.append(f"{k[:34]:34s} {max(nf, nw):5d} {mf:12.1f} {mw:12.1f} {mb:12.2f}")
    lines.append("#")
    for k, v in sorted(per.items()):
        nf, sf = v["fetch"]
        nw, sw = v["write"]
        if nf and nw:
            b = (sf / nf * ff + sw / nw * fw) * 1024
            lines.append(f"traffic_bytes_per_dispatch {k} {b:.0f}")
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")


def traffic_for(group, size):
    """HBM bytes per dispatch of the group's kernel from the committed PMC
    summary (same workload), or None."""
    k = GROUP_KERNEL.get(group)
    if k is None or size != 4096 or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        for line in f:
            p = line.split()
            if len(p) == 3 and p[0] == "traffic_bytes_per_dispatch" and p[1] == k:
                return float(p[2])
    return None


def ensemble_member(args):
    """one instance of the 64-member ensemble: seeded size, surface and
    parameters (K log-uniform 1e-6..1e-4, m/n uniform 0.3..0.6 with n = 1,
    D log-uniform 1e-3..1e-1 (Roering, Sc = 0.8), U log-uniform 1e-4..1e-3)"""
    i, seed = args
    from lsdtt3_amd.synthetic import lem_surface
    rng = np.random.default_rng(seed * 1000 + i)
    cells = float(np.exp(rng.uniform(np.log(256.0 ** 2), np.log(2048.0 ** 2))))
    aspect = float(rng.uniform(0.5, 2.0))
    nr = max(32, int(round(np.sqrt(cells * aspect))))
    nc = max(32, int(round(cells / nr)))
    logu = lambda a, b: float(np.exp(rng.uniform(np.log(a), np.log(b))))  # noqa: E731
    K, theta, D, U = logu(1e-6, 1e-4), float(rng.uniform(0.3, 0.6)), logu(1e-3, 1e-1), logu(1e-4, 1e-3)
    z = lem_surface(nr, nc, dx=30.0, seed=seed * 1000 + i, max_slope=0.75 * 0.8)
    return dict(dem=z, cell_size=30.0, timestep_years=500.0, erodibility=K, m_exponent=theta, n_exponent=1.0,
                uplift_rate=U, hillslope_diffusivity=D, nonlinear_hillslopes=True, critical_hillslope_slope=0.8,
                boundary_conditions="nnon", newton_rel_tol=1e-8)


def bench_ensemble(args):
    import lsdtt3_amd
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(max_workers=8) as ex:
        ms = list(ex.map(ensemble_member, [(i, 2024) for i in range(args.instances)]))
    cells = [m["dem"].size for m in ms]
    ens = lsdtt3_amd.Ensemble(ms, workers=args.workers)
    ens.run(args.warmup)
    stats, sec = ens.run(args.steps)
    total = sum(ens.n_valid)
    value = total * args.steps / sec
    worst = max((s["hillslope_residual"] / s["hillslope_initial_residual"]) if s["hillslope_initial_residual"] > 0
                else 0.0 for s in stats)
    out = {"metric": METRIC, "value": round(value, 1), "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(1e3 * sec / args.steps, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"batched ensemble of {len(ms)} LEM instances, cells log-uniform "
                                  f"256^2..2048^2 (total {total}), per-instance K, m/n, D, U; Roering Newton "
                                  f"r<=1e-8, stream power n=1; {args.workers} host workers, one stream each",
                      "instances": len(ms), "total_cells": total, "min_cells": min(cells), "max_cells": max(cells),
                      "worst_relative_residual": worst,
                      "picard_fallbacks": sum(1 for s in stats if s["hillslope_method"] == 1)}}
    if args.sequential:
        seq = lsdtt3_amd.Ensemble(ms, workers=1)
        seq.run(args.warmup)
        _, s1 = seq.run(args.steps)
        out["sequential_one_stream"] = {"value": round(total * args.steps / s1, 1), "unit": "cell-steps/s",
                                        "seconds": round(s1, 3)}
    print(json.dumps(out))


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--summarize-trace":
        summarize_trace(sys.argv[2], sys.argv[3])
        return
    if len(sys.argv) > 1 and sys.argv[1] == "--summarize-pmc":
        summarize_pmc(sys.argv[2], sys.argv[3], sys.argv[4])
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--workload", choices=["lem4096", "ensemble"], default="lem4096")
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--sequential", action="store_true", help="ensemble: also time one instance at a time")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: this build has no multi-GPU path (asked for {args.gpus}); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import lsdtt3_amd
    if not lsdtt3_amd.gpu_available():
        print("no GPU visible to liblsdtt3_gpu.so", file=sys.stderr)
        sys.exit(1)
    if args.workload == "ensemble":
        bench_ensemble(args)
        return
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
    roofline = {"bound": "hbm", "kernel": dname, "hip_kernel": GROUP_KERNEL.get(dname),
                "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_for(dname, args.size),
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
