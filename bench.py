"""LEM timestep throughput on one MI355X.

Workload (the plain-GPU LEM configuration the metric is quoted on):
4096 x 4096 cells at dx = 30 m, FP64 elevation state; a tilted plane (0.01)
plus seeded power-law noise (beta = 2, 50 m std) plus 200 pits, smoothed
until the steepest centred slope is <= 0.75 Sc so that the Roering
hillslope starts sub-critical; Roering hillslope D = 0.005, Sc = 0.8 by
Newton (FP64 residual, FP32 BiCGSTAB) to a 1e-8 relative residual; stream
power K log-uniform in [1e-6, 1e-4] (seeded), m = 0.5, n = 1.5; U = 5e-4
m/yr, dt = 500 yr; boundary "nnon" (south edge fixed).

One step = one run_lem_step on the resident state, 500 steps.  value =
valid cells x timed steps / wall seconds, inputs resident in HBM before the clock starts,
with no instrumentation in the stream.  The per-kernel breakdown and the
roofline object (the kernel group with the largest share) come from a
separate run of --profile-steps steps timed with HIP events recorded on
the step's stream (the events add gaps of their own, so that run is not the
timed one); the CPU
baseline is the oracle's single-threaded FP64 step on a 512 x 512 crop of
the same recipe.

The line also carries learned_warm_start: the same step with the Newton
solve started from the explicit-Euler prediction and from the U-Net's
prediction (trained here first, on tile pairs captured from seeded 2048^2
runs of the same physics), each to the same r <= 1e-8, with the network's
training figures and its MFMA utilisation at the 4096^2 inference shape
(256 windows of 288^2).  value stays the plain step, the fastest of the
three.  The line also carries batched: the 64-instance ensemble (cells
log-uniform 256^2..2048^2, per-instance K, m/n, D, U) on the batched engine
-- one launch per kernel for all instances, the step replayed from a
captured hipGraph -- with its throughput relative to the plain step, and
the same ensemble run again with each instance's Newton solve started from
its previous step's hillslope change (previous_step_start).
The relative residual r reported is ||F||inf over ||F(z^n)||inf of the
sub-step the Newton stop judged (the stop is r <= 1e-8).
The timed steps run from the initial surface (warm-up and tile autotuning
steps are discarded by resetting the elevation), so the default 500 timed
steps are the workload's whole trajectory, reported per 100-step segment
as well (Newton and Krylov counts grow as relief builds).  unsmoothed is the plain CPU
workload's recipe at the same size: the raw surface (steepest centred slope
about 3, far above Sc, which is why the Roering workload smooths it first)
with linear diffusion D = 0.01, K = 1e-5, n = 1, U = 1e-3, dt = 1000 yr.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size N] [--no-learned]
       python bench.py --workload ensemble   (64-instance batched ensemble)
       python bench.py --workload cnn        (the U-Net alone: MFMA utilisation)
       python bench.py --workload routing    (fill + D8 flow object + chi on a 4096^2 FP32 DEM)
       python bench.py --workload plain512   (the plain CPU workload: 512^2 raw surface, linear,
                                              200 steps, GPU and OpenMP oracle)

Profile summaries (text, for profiles/) from rocprofv3 CSV output:
  python bench.py --summarize-trace DIR OUT [TITLE CMD NOTE]   (--kernel-trace --stats run)
  python bench.py --summarize-pmc FETCH_DIR WRITE_DIR OUT [TITLE CMD]
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


def progress(msg):
    """One line per finished stage on stderr (the JSON line stays alone on stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)

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


# variant of the step (the metric's workload is the default: D8 incision,
# no channel threshold); set from the command line
STEP_VARIANT = {"fluvial_method": "d8", "incision_threshold_specific_area": 0.0}


def lem_kwargs(K):
    return dict(timestep_years=500.0, erodibility=K, m_exponent=0.5, n_exponent=1.5, uplift_rate=5e-4,
                hillslope_diffusivity=0.005, nonlinear_hillslopes=True, critical_hillslope_slope=0.8,
                boundary_conditions="nnon", newton_rel_tol=1e-8, **STEP_VARIANT)


def host_threads():
    """threads the CPU baseline uses: the cores in this process's affinity
    mask, capped by OMP_NUM_THREADS when that is set"""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n


# the plain CPU workload's physics: linear diffusion and n = 1 incision on
# the raw (unsmoothed) surface
PLAIN_CPU_KW = dict(timestep_years=1000.0, erodibility=1e-5, m_exponent=0.5, n_exponent=1.0, uplift_rate=1e-3,
                    hillslope_diffusivity=0.01, boundary_conditions="nnon", newton_rel_tol=1e-8)


def unsmoothed_measure(args, sync, steps=50):
    """the plain CPU workload's recipe scaled to size^2: raw surface (no
    smoothing: steepest centred slope ~3), linear D = 0.01, K = 1e-5, n = 1,
    U = 1e-3, dt = 1000 yr"""
    import lsdtt3_amd
    from lsdtt3_amd.synthetic import lem_surface, max_centred_slope
    z = lem_surface(args.size, args.size, dx=30.0, seed=2024)
    lem = lsdtt3_amd.LandscapeModel(z, 30.0, **PLAIN_CPU_KW)
    for _ in range(3):
        lem.step()
    dt, a, st, _ = run_steps(lem, steps, sync)
    nv = int(st["n_valid"])
    return {"value": round(nv * steps / dt, 1), "unit": "cell-steps/s", "steps": steps,
            "ms_per_step": round(1e3 * dt / steps, 3),
            "workload": f"raw {args.size}x{args.size} surface (max centred slope "
                        f"{max_centred_slope(z, 30.0):.2f}), linear D=0.01, K=1e-5 m=0.5 n=1, U=1e-3, "
                        f"dt=1000, bc=nnon, r<=1e-8",
            "newton_iters_per_step": round(a["newton"] / steps, 2),
            "krylov_iters_per_step": round(a["krylov"] / steps, 2),
            "worst_relative_residual": a["worst_r"]}


def bench_plain_cpu_workload(args):
    """The plain CPU workload as written (512^2 raw surface, linear D = 0.01,
    K = 1e-5, n = 1, U = 1e-3, dt = 1000 yr, 200 steps) on the GPU, with the
    oracle's OpenMP FP64 run of the same 200 steps beside it and the largest
    elevation difference between the two end states."""
    import lsdtt3_amd
    from lsdtt3_amd import _native as N
    from lsdtt3_amd.synthetic import lem_surface
    from oracle import oracle as orc
    n, steps = 512, 200
    z0 = lem_surface(n, n, dx=30.0, seed=2024)
    lem = lsdtt3_amd.LandscapeModel(z0, 30.0, **PLAIN_CPU_KW)
    lem.step()                      # warm the library; the timed run restarts from z0
    lem.elevation = z0
    dt, a, st, _ = run_steps(lem, steps, N.synchronize)
    zg = lem.elevation
    out = {"metric": "LEM cell-steps/s, plain CPU workload (512^2, 200 steps)",
           "value": round(n * n * steps / dt, 1), "unit": "cell-steps/s", "n_gpus": 1, "steps": steps,
           "warmup": 0, "ms_per_step": round(1e3 * dt / steps, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "raw 512x512 surface dx=30 (0.01 tilt, beta 2, 50 m std, 200 pits), linear "
                                  "D=0.01, K=1e-5 m=0.5 n=1, U=1e-3, dt=1000, bc=nnon, r<=1e-8",
                      "newton_iters_per_step": round(a["newton"] / steps, 2),
                      "krylov_iters_per_step": round(a["krylov"] / steps, 2)}}
    if not args.no_cpu_baseline:
        threads = orc.set_threads(host_threads())
        mask = np.ones(z0.shape, np.uint8)
        p = orc.lem_params(dt=1000.0, K=1e-5, m=0.5, n=1.0, U=1e-3, D=0.01, bc="nnon", rel_tol=1e-8)
        bf = orc.boundary_flats(mask, "nnon")
        zo = z0.copy()
        t0 = time.perf_counter()
        for _ in range(steps):
            zo, so, rc = orc.lem_step(zo, mask, 30.0, p, bf)
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(n * n * steps / el, 1), "unit": "cell-steps/s", "cores": threads,
                               "kind": "port", "sample": f"the whole workload (200 steps) by the oracle, "
                                                         f"{threads} OpenMP threads, {el:.1f} s"}
        out["config"]["max_abs_dz_vs_oracle_after_200_steps"] = float(np.abs(zg - zo).max())
    print(json.dumps(out))


def cpu_baseline(K, budget_s=15.0, size=512):
    """oracle FP64 LEM step, OpenMP over the host cores, on a crop of the same
    recipe (results do not depend on the thread count)"""
    from oracle import oracle as orc
    threads = orc.set_threads(host_threads())
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
    cpu = ""
    try:
        with open("/proc/cpuinfo") as f:
            cpu = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": size * size * steps / el, "unit": "cell-steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle FP64 LEM step (orc_lem_step: Newton with OpenMP GMRES, serial fill/routing/"
                      f"stream power as in the reference) on a {size}x{size} crop of the same recipe, "
                      f"{steps} steps in {el:.1f} s, {threads} OpenMP threads"
                      + (f" on {cpu}" if cpu else "")}


# timer group -> the kernel whose dispatches it times
GROUP_KERNEL = {"stream_power": "k_tile_sp", "upstream_counts": "k_tile_counts", "tile_depths": "k_tile_order",
                "hillslope_residual": "k_trial", "hillslope_prepare": "k_prepare", "receivers": "k_receivers",
                "bicgstab_pv": "k_bi_pv", "bicgstab_st": "k_bi_st", "bicgstab_x": "k_bi_x",
                "step_stats": "k_stats"}
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


def summarize_trace(d, out, title=None, cmd=None, note=None):
    rows = _csv(_find(d, "kernel_stats.csv"))
    title = title or "LEM step at 4096^2 (bench.py workload)"
    cmd = cmd or "python3 bench.py --steps 3 --warmup 1 --no-cpu-baseline --no-learned --autotune-steps 0"
    note = note or "totals cover the 4 LEM steps of the run (1 warm-up + 3 timed) plus the context set-up"
    lines = [f"# rocprofv3 kernel statistics, {title}", "#",
             "# made by: cd /tmp && rocprofv3 --kernel-trace --stats --output-format csv -d <dir> -o run -- " + cmd,
             "# summarized by: python bench.py --summarize-trace <dir> <this file>"
             + ("" if "--no-learned" in cmd and title.startswith("LEM") else " <title> <command> <note>"),
             f"# {note}", "#",
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


PMC_CMD = "python3 bench.py --steps 2 --warmup 1 --no-cpu-baseline --no-learned --no-batched --no-unsmoothed " \
          "--autotune-steps 0 --profile-steps 1"


def summarize_pmc(fetch_dir, write_dir, out, title=None, cmd=None, n_cells=4096 * 4096):
    per = {}
    for key, d in (("fetch", fetch_dir), ("write", write_dir)):
        for r in _csv(_find(d, "counter_collection.csv")):
            k = _short(r["Kernel_Name"])
            a = per.setdefault(k, {"fetch": [0, 0.0], "write": [0, 0.0]})[key]
            a[0] += 1
            a[1] += float(r["Counter_Value"])
    copy_kib = 8.0 * n_cells / 1024.0
    if "k_copy" in per and per["k_copy"]["fetch"][0] and per["k_copy"]["write"][0]:
        cf = per["k_copy"]["fetch"]
        cw = per["k_copy"]["write"]
        ff = copy_kib / (cf[1] / cf[0])
        fw = copy_kib / (cw[1] / cw[0])
        calib = (f"# calibration: k_copy reads and writes {copy_kib:.0f} KiB per dispatch; FETCH_SIZE reads "
                 f"{cf[1] / cf[0]:.1f}, WRITE_SIZE {cw[1] / cw[0]:.1f} -> scale factors fetch {ff:.4f}, "
                 f"write {fw:.4f}")
    else:
        # the D8 step copies the surface inside the receivers kernel, so no
        # plain copy runs: keep the factors of the last calibration
        import re
        old = open(TRAFFIC_FILE).read()
        m = re.search(r"scale factors fetch ([0-9.]+), write ([0-9.]+)", old)
        if not m:
            raise SystemExit("no k_copy dispatch to calibrate against and no earlier calibration")
        ff, fw = float(m.group(1)), float(m.group(2))
        calib = (f"# calibration: no plain copy in this run; scale factors fetch {ff:.4f}, write {fw:.4f} "
                 f"carried over from the previous calibration against k_copy")
    title = title or "LEM step at 4096^2 (bench.py workload)"
    cmd = cmd or PMC_CMD
    lines = [f"# HBM traffic per dispatch from rocprofv3 PMC counters, {title}", "#",
             "# made by: cd /tmp && rocprofv3 --pmc FETCH_SIZE --output-format csv -d <fetch_dir> -o run -- " + cmd,
             "#          cd /tmp && rocprofv3 --pmc WRITE_SIZE --output-format csv -d <write_dir> -o run -- " + cmd,
             "# summarized by: python bench.py --summarize-pmc <fetch_dir> <write_dir> <this file>"
             + ("" if cmd == PMC_CMD else " <title> <command>"),
             calib,
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


def ensemble_measure(args):
    """the 64-instance ensemble on the batched engine, tiles autotuned, graph
    replay, --ensemble-steps timed steps from the plain start z^n and again
    with each instance's previous hillslope change as the start (same
    surfaces, same tile height); returns the measurement dict"""
    import lsdtt3_amd
    from lsdtt3_amd.learned import WARM_PREVIOUS
    from concurrent.futures import ProcessPoolExecutor
    t0 = time.perf_counter()
    with ProcessPoolExecutor(max_workers=8) as ex:
        ms = list(ex.map(ensemble_member, [(i, 2024) for i in range(args.instances)]))
    gen_s = time.perf_counter() - t0
    cells = [m["dem"].size for m in ms]

    def timed(ens, tune):
        ens.run(max(1, args.warmup))
        tuning = None
        if tune and args.autotune_steps > 0:
            th, tms = ens.autotune_tiles(args.autotune_steps)
            tuning = {"tile_height": th, "ms_per_step": {str(h): round(v, 3) for h, v in tms.items()},
                      "steps_per_trial": args.autotune_steps}
        ens.run(1)  # the pattern of the chosen height (tuning restores the surfaces)
        stats, sec = ens.run(args.ensemble_steps)
        # relative residual as the Newton stop measures it: ||F|| over ||F(z^n)|| of
        # the sub-step start (tolerance = rel_tol * ||F(z^n)||)
        worst = max(s["hillslope_residual"] / max(s["hillslope_tolerance"] / 1e-8, 1e-300) for s in stats)
        return stats, sec, worst, tuning

    ens = lsdtt3_amd.Ensemble(ms)
    total = sum(ens.n_valid)
    stats, sec, worst, tuning = timed(ens, True)
    value = total * args.ensemble_steps / sec
    pat = ens.pattern()
    out = {"value": round(value, 1), "unit": "cell-steps/s", "ms_per_step": round(1e3 * sec / args.ensemble_steps, 3),
           "steps": args.ensemble_steps,
           "workload": f"batched ensemble of {len(ms)} LEM instances, cells log-uniform 256^2..2048^2 "
                       f"(total {total}), per-instance K, m/n, D, U; Roering Newton r<=1e-8, stream power n=1; "
                       f"one launch per kernel for all instances, step replayed from a captured hipGraph",
           "instances": len(ms), "total_cells": total, "min_cells": min(cells), "max_cells": max(cells),
           "worst_relative_residual": worst,
           "newton_iters_per_step_max": max(s["newton_iterations"] for s in stats),
           "krylov_iters_per_step_max": max(s["krylov_iterations"] for s in stats),
           "picard_fallbacks": sum(1 for s in stats if s["hillslope_method"] == 1),
           "launch_pattern": pat, "autotuned_tiles": tuning, "generate_s": round(gen_s, 2)}
    if args.no_ensemble_warm_start:
        return out
    # the same run from each instance's previous hillslope change
    plain = [ens.elevation(i) for i in range(len(ms))]
    th = ens.tile_height
    del ens
    ew = lsdtt3_amd.Ensemble(ms)
    ew.tile_height = th
    ew.set_warm_start(WARM_PREVIOUS)
    wstats, wsec, wworst, _ = timed(ew, False)
    # the two trajectories agree step by step to the Newton tolerance; over
    # hundreds of steps a tolerance-sized difference can flip the D8 receiver
    # of a near-tied cell and move its channel, as any other valid solve of
    # the same tolerance does (another Krylov forcing, another thread count)
    diffs = [np.abs(ew.elevation(i) - plain[i]) for i in range(len(ms))]
    diff = max(float(d.max()) for d in diffs)
    moved = sum(int((d > 1e-3).sum()) for d in diffs)
    moved_inst = sum(1 for d in diffs if d.max() > 1e-3)
    del diffs
    wvalue = total * args.ensemble_steps / wsec
    out["previous_step_start"] = {
        "value": round(wvalue, 1), "ms_per_step": round(1e3 * wsec / args.ensemble_steps, 3),
        "vs_plain_start": round(wvalue / value, 3), "worst_relative_residual": wworst,
        "accepted_last_step": sum(s["warm_start_accepted"] for s in wstats),
        "newton_iters_per_step_max": max(s["newton_iterations"] for s in wstats),
        "krylov_iters_per_step_max": max(s["krylov_iterations"] for s in wstats),
        "picard_fallbacks": sum(1 for s in wstats if s["hillslope_method"] == 1),
        "max_abs_diff_vs_plain_m": diff, "cells_over_1mm_vs_plain": moved, "instances_over_1mm_vs_plain": moved_inst,
        "launch_pattern": ew.pattern()}
    return out


def bench_ensemble(args):
    m = ensemble_measure(args)
    out = {"metric": METRIC, "value": m["value"], "unit": "cell-steps/s", "n_gpus": 1, "steps": args.ensemble_steps,
           "warmup": args.warmup, "ms_per_step": m["ms_per_step"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {k: v for k, v in m.items() if k not in ("value", "unit", "ms_per_step", "steps")}}
    print(json.dumps(out))


BF16_PEAK_TFLOPS = 2500.0


def run_steps(lem, n, sync, profile=False):
    """n steps; wall seconds (inputs resident, synchronised on both sides),
    Newton / Krylov totals, worst relative residual, warm-start figures and,
    with profile, the per-kernel-group HIP-event totals.  The event records
    between kernel groups add their own gaps to the stream, so the timed
    throughput comes from a run without them and the breakdown from a
    separate profiled run."""
    if profile:
        lem.profile(True)
    acc = {"newton": 0, "krylov": 0, "worst_r": 0.0, "accepted": 0, "start_ratio": 0.0}
    sync()
    t0 = time.perf_counter()
    st = None
    for _ in range(n):
        st = lem.step()
        acc["newton"] += st["newton_iterations"]
        acc["krylov"] += st["krylov_iterations"]
        acc["accepted"] += st["warm_start_accepted"]
        # relative residual as the Newton stop measures it: ||F(z)||inf over
        # ||F(z^n)||inf of the last sub-step's start (tolerance = rel_tol times it)
        den = st["hillslope_tolerance"] / lem.options.newton_rel_tol
        if den > 0:
            acc["worst_r"] = max(acc["worst_r"], st["hillslope_residual"] / den)
        if st["hillslope_initial_residual"] > 0:
            acc["start_ratio"] = max(acc["start_ratio"], st["warm_start_residual"] / st["hillslope_initial_residual"])
        if not st["hillslope_converged"]:
            raise RuntimeError("hillslope Newton did not converge")
    sync()
    dt = time.perf_counter() - t0
    prof = None
    if profile:
        prof = lem.profile_totals()
        lem.profile(False)
    return dt, acc, st, prof


def learned_comparison(args, z, K, sync, plain_value):
    """Config 3: the learned warm start on the same 4096^2 workload.

    Training data: seeded 2048^2 LEM runs of the same physics (K
    log-uniform 1e-6..1e-4 per run) captured step by step; one snapshot
    holds (2048 / 256)^2 = 64 tile pairs, so `pairs` tile pairs need
    pairs / 64 snapshots.  Training: Adam lr 1e-3, batch 32 of 256^2 tiles
    + 16-cell halos, one pass over the pairs.  Timed: the 4096^2 step with
    the explicit-Euler start and with the learned start, each from the
    same surface as the plain run, to the same r <= 1e-8."""
    import lsdtt3_amd
    from lsdtt3_amd import learned as Lm
    from lsdtt3_amd.synthetic import lem_surface
    t0 = time.perf_counter()
    ds = Lm.Dataset()
    runs = 4
    per_run = max(1, int(np.ceil(args.pairs / 64 / runs)))
    rng = np.random.default_rng(77)
    for r in range(runs):
        Kr = float(10.0 ** rng.uniform(-6.0, -4.0))
        zr = lem_surface(args.train_size, args.train_size, dx=30.0, seed=100 + r, max_slope=0.75 * 0.8)
        m = lsdtt3_amd.LandscapeModel(zr, 30.0, **lem_kwargs(Kr))
        Lm.generate([m], ds, per_run)
        del m
    sync()
    t_gen = time.perf_counter() - t0
    net = Lm.UNet(4, 32, 4, seed=1, tile=256, halo=16)
    zero_loss = net.evaluate(ds, 2, 32, seed=5)[1]
    iters = max(1, int(np.ceil(len(ds) * 64 / 32)))
    t0 = time.perf_counter()
    losses = net.train(ds, iters, batch=32, lr=1e-3, seed=3)
    sync()
    t_train = time.perf_counter() - t0
    loss = net.evaluate(ds, 2, 32, seed=5)[0]
    win = 256 + 2 * 16
    tiles = ((args.size + 255) // 256) ** 2
    bm = net.benchmark(tiles, win, win, reps=2)
    out = {"training": {"snapshots": len(ds), "tile_pairs": len(ds) * 64, "snapshot_grid": args.train_size,
                        "runs": runs, "iterations": iters, "batch": 32, "lr": 1e-3,
                        "generate_s": round(t_gen, 2), "train_s": round(t_train, 2),
                        "ms_per_iteration": round(1e3 * t_train / iters, 2),
                        "loss_zero_prediction": zero_loss, "loss_trained": loss,
                        "first_batch_loss": float(losses[0]), "last_batch_loss": float(losses[-1])}}
    flops = bm["conv_flops"]
    tf = flops / (bm["conv_ms"] * 1e-3) / 1e12
    out["inference"] = {"windows": tiles, "window": win, "conv_flops": flops, "forward_ms": round(bm["forward_ms"], 3),
                        "conv_ms": round(bm["conv_ms"], 3),
                        "mfma_roofline": {"bound": "mfma", "achieved": round(tf, 1), "peak": BF16_PEAK_TFLOPS,
                                          "unit": "TFLOP/s", "frac": round(tf / BF16_PEAK_TFLOPS, 4),
                                          "kernel": "k_conv_tile (3x3 conv stack, forward)"}}
    for name, mode in (("explicit_start", Lm.WARM_EXPLICIT), ("previous_step_start", Lm.WARM_PREVIOUS),
                       ("learned_start", Lm.WARM_LEARNED)):
        lem = lsdtt3_amd.LandscapeModel(z, 30.0, **lem_kwargs(K))
        lem.tile_height = getattr(args, "tile_height", 16)  # as the plain run
        Lm.set_warm_start(lem, mode, net if mode == Lm.WARM_LEARNED else None)
        for _ in range(args.warmup):
            lem.step()
        lem.elevation = z            # the same trajectory as the plain run
        dt, a, st, _ = run_steps(lem, args.steps, sync)
        _, _, _, prof = run_steps(lem, min(args.steps, args.profile_steps), sync, profile=True)
        newton, krylov, acc = a["newton"], a["krylov"], a["accepted"]
        worst_r, ratio = a["worst_r"], a["start_ratio"]
        v = int(st["n_valid"]) * args.steps / dt
        out[name] = {"value": round(v, 1), "unit": "cell-steps/s", "vs_plain": round(v / plain_value, 4),
                     "ms_per_step": round(1e3 * dt / args.steps, 3),
                     "newton_iters_per_step": round(newton / args.steps, 2),
                     "krylov_iters_per_step": round(krylov / args.steps, 2),
                     "starts_accepted": acc, "worst_start_residual_ratio": ratio,
                     "worst_relative_residual": worst_r,
                     "predictor_ms_per_step": round(prof.get("warm_start", (0.0,))[0]
                                                    / min(args.steps, args.profile_steps), 3)}
        del lem
    return out


def bench_cnn(args):
    """the warm-start U-Net alone (random weights and inputs): forward at the
    4096^2 inference shape (256 windows of 288^2) and a training iteration
    at batch 32; MFMA utilisation of the 3x3 convolution stack"""
    from lsdtt3_amd import learned as Lm
    net = Lm.UNet(4, 32, 4, seed=1, tile=256, halo=16)
    win = 288
    tiles = ((args.size + 255) // 256) ** 2
    inf = net.benchmark(tiles, win, win, reps=max(1, args.steps))
    tr = net.benchmark(32, win, win, reps=max(1, args.steps))
    tf = inf["conv_flops"] / (inf["conv_ms"] * 1e-3) / 1e12
    out = {"metric": "U-Net 3x3 convolution TFLOP/s (forward, BF16 MFMA)", "value": round(tf, 1),
           "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
           "ms_per_step": round(inf["forward_ms"], 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "bf16", "data": "synthetic",
           "config": {"workload": f"warm-start U-Net (4 levels, 32 base channels, 3x3 convs) forward on {tiles} "
                                  f"windows of {win}x{win}x4 (the {args.size}^2 inference), and training at batch 32",
                      "params": net.num_params, "conv_flops_forward": inf["conv_flops"],
                      "forward_ms": round(inf["forward_ms"], 3), "conv_ms": round(inf["conv_ms"], 3),
                      "train_ms_batch32": round(tr["train_ms"], 3),
                      "forward_ms_batch32": round(tr["forward_ms"], 3)},
           "roofline": {"bound": "mfma", "kernel": "conv stack", "hip_kernel": "k_conv_tile",
                        "achieved": round(tf, 1), "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": round(tf / BF16_PEAK_TFLOPS, 4), "traffic": None}}
    print(json.dumps(out))


def bench_routing(args):
    """The routing pipeline on a device-resident float32 DEM: epsilon-free
    fill -> D8 flow object (pits as outlets; donors, Kahn order, s-vector,
    int64 counts) -> chi raster (theta 0.45, A0 = 1 m^2) from the counts
    (area = count x 100 m^2).  Surface: power-law noise (beta 2.2, 100 m std)
    with closed bowls on ~1% of the cells (depth U(0, 20] m) and exactly flat
    plateaus on ~5%, dx = 10 m, boundary "oooo"."""
    import ctypes as C
    from lsdtt3_amd import _native as N
    from lsdtt3_amd.synthetic import routing_surface
    L = N.lib()
    n = args.size
    z = routing_surface(n, n, seed=2024).astype(np.float32)
    dz = N.DeviceArray.from_host(z)
    dm = N.DeviceArray.from_host(np.ones((n, n), np.uint8))
    dfill = N.DeviceArray((n, n), np.float32)
    dchi = N.DeviceArray((n, n), np.float32)
    o = N.FlowOptions()
    L.lsdtt_gpu_flow_options_init(C.byref(o), n, n, 10.0)
    o.boundary_conditions = b"oooo"
    o.allow_pits_as_outlets = 1
    fh = L.lsdtt_gpu_flow_create(C.byref(o))
    if not fh:
        raise RuntimeError((L.lsdtt_gpu_last_error() or b"").decode())

    def stage_fill():
        N.check(L.lsdtt_gpu_fill_device(dz.ptr, dm.ptr, n, n, -9999.0, b"oooo", dfill.ptr, None))

    def stage_flow():
        N.check(L.lsdtt_gpu_flow_build_device(fh, dfill.ptr, 0, dm.ptr, None))

    def stage_chi():
        N.check(L.lsdtt_gpu_flow_chi_device(fh, 0.45, 1.0, -9999.0, dchi.ptr, None))

    stages = [("fill", stage_fill), ("flow_object", stage_flow), ("chi", stage_chi)]
    for _ in range(max(1, args.warmup)):
        for _, f in stages:
            f()
    N.synchronize()
    # whole pipeline, no synchronisation between stages
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for _, f in stages:
            f()
    N.synchronize()
    dt = time.perf_counter() - t0
    # stage breakdown (each stage timed alone, synchronised)
    per = {}
    for k, f in stages:
        t1 = time.perf_counter()
        for _ in range(args.steps):
            f()
        N.synchronize()
        per[k] = 1e3 * (time.perf_counter() - t1) / args.steps
    # the result is the oracle's (bit-exact parity is the routing tests' job);
    # here only a sanity check that chi was written
    chi = dchi.to_host()
    assert np.isfinite(chi).any()
    nv, no, md = C.c_int64(), C.c_int64(), C.c_int64()
    N.check(L.lsdtt_gpu_flow_sizes(fh, C.byref(nv), C.byref(no), C.byref(md)))
    L.lsdtt_gpu_flow_destroy(fh)
    cells = n * n
    value = cells * args.steps / dt
    # algorithmic bytes per cell of the pipeline (each array element a
    # kernel must read or write counted once): fill 9 (z 4, mask 1, filled
    # 4), receivers 10, donors + outlets 5, Kahn order 13, counts 24,
    # s-vector 28, chi on every valid cell 29
    bpc = 9 + 10 + 5 + 13 + 24 + 28 + 29
    achieved = bpc * cells / (dt / args.steps) / 1e9
    out = {"metric": "routing pipeline cells/s at 4096^2 (fill + D8 flow object + chi, bit-exact)",
           "value": round(value, 1), "unit": "cells/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(1e3 * dt / args.steps, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32 z, int64 counts", "data": "synthetic",
           "config": {"workload": f"routing surface {n}x{n} dx=10 (beta 2.2, 100 m std, 1% closed bowls, 5% "
                                  f"flat plateaus), bc oooo, pits as outlets, theta 0.45, A0 1",
                      "n_valid": int(nv.value), "n_outlets": int(no.value), "max_depth": int(md.value),
                      "stage_ms": {k: round(v, 3) for k, v in per.items()}},
           "roofline": {"bound": "hbm", "kernel": "whole pipeline", "achieved": round(achieved, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "traffic": None, "algorithmic_bytes_per_cell": bpc}}
    if not args.no_cpu_baseline:
        out["cpu_baseline"], ref = routing_cpu_baseline(z, per["fill"])
        if ref is not None:
            out["reference_fill"] = ref
    print(json.dumps(out))


def routing_cpu_baseline(z, gpu_fill_ms):
    """The same pipeline on the host: the oracle's fill, flow object and chi
    (serial, as the reference's fill and accumulation are), once over the
    full surface; and, where it was built, the reference's own fill
    (lsdtt::fill_pits, compiled from its sources) on the same surface."""
    from oracle import oracle as orc
    n = z.shape[0] * z.shape[1]
    mask = np.ones(z.shape, np.uint8)
    t0 = time.perf_counter()
    zf = orc.fill(z, mask)
    t1 = time.perf_counter()
    flow = orc.flow_build(zf, mask, 10.0, "oooo", allow_pits_as_outlets=True)
    orc.chi_raster(flow, mask, 10.0, 0.45, 1.0)
    t2 = time.perf_counter()
    base = {"value": round(n / (t2 - t0), 1), "unit": "cells/s", "cores": 1, "kind": "port",
            "sample": f"oracle fill + D8 flow object + chi on the same {z.shape[0]}x{z.shape[1]} surface, "
                      f"once, 1 thread (fill {1e3 * (t1 - t0):.0f} ms, flow+chi {1e3 * (t2 - t1):.0f} ms)"}
    ref = None
    if orc.ref_fill_available():
        t0 = time.perf_counter()
        zr = orc.ref_fill(z, mask, 10.0, "oooo", n_threads=1)
        el = time.perf_counter() - t0
        ref = {"value": round(n / el, 1), "unit": "cells/s", "cores": 1, "kind": "reference",
               "ms": round(1e3 * el, 1), "gpu_fill_ms": round(gpu_fill_ms, 3),
               "gpu_speedup": round(1e3 * el / gpu_fill_ms, 1),
               "bit_exact_vs_oracle": bool(np.array_equal(zr.view(np.uint32), zf.view(np.uint32))),
               "sample": "lsdtt::fill_pits(dem, 0, \"oooo\", 1) built from the reference's sources, "
                         "once over the full surface"}
    return base, ref


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--summarize-trace":
        summarize_trace(*sys.argv[2:7])
        return
    if len(sys.argv) > 1 and sys.argv[1] == "--summarize-pmc":
        summarize_pmc(*sys.argv[2:7])
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--workload", choices=["lem4096", "ensemble", "cnn", "routing", "plain512"], default="lem4096")
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--fluvial", choices=["d8", "dinf"], default="d8",
                    help="lem4096: D8 (the metric's workload) or D-infinity incision")
    ap.add_argument("--channel-threshold", type=float, default=0.0,
                    help="lem4096: incision_threshold_area (m, on area/dx); > 0 derives the hillslope's "
                         "channel set from the D-infinity area each step")
    ap.add_argument("--autotune-steps", type=int, default=2,
                    help="timed steps per tile height when autotuning the routing tiles (0: default height)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned warm-start comparison")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-ensemble measurement")
    ap.add_argument("--ensemble-steps", type=int, default=300, help="batched ensemble: timed steps")
    ap.add_argument("--no-ensemble-warm-start", action="store_true",
                    help="batched ensemble: skip the run from the previous hillslope change")
    ap.add_argument("--no-unsmoothed", action="store_true",
                    help="skip the raw-surface (linear, n = 1) measurement at the same size")
    ap.add_argument("--profile-steps", type=int, default=10,
                    help="steps of the separate HIP-event-profiled run (per-kernel breakdown, roofline)")
    ap.add_argument("--pairs", type=int, default=20000, help="learned: training tile pairs")
    ap.add_argument("--train-size", type=int, default=2048, help="learned: training LEM grid side")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: the library has no multi-GPU path (asked for {args.gpus}); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import lsdtt3_amd
    if not lsdtt3_amd.gpu_available():
        print("no GPU visible to liblsdtt3_gpu.so", file=sys.stderr)
        sys.exit(1)
    if args.workload == "ensemble":
        bench_ensemble(args)
        return
    if args.workload == "cnn":
        bench_cnn(args)
        return
    if args.workload == "routing":
        bench_routing(args)
        return
    if args.workload == "plain512":
        bench_plain_cpu_workload(args)
        return
    from lsdtt3_amd import _native as N
    sync = N.synchronize

    STEP_VARIANT["fluvial_method"] = args.fluvial
    STEP_VARIANT["incision_threshold_specific_area"] = args.channel_threshold
    variant = ("" if args.fluvial == "d8" else ", D-infinity incision") + \
        (f", channels from D-inf area/dx >= {args.channel_threshold:g} m" if args.channel_threshold > 0 else "")
    z, K, passes = workload(args.size)
    lem = lsdtt3_amd.LandscapeModel(z, 30.0, **lem_kwargs(K))
    for _ in range(args.warmup):
        lem.step()
    tuned = None
    if args.autotune_steps > 0:
        th, tms = lem.autotune_tiles(args.autotune_steps)
        args.tile_height = th
        tuned = {"tile_height": th, "ms_per_step": {str(h): round(v, 3) for h, v in tms.items()},
                 "steps_per_trial": args.autotune_steps}
    # the timed run is the workload's trajectory from the initial surface:
    # warm-up and autotuning steps only warmed the library
    lem.elevation = z
    seg = 100 if args.steps % 100 == 0 else args.steps
    dt, newton, krylov, worst_r, segs = 0.0, 0, 0, 0.0, []
    for _ in range(args.steps // seg):
        sdt, a, st, _ = run_steps(lem, seg, sync)
        dt += sdt
        newton, krylov, worst_r = newton + a["newton"], krylov + a["krylov"], max(worst_r, a["worst_r"])
        segs.append({"ms_per_step": round(1e3 * sdt / seg, 3), "newton_iters_per_step": round(a["newton"] / seg, 2),
                     "krylov_iters_per_step": round(a["krylov"] / seg, 2)})
        progress(f"steps {len(segs) * seg}/{args.steps}: {segs[-1]}")
    psteps = min(args.steps, args.profile_steps)
    _, _, _, prof = run_steps(lem, psteps, sync, profile=True)

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
                               f"K={K:.3e}, U=5e-4 dt=500 bc=nnon{variant}",
                   "n_valid": n_valid, "smoothing_passes": passes,
                   "newton_iters_per_step": round(newton / args.steps, 2),
                   "krylov_iters_per_step": round(krylov / args.steps, 2),
                   "worst_relative_residual": worst_r,
                   "timed_steps_from_initial_surface": args.steps,
                   "segments": segs if len(segs) > 1 else None,
                   "count_passes_last_step": int(st["count_passes"]),
                   "stream_power_passes_last_step": int(st["stream_power_passes"]),
                   "count_tile_runs_last_step": int(st["count_tile_runs"]),
                   "stream_power_tile_runs_last_step": int(st["stream_power_tile_runs"]),
                   "tile_height": lem.tile_height, "autotuned_tiles": tuned},
        "roofline": roofline,
        "kernels_ms_per_step": {k: round(v[0] / psteps, 4) for k, v in sorted(prof.items(), key=lambda kv: -kv[1][0])},
        "profiled_steps": psteps,
    }
    progress(f"plain trajectory done: {out['value']:.4g} cell-steps/s")
    if not args.no_unsmoothed:
        out["unsmoothed"] = unsmoothed_measure(args, sync)
        progress("raw-surface workload done")
    if not args.no_batched:
        del lem
        b = ensemble_measure(args)
        b["vs_plain"] = round(b["value"] / value, 4)
        out["batched"] = b
        progress("batched ensemble done")
    if not args.no_learned:
        out["learned_warm_start"] = learned_comparison(args, z, K, sync, value)
        progress("learned warm start done")
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(K, args.cpu_budget)
        progress("cpu baseline done")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
