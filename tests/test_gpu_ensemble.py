"""Batched ensemble (lsdtt_gpu_ensemble_*): instances of uneven sizes and
parameters stepped together -- one launch per kernel for all of them, each
instance's solver decisions on the device, the step replayed from a
captured hipGraph.

What is checked, and to which bar:
  * graph replay and the same launches enqueued one by one: bit-identical
    elevations and statistics over several steps;
  * an instance alone (a batch of one) and inside the batch: bit-identical
    (per-instance reductions, so no dependence on the batch);
  * fluvial-only instances against the single-grid engine
    (lsdtt_gpu_lem_*): bit-identical elevations (same per-cell arithmetic);
  * with the Newton hillslope, against the oracle's FP64 step (and the
    single-grid engine): within 4 dt tol, as in test_gpu_lem.py (each side
    solves to its own residual tolerance; the Krylov iterates differ by
    rounding);
  * the Picard fallback of a failing Newton solve against the oracle
    (1e-7 m, as test_gpu_lem.py);
  * the warm start from each instance's previous hillslope change: tried
    from the second step on, forgotten when a surface is set, graph replay
    and direct launches still bit-identical, and the surfaces within the
    Newton tolerance of the plain start's (4 dt tol per step, accumulated)."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import lsdtt3_amd
    assert lsdtt3_amd.gpu_available()
    return lsdtt3_amd


def models(hillslope=True):
    from lsdtt3_amd.synthetic import lem_surface
    shapes = [(40, 52), (120, 96), (64, 64), (33, 150), (90, 70), (150, 41), (257, 300)]
    out = []
    for i, (nr, nc) in enumerate(shapes):
        rng = np.random.default_rng(100 + i)
        nonlinear = i % 2 == 1
        out.append(dict(dem=lem_surface(nr, nc, dx=30.0, seed=60 + i, max_slope=0.6), cell_size=30.0,
                        timestep_years=500.0, erodibility=float(np.exp(rng.uniform(np.log(1e-6), np.log(1e-4)))),
                        m_exponent=float(rng.uniform(0.3, 0.6)) if i != 2 else 0.5,
                        n_exponent=1.0 if i % 3 else 1.5,
                        uplift_rate=float(np.exp(rng.uniform(np.log(1e-4), np.log(1e-3)))),
                        hillslope_diffusivity=float(np.exp(rng.uniform(np.log(1e-3), np.log(1e-1)))),
                        nonlinear_hillslopes=nonlinear, critical_hillslope_slope=0.8,
                        hillslope_model="roering_implicit" if hillslope else "none",
                        baselevel_fall_rate=1e-4 if i == 4 else 0.0,
                        boundary_conditions=["nnon", "oooo", "popo", "nnnn", "onon", "bnnn", "nnon"][i]))
    return out


def oracle_params(m):
    return orc.lem_params(m["timestep_years"], m["erodibility"], m["m_exponent"], m["n_exponent"],
                          m["uplift_rate"], m["hillslope_diffusivity"], Sc=m["critical_hillslope_slope"],
                          nonlinear=m["nonlinear_hillslopes"], bc=m["boundary_conditions"],
                          hillslope=m["hillslope_model"] != "none", baselevel_fall=m["baselevel_fall_rate"])


def test_gpu_ensemble_graph_replay_equals_direct_launches(gpu):
    print("[gpu ensemble] graph vs direct launches, 7 instances x 4 steps", flush=True)
    ms = models()
    eg = gpu.Ensemble(ms)
    ed = gpu.Ensemble(ms)
    ed.set_graph(False)
    for step in range(4):
        sg, _ = eg.run(1)
        sd, _ = ed.run(1)
        for i in range(len(ms)):
            # host work between the replays: read every surface back
            assert np.array_equal(eg.elevation(i), ed.elevation(i)), (step, i)
            assert sg[i] == sd[i], (step, i)
    p = eg.pattern()
    assert p["graph_launches"] == 4 and p["graph_captures"] >= 1
    assert ed.pattern()["graph_launches"] == 0


def test_gpu_ensemble_instance_independent_of_batch(gpu):
    print("[gpu ensemble] each instance alone vs in the batch", flush=True)
    ms = models()
    ens = gpu.Ensemble(ms)
    stats, sec = ens.run(3)
    assert sec > 0 and len(stats) == len(ms)
    for i, m in enumerate(ms):
        one = gpu.Ensemble([m])
        s1, _ = one.run(3)
        assert np.array_equal(one.elevation(0), ens.elevation(i)), i
        # every per-instance figure; the sweeps' pass counts are the batch's
        batch_wide = ("count_passes", "stream_power_passes", "count_tile_runs", "stream_power_tile_runs")
        assert {k: v for k, v in s1[0].items() if k not in batch_wide} == \
            {k: v for k, v in stats[i].items() if k not in batch_wide}, i


def test_gpu_ensemble_fluvial_equals_single_engine(gpu):
    print("[gpu ensemble] fluvial only vs single-grid engine", flush=True)
    ms = models(hillslope=False)
    ens = gpu.Ensemble(ms)
    for step in range(3):
        stats, _ = ens.run(1)
        for i, m in enumerate(ms):
            if step == 0:
                kw = dict(m)
                z = kw.pop("dem")
                m["_lem"] = gpu.LandscapeModel(z, kw.pop("cell_size"), **kw)
            st = m["_lem"].step()
            assert np.array_equal(ens.elevation(i), m["_lem"].elevation), (step, i)
            assert stats[i]["cells_eroded"] == st["cells_eroded"], (step, i)
            assert stats[i]["max_local_elevation_change"] == st["max_local_elevation_change"], (step, i)
            assert stats[i]["p999_local_elevation_change"] == st["p999_local_elevation_change"], (step, i)
            for k in ("delta_mean_elevation", "mean_erosion_rate"):
                assert abs(stats[i][k] - st[k]) <= 1e-12 * max(abs(st[k]), 1.0), (step, i, k)


def test_gpu_ensemble_matches_oracle_and_single_engine(gpu):
    print("[gpu ensemble] vs oracle, 2 steps", flush=True)
    ms = models()
    ens = gpu.Ensemble(ms)
    zo = [m["dem"].copy() for m in ms]
    for step in range(2):
        stats, _ = ens.run(1)
        for i, m in enumerate(ms):
            mask = np.ones(m["dem"].shape, np.uint8)
            zo[i], so, rc = orc.lem_step(zo[i], mask, 30.0, oracle_params(m))
            assert rc == 0
            st = stats[i]
            assert st["hillslope_converged"] == 1 and st["hillslope_method"] == 0, (step, i)
            assert st["hillslope_residual"] <= st["hillslope_tolerance"], (step, i)
            assert st["n_valid"] == so["n_valid"]
            tol = max(4 * 500.0 * st["hillslope_tolerance"], 1e-8)
            assert np.abs(ens.elevation(i) - zo[i]).max() <= tol, (step, i, np.abs(ens.elevation(i) - zo[i]).max())
            for k in ("delta_max_elevation", "max_local_elevation_change"):
                assert abs(st[k] - so[k]) <= 2 * tol, (step, i, k)
    # and against the single-grid engine after the same two steps
    for i, m in enumerate(ms):
        kw = dict(m)
        z = kw.pop("dem")
        lem = gpu.LandscapeModel(z, kw.pop("cell_size"), **kw)
        s = lem.run(2)
        tol = max(4 * 500.0 * s["hillslope_tolerance"], 1e-8)
        assert np.abs(ens.elevation(i) - lem.elevation).max() <= 2 * tol, i


def test_gpu_ensemble_picard_fallback_matches_oracle(gpu):
    """one Newton iteration allowed in one instance: its sub-steps fail down
    to the minimum, it restarts from the post-incision surface with the
    damped Picard solver (as the single-grid engine and the oracle); the
    other instance is unaffected"""
    print("[gpu ensemble] Picard fallback", flush=True)
    from lsdtt3_amd.synthetic import lem_surface
    base = dict(cell_size=30.0, timestep_years=500.0, erodibility=3e-5, m_exponent=0.5, n_exponent=1.5,
                uplift_rate=5e-4, hillslope_diffusivity=0.005, nonlinear_hillslopes=True,
                critical_hillslope_slope=0.8, boundary_conditions="nnon", picard_tol=1e-9, picard_max_iter=200)
    m0 = dict(base, dem=lem_surface(40, 48, dx=30.0, seed=52, max_slope=0.6), newton_max_iterations=1)
    m1 = dict(base, dem=lem_surface(50, 44, dx=30.0, seed=53, max_slope=0.6))
    ens = gpu.Ensemble([m0, m1])
    stats, _ = ens.run(1)
    assert stats[0]["hillslope_method"] == 1 and stats[1]["hillslope_method"] == 0
    pp = orc.picard_params(0.005, Sc=0.8, nonlinear=True, picard_tol=1e-9, picard_max_iter=200)
    p = orc.lem_params(500.0, 3e-5, 0.5, 1.5, 5e-4, 0.005, Sc=0.8, nonlinear=True, bc="nnon", picard=pp,
                       max_iterations=1)
    zo, so, rc = orc.lem_step(m0["dem"], np.ones(m0["dem"].shape, np.uint8), 30.0, p)
    assert so["hillslope_method"] == 1
    assert np.abs(ens.elevation(0) - zo).max() <= 1e-7
    one = gpu.Ensemble([m1])
    one.run(1)
    assert np.array_equal(one.elevation(0), ens.elevation(1))


def test_gpu_ensemble_solver_rounds_beyond_the_pattern(gpu):
    """a first step whose slowest instance needs more solver rounds than
    the starting pattern holds: the rest run outside the graph, the next
    pattern grows, and the results equal the direct launches"""
    print("[gpu ensemble] rounds beyond the pattern", flush=True)
    ms = models()[:3]
    for m in ms:
        m["krylov_rel_tol"] = 1e-7   # many BiCGSTAB iterations per Newton iteration
        m["krylov_forcing"] = 0      # at every Newton iteration
    eg = gpu.Ensemble(ms)
    ed = gpu.Ensemble(ms)
    ed.set_graph(False)
    eg.run(1)
    ed.run(1)
    p = eg.pattern()
    assert p["extra_rounds"] > 0
    eg.run(2)
    ed.run(2)
    assert eg.pattern()["rounds"] > 8
    for i in range(len(ms)):
        assert np.array_equal(eg.elevation(i), ed.elevation(i)), i


def test_gpu_ensemble_erosion_rate_field(gpu):
    ms = models()[:2]
    for m in ms:
        m["compute_erosion_rate_field"] = 1
    ens = gpu.Ensemble(ms)
    z0 = [m["dem"] for m in ms]
    stats, _ = ens.run(1)
    for i, m in enumerate(ms):
        er = ens.erosion_rate(i)
        want = ((z0[i] - (ens.elevation(i) - m["uplift_rate"] * 500.0)) / 500.0).astype(np.float32)
        np.testing.assert_allclose(er, want, rtol=1e-6, atol=1e-12)
        pos = np.clip(er.astype(np.float64), 0.0, None)
        assert abs(pos.max() - stats[i]["max_erosion_rate"]) <= 1e-6 * stats[i]["max_erosion_rate"]


def test_gpu_ensemble_rejects_unsupported_options(gpu):
    from lsdtt3_amd._native import GpuError
    ms = models()[:2]
    ms[1]["drainage_area_kernel_size"] = 2
    with pytest.raises(GpuError):
        gpu.Ensemble(ms)
    ms = models()[:2]
    ms[0]["fluvial_method"] = "dinf"
    with pytest.raises(GpuError):
        gpu.Ensemble(ms)


def test_gpu_ensemble_damped_line_search_matches_single_engine(gpu):
    """an instance whose Newton iterations need line-search steps below
    lambda = 1 after a few steps (a Roering surface steepened by incision):
    the batched solve takes the same sub-steps as the single-grid engine
    (no rejections) and agrees with it within 4 dt tol every step"""
    print("[gpu ensemble] damped line search, 10 steps", flush=True)
    from lsdtt3_amd.synthetic import lem_surface
    m = dict(dem=lem_surface(434, 345, dx=30.0, seed=2024006, max_slope=0.6), cell_size=30.0,
             timestep_years=500.0, erodibility=7.988049563613928e-05, m_exponent=0.5687867749125224,
             n_exponent=1.0, uplift_rate=0.0009256564972093344, hillslope_diffusivity=0.007291700011174901,
             nonlinear_hillslopes=True, critical_hillslope_slope=0.8, boundary_conditions="nnon")
    ens = gpu.Ensemble([m])
    kw = dict(m)
    z = kw.pop("dem")
    lem = gpu.LandscapeModel(z, kw.pop("cell_size"), **kw)
    for step in range(10):
        sb, _ = ens.run(1)
        ss = lem.step()
        assert sb[0]["hillslope_rejected"] == ss["hillslope_rejected"] == 0, (step, sb[0], ss)
        assert sb[0]["hillslope_substeps"] == ss["hillslope_substeps"] == 1, step
        tol = max(4 * 500.0 * ss["hillslope_tolerance"], 1e-8)
        assert np.abs(ens.elevation(0) - lem.elevation).max() <= 2 * tol, step
        lem.elevation = ens.elevation(0)  # keep both on one trajectory


def test_gpu_ensemble_warm_start_previous_step(gpu):
    print("[gpu ensemble] warm start from the previous hillslope change, 7 instances x 6 steps", flush=True)
    ms = models()
    a, b, d = gpu.Ensemble(ms), gpu.Ensemble(ms), gpu.Ensemble(ms)
    b.set_warm_start(3)
    d.set_warm_start(3)
    d.set_graph(False)
    bound = [0.0] * len(ms)
    accepted = 0
    for k in range(6):
        sa, _ = a.run(1)
        sb, _ = b.run(1)
        sd, _ = d.run(1)
        for i in range(len(ms)):
            assert sb[i] == sd[i], (k, i)
            assert np.array_equal(b.elevation(i), d.elevation(i)), (k, i)
            s = sb[i]
            assert s["hillslope_converged"] == 1 and s["hillslope_residual"] <= s["hillslope_tolerance"], (k, i)
            assert s["warm_start_mode"] == (0 if k == 0 else 3), (k, i)
            if s["warm_start_accepted"]:
                accepted += 1
                assert s["warm_start_residual"] < s["hillslope_initial_residual"], (k, i)
            bound[i] += 4 * 500.0 * max(sa[i]["hillslope_tolerance"], s["hillslope_tolerance"])
    assert accepted >= 3 * len(ms)  # most instances keep the predicted start
    for i in range(len(ms)):
        assert np.abs(a.elevation(i) - b.elevation(i)).max() <= max(bound[i], 1e-8), i
    # a new surface forgets its instance's change only
    b.set_elevation(0, b.elevation(0))
    sb, _ = b.run(1)
    assert sb[0]["warm_start_mode"] == 0
    assert all(s["warm_start_mode"] == 3 for s in sb[1:])
    from lsdtt3_amd._native import GpuError
    with pytest.raises(GpuError):
        b.set_warm_start(1)
