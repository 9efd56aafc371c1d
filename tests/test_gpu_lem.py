"""GPU LEM timestep (through the C ABI) against the oracle's FP64 step.

Tolerances, stated per output:
  * fluvial only: |dz| <= 1e-9 m.  The device and host evaluate the same
    per-node expressions in the same order with no FMA contraction; only
    pow() differs (ocml vs glibc, <= 1 ulp), and the n != 1 Newton stops at
    slope_clamp * dx_eff, so a 1-ulp difference can at most shift one
    iteration.
  * with hillslope: both solve F(z) = 0 to ||F||inf <= rel_tol ||F(z^n)||inf,
    so each result is within dt * tol / (1 - rho) of the exact root, rho < 1
    the Jacobi contraction of I + dt J_L; the test allows 4 * dt * tol.
"""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import lsdtt3_amd
    assert lsdtt3_amd.gpu_available()
    return lsdtt3_amd


def run_both(gpu, z, dx, steps, kw, okw, mask=None):
    mask = np.ones(z.shape, np.uint8) if mask is None else mask
    lem = gpu.LandscapeModel(z, dx, mask=mask, **kw)
    p = orc.lem_params(**okw)
    zo = z.copy()
    for i in range(steps):
        st_g = lem.step()
        zo, st_o, rc = orc.lem_step(zo, mask, dx, p)
        assert rc == 0
    return lem.elevation, zo, st_g, st_o


@pytest.mark.parametrize("nexp", [1.0, 1.5])
def test_gpu_fluvial_only_matches_oracle(gpu, nexp):
    print(f"[gpu lem fluvial] n={nexp}", flush=True)
    from lsdtt3_amd.synthetic import lem_surface
    z = lem_surface(150, 170, dx=30.0, seed=5)
    kw = dict(timestep_years=1000.0, erodibility=1e-5, m_exponent=0.5, n_exponent=nexp, uplift_rate=1e-3,
              hillslope_model="none", boundary_conditions="nnon")
    okw = dict(dt=1000.0, K=1e-5, m=0.5, n=nexp, U=1e-3, D=0.0, bc="nnon", hillslope=False)
    zg, zo, sg, so = run_both(gpu, z, 30.0, 3, kw, okw)
    assert np.abs(zg - zo).max() <= 1e-9
    assert sg["cells_eroded"] == so["cells_eroded"]
    assert sg["n_valid"] == so["n_valid"]
    assert abs(sg["max_local_elevation_change"] - so["max_local_elevation_change"]) <= 1e-9
    assert abs(sg["p999_local_elevation_change"] - so["p999_local_elevation_change"]) <= 1e-6


def test_gpu_fluvial_erodibility_field(gpu):
    from lsdtt3_amd.synthetic import lem_surface
    z = lem_surface(64, 80, dx=30.0, seed=6)
    K = np.random.default_rng(1).uniform(1e-6, 1e-4, z.shape)
    mask = np.ones(z.shape, np.uint8)
    lem = gpu.LandscapeModel(z, 30.0, timestep_years=500.0, m_exponent=0.5, n_exponent=1.5,
                             hillslope_model="none", boundary_conditions="nnon", erodibility_field=K)
    lem.step()
    zo, _, _ = orc.lem_step(z, mask, 30.0, orc.lem_params(500.0, 1e-3, 0.5, 1.5, 1e-4, 0.0, bc="nnon",
                                                          hillslope=False), K_field=K)
    assert np.abs(lem.elevation - zo).max() <= 1e-9


def test_gpu_linear_lem_matches_oracle(gpu):
    """config-0 physics (linear D, n = 1) on a 128 x 160 grid, 4 steps"""
    print("[gpu lem linear]", flush=True)
    from lsdtt3_amd.synthetic import lem_surface
    z = lem_surface(128, 160, dx=30.0, seed=7)
    kw = dict(timestep_years=1000.0, erodibility=1e-5, m_exponent=0.5, n_exponent=1.0, uplift_rate=1e-3,
              hillslope_diffusivity=0.01, boundary_conditions="nnon")
    okw = dict(dt=1000.0, K=1e-5, m=0.5, n=1.0, U=1e-3, D=0.01, bc="nnon")
    zg, zo, sg, so = run_both(gpu, z, 30.0, 4, kw, okw)
    tol = 4 * 1000.0 * max(sg["hillslope_tolerance"], 1e-12)
    assert sg["hillslope_converged"] == 1
    assert sg["hillslope_residual"] <= sg["hillslope_tolerance"]
    assert np.abs(zg - zo).max() <= max(tol, 1e-8), (np.abs(zg - zo).max(), tol)


def test_gpu_plain_cpu_workload_full_trajectory(gpu):
    """the plain CPU workload as written: 512 x 512 raw (unsmoothed) surface
    at dx = 30 m (0.01 tilt, beta-2 noise with 50 m std, 200 pits 1-5 m),
    linear D = 0.01, K = 1e-5, m = 0.5, n = 1, U = 1e-3, dt = 1000 yr, south
    edge fixed, 200 steps.  Each step solves to ||F|| <= tol, so the states
    may drift apart by at most about 4 dt tol per step; a receiver choice
    flipped by a last-bit difference would show as a jump far above that."""
    print("[gpu lem plain cpu workload 512^2 x 200]", flush=True)
    from lsdtt3_amd.synthetic import lem_surface
    z = lem_surface(512, 512, dx=30.0, seed=2024)
    mask = np.ones(z.shape, np.uint8)
    kw = dict(timestep_years=1000.0, erodibility=1e-5, m_exponent=0.5, n_exponent=1.0, uplift_rate=1e-3,
              hillslope_diffusivity=0.01, boundary_conditions="nnon")
    lem = gpu.LandscapeModel(z, 30.0, mask=mask, **kw)
    p = orc.lem_params(dt=1000.0, K=1e-5, m=0.5, n=1.0, U=1e-3, D=0.01, bc="nnon")
    bf = orc.boundary_flats(mask, "nnon")
    zo = z.copy()
    bound = 0.0
    for i in range(200):
        sg = lem.step()
        zo, so, rc = orc.lem_step(zo, mask, 30.0, p, bf)
        assert rc == 0 and sg["hillslope_converged"] == 1
        assert sg["cells_eroded"] == so["cells_eroded"], i
        bound += 4 * 1000.0 * max(sg["hillslope_tolerance"], 1e-12)
    d = np.abs(lem.elevation - zo).max()
    print(f"[gpu lem plain cpu workload] max |dz| after 200 steps {d:.3e} (bound {bound:.3e})", flush=True)
    assert d <= max(bound, 1e-7)
    assert abs(sg["mean_erosion_rate"] - so["mean_erosion_rate"]) <= 1e-9


def test_gpu_roering_lem_matches_oracle(gpu):
    """config-2 physics (Roering, Sc = 0.8, n = 1.5) on a sub-critical start"""
    print("[gpu lem roering]", flush=True)
    from lsdtt3_amd.synthetic import lem_surface
    z = lem_surface(96, 112, dx=30.0, seed=8, max_slope=0.6)
    kw = dict(timestep_years=500.0, erodibility=3e-5, m_exponent=0.5, n_exponent=1.5, uplift_rate=5e-4,
              hillslope_diffusivity=0.005, nonlinear_hillslopes=True, critical_hillslope_slope=0.8,
              boundary_conditions="nnon")
    okw = dict(dt=500.0, K=3e-5, m=0.5, n=1.5, U=5e-4, D=0.005, Sc=0.8, nonlinear=True, bc="nnon")
    zg, zo, sg, so = run_both(gpu, z, 30.0, 3, kw, okw)
    assert sg["hillslope_converged"] == 1
    assert sg["hillslope_residual"] <= sg["hillslope_tolerance"]
    tol = 4 * 500.0 * sg["hillslope_tolerance"]
    assert np.abs(zg - zo).max() <= max(tol, 1e-8), (np.abs(zg - zo).max(), tol)


def test_gpu_boundary_restore_and_mask(gpu):
    rng = np.random.default_rng(3)
    z = rng.random((40, 50)) * 5 + 100
    mask = np.ones(z.shape, np.uint8)
    mask[10:14, 20:26] = 0
    lem = gpu.LandscapeModel(z, 10.0, mask=mask, timestep_years=100.0, uplift_rate=1e-3,
                             hillslope_diffusivity=0.02, boundary_conditions="onon", baselevel_fall_rate=1e-4)
    lem.step()
    zg = lem.elevation
    bf = gpu.build_boundary_flats(mask, "onon")
    assert np.array_equal(bf, orc.boundary_flats(mask, "onon"))
    np.testing.assert_allclose(zg.ravel()[bf], z.ravel()[bf] - 0.01, rtol=0, atol=1e-12)
    zo, so, rc = orc.lem_step(z, mask, 10.0, orc.lem_params(100.0, 1e-3, 0.45, 1.0, 1e-3, 0.02, bc="onon",
                                                            baselevel_fall=1e-4))
    assert np.abs(zg - zo)[mask != 0].max() <= 1e-6


@pytest.mark.parametrize("bc,holes", [("popo", False), ("pnpn", True), ("nnnn", True), ("bnbn", False)])
def test_gpu_fluvial_boundaries_and_holes(gpu, bc, holes):
    """tile passes across periodic wraps, masked holes and base-level edges;
    grid sizes not multiples of the 32 x 32 tile"""
    print(f"[gpu lem fluvial bc] {bc} holes={holes}", flush=True)
    from lsdtt3_amd.synthetic import lem_surface
    z = lem_surface(101, 150, dx=30.0, seed=11)
    mask = np.ones(z.shape, np.uint8)
    if holes:
        mask[np.random.default_rng(2).random(z.shape) < 0.03] = 0
    kw = dict(timestep_years=1000.0, erodibility=2e-5, m_exponent=0.5, n_exponent=1.5, uplift_rate=1e-3,
              hillslope_model="none", boundary_conditions=bc)
    okw = dict(dt=1000.0, K=2e-5, m=0.5, n=1.5, U=1e-3, D=0.0, bc=bc, hillslope=False)
    zg, zo, sg, so = run_both(gpu, z, 30.0, 2, kw, okw, mask=mask)
    assert np.abs(zg - zo)[mask != 0].max() <= 1e-9
    assert sg["cells_eroded"] == so["cells_eroded"]


def test_gpu_fluvial_long_channel(gpu):
    """one 1-D channel of 3000 cells crosses ~94 tiles; the ridges beside it
    fall eastward faster than the channel rises, so the tiles' mean
    elevations run against the flow and every tile crossing is visited in
    the wrong order: the sweeps need many passes and must still agree"""
    x = np.arange(3000, dtype=np.float64)
    z = np.empty((3, 3000))
    z[1] = 100.0 + 0.3 * x
    z[0] = z[2] = 5000.0 - 1.0 * x
    kw = dict(timestep_years=100.0, erodibility=1e-4, m_exponent=0.5, n_exponent=1.0, uplift_rate=1e-3,
              hillslope_model="none", boundary_conditions="nnno")
    okw = dict(dt=100.0, K=1e-4, m=0.5, n=1.0, U=1e-3, D=0.0, bc="nnno", hillslope=False)
    mask = np.ones(z.shape, np.uint8)
    lem = gpu.LandscapeModel(z, 10.0, mask=mask, **kw)
    p = orc.lem_params(**okw)
    zo = z.copy()
    for step in range(2):
        sg = lem.step()
        zo, so, rc = orc.lem_step(zo, mask, 10.0, p)
        assert rc == 0
        assert np.abs(lem.elevation - zo).max() <= 1e-9
        assert sg["cells_eroded"] == so["cells_eroded"]
        assert sg["stream_power_passes"] >= 2
        if step == 0:  # counts from zero: one pass per tile crossing
            assert sg["count_passes"] >= 2 and sg["count_tile_runs"] > 94
        else:  # same receivers: the previous counts are already the answer
            assert sg["count_passes"] == 1


def test_gpu_fluvial_long_channel_in_order(gpu):
    """the same channel with tile means falling along the flow settles in
    one pass of each sweep (tiles wait for their receivers' tiles)"""
    z = (np.arange(3000, dtype=np.float64) * 0.3)[None, :].repeat(3, 0) + 100.0
    z[0] += 50.0
    z[2] += 50.0
    kw = dict(timestep_years=100.0, erodibility=1e-4, m_exponent=0.5, n_exponent=1.5, uplift_rate=1e-3,
              hillslope_model="none", boundary_conditions="nnno")
    okw = dict(dt=100.0, K=1e-4, m=0.5, n=1.5, U=1e-3, D=0.0, bc="nnno", hillslope=False)
    zg, zo, sg, so = run_both(gpu, z, 10.0, 2, kw, okw)
    assert np.abs(zg - zo).max() <= 1e-9
    assert sg["stream_power_passes"] >= 1


def test_gpu_counts_start_from_unrelated_counts(gpu):
    """each step's counts start from the previous step's; after the
    elevation is replaced by an unrelated surface the step still matches the
    oracle's step on that surface"""
    from lsdtt3_amd.synthetic import lem_surface
    za = lem_surface(120, 140, dx=30.0, seed=31)
    zb = lem_surface(120, 140, dx=30.0, seed=32)[:, ::-1].copy()
    kw = dict(timestep_years=1000.0, erodibility=2e-5, m_exponent=0.5, n_exponent=1.5, uplift_rate=1e-3,
              hillslope_model="none", boundary_conditions="oooo")
    lem = gpu.LandscapeModel(za, 30.0, **kw)
    lem.step()
    lem.elevation = zb
    st = lem.step()
    mask = np.ones(zb.shape, np.uint8)
    zo, so, rc = orc.lem_step(zb, mask, 30.0, orc.lem_params(1000.0, 2e-5, 0.5, 1.5, 1e-3, 0.0, bc="oooo",
                                                             hillslope=False))
    assert rc == 0
    assert np.abs(lem.elevation - zo).max() <= 1e-9
    assert st["cells_eroded"] == so["cells_eroded"]


@pytest.mark.parametrize("k,bc", [(3, "nnon"), (5, "popo"), (7, "oooo")])
def test_gpu_area_smoothing_matches_oracle(gpu, k, bc):
    """drainage_area_moving_average_kernel_size_pixels > 1: the smoothed area
    (rolling sums in the reference's order) feeds the stream power"""
    print(f"[gpu lem smoothing] k={k} bc={bc}", flush=True)
    from lsdtt3_amd.synthetic import lem_surface
    z = lem_surface(90, 110, dx=30.0, seed=40 + k)
    mask = np.ones(z.shape, np.uint8)
    mask[np.random.default_rng(k).random(z.shape) < 0.02] = 0
    kw = dict(timestep_years=1000.0, erodibility=2e-5, m_exponent=0.5, n_exponent=1.5, uplift_rate=1e-3,
              hillslope_model="none", boundary_conditions=bc, drainage_area_kernel_size=k)
    okw = dict(dt=1000.0, K=2e-5, m=0.5, n=1.5, U=1e-3, D=0.0, bc=bc, hillslope=False, area_kernel=k)
    zg, zo, sg, so = run_both(gpu, z, 30.0, 2, kw, okw, mask=mask)
    assert np.abs(zg - zo)[mask != 0].max() <= 1e-9
    assert sg["cells_eroded"] == so["cells_eroded"]


def test_gpu_width_factor_matches_oracle(gpu):
    """multiply_stream_power_term_by_w_over_dx: K scaled by w / dx"""
    from lsdtt3_amd.synthetic import lem_surface
    z = lem_surface(70, 90, dx=30.0, seed=44)
    kw = dict(timestep_years=1000.0, erodibility=2e-5, m_exponent=0.5, n_exponent=1.0, uplift_rate=1e-3,
              hillslope_model="none", boundary_conditions="nnon", multiply_by_w_over_dx=1, channel_width_w=12.0)
    okw = dict(dt=1000.0, K=2e-5, m=0.5, n=1.0, U=1e-3, D=0.0, bc="nnon", hillslope=False, width_factor=12.0 / 30.0)
    zg, zo, sg, so = run_both(gpu, z, 30.0, 2, kw, okw)
    assert np.abs(zg - zo).max() <= 1e-9


def test_gpu_explicit_channel_mask_is_dirichlet(gpu):
    """explicit_channel_mask: channel pixels are fixed during the hillslope
    step (the fluvial step still erodes them)"""
    print("[gpu lem channel mask]", flush=True)
    from lsdtt3_amd.synthetic import lem_surface
    z = lem_surface(80, 96, dx=30.0, seed=45)
    mask = np.ones(z.shape, np.uint8)
    ch = np.zeros(z.shape, np.uint8)
    ch[40, 10:80] = 1
    ch[20:60, 48] = 1
    kw = dict(timestep_years=1000.0, erodibility=1e-5, m_exponent=0.5, n_exponent=1.0, uplift_rate=1e-3,
              hillslope_diffusivity=0.01, boundary_conditions="nnon")
    lem = gpu.LandscapeModel(z, 30.0, **kw)
    lem.set_channel_mask(ch)
    p = orc.lem_params(1000.0, 1e-5, 0.5, 1.0, 1e-3, 0.01, bc="nnon", channel=ch)
    zo = z.copy()
    for _ in range(2):
        st = lem.step()
        zo, so, rc = orc.lem_step(zo, mask, 30.0, p)
        assert rc == 0
    tol = 4 * 1000.0 * max(st["hillslope_tolerance"], 1e-12)
    assert np.abs(lem.elevation - zo).max() <= max(tol, 1e-8)
    # the coupled regime ignores the mask: same as no mask
    lem2 = gpu.LandscapeModel(z, 30.0, hillslope_coupling_regime=1, **kw)
    lem2.set_channel_mask(ch)
    lem3 = gpu.LandscapeModel(z, 30.0, **kw)
    lem2.step()
    lem3.step()
    assert np.array_equal(lem2.elevation, lem3.elevation)


def test_gpu_lem_option_errors(gpu):
    from lsdtt3_amd._native import GpuError
    z = np.random.default_rng(0).random((20, 24)) * 10 + 100
    for bad in (dict(drainage_area_kernel_size=2), dict(drainage_area_kernel_size=0),
                dict(multiply_by_w_over_dx=1, channel_width_w=0.0), dict(hillslope_coupling_regime=3)):
        with pytest.raises(GpuError):
            gpu.LandscapeModel(z, 10.0, **bad)


@pytest.mark.parametrize("nonlinear", [False, True])
def test_gpu_forced_picard_matches_oracle(gpu, nonlinear):
    """hillslope_force_picard: damped Picard (FP64 CG solves to 1e-14) on
    both sides; tolerance: both stop at max |dz| < picard_tol = 1e-9 m
    between iterates of a contraction, so each is within ~1e-9 of the
    same fixed point"""
    print(f"[gpu lem picard] nonlinear={nonlinear}", flush=True)
    from lsdtt3_amd.synthetic import lem_surface
    z = lem_surface(72, 88, dx=30.0, seed=51, max_slope=0.6)
    mask = np.ones(z.shape, np.uint8)
    mask[30:33, 40:44] = 0
    kw = dict(timestep_years=500.0, erodibility=3e-5, m_exponent=0.5, n_exponent=1.5, uplift_rate=5e-4,
              hillslope_diffusivity=0.005, nonlinear_hillslopes=nonlinear, critical_hillslope_slope=0.8,
              boundary_conditions="nnon", hillslope_force_picard=1, picard_tol=1e-9, picard_max_iter=200)
    pp = orc.picard_params(0.005, Sc=0.8, nonlinear=nonlinear, picard_tol=1e-9, picard_max_iter=200)
    okw = dict(dt=500.0, K=3e-5, m=0.5, n=1.5, U=5e-4, D=0.005, Sc=0.8, nonlinear=nonlinear, bc="nnon",
               force_picard=True, picard=pp)
    zg, zo, sg, so = run_both(gpu, z, 30.0, 2, kw, okw, mask=mask)
    assert sg["hillslope_method"] == 1 and so["hillslope_method"] == 1
    assert sg["picard_cg_iterations"] > 0
    assert sg["picard_substeps"] == so["picard_substeps"]
    assert np.abs(zg - zo)[mask != 0].max() <= 1e-7


def test_gpu_newton_failure_falls_back_to_picard(gpu):
    """one Newton iteration allowed: every Newton sub-step fails down to the
    minimum sub-step, the step restarts from the post-incision surface with
    Picard, on the device as in the oracle (lsdtt_lem_cpu.cpp:563-582)"""
    print("[gpu lem newton->picard]", flush=True)
    from lsdtt3_amd.synthetic import lem_surface
    z = lem_surface(40, 48, dx=30.0, seed=52, max_slope=0.6)
    kw = dict(timestep_years=500.0, erodibility=3e-5, m_exponent=0.5, n_exponent=1.5, uplift_rate=5e-4,
              hillslope_diffusivity=0.005, nonlinear_hillslopes=True, critical_hillslope_slope=0.8,
              boundary_conditions="nnon", newton_max_iterations=1, picard_tol=1e-9, picard_max_iter=200)
    pp = orc.picard_params(0.005, Sc=0.8, nonlinear=True, picard_tol=1e-9, picard_max_iter=200)
    okw = dict(dt=500.0, K=3e-5, m=0.5, n=1.5, U=5e-4, D=0.005, Sc=0.8, nonlinear=True, bc="nnon",
               picard=pp, max_iterations=1)
    zg, zo, sg, so = run_both(gpu, z, 30.0, 1, kw, okw)
    assert sg["hillslope_method"] == 1 and so["hillslope_method"] == 1
    assert np.abs(zg - zo).max() <= 1e-7


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_gpu_p999_is_the_exact_order_statistic(gpu, seed):
    """the step's p99.9 |dz| (radix selection on the device) is element
    floor(0.999 (n - 1)) of the sorted float |z_after - z_before| of the
    device's own surfaces (lsdtt_statistics.cpp:420-429), bit for bit"""
    from lsdtt3_amd.synthetic import lem_surface
    z = lem_surface(97, 113, dx=30.0, seed=seed, max_slope=0.6)
    mask = np.ones(z.shape, np.uint8)
    mask[40:44, 50:55] = 0
    lem = gpu.LandscapeModel(z, 30.0, mask=mask, timestep_years=800.0, erodibility=3e-5, m_exponent=0.5,
                             n_exponent=1.0, uplift_rate=1e-3, hillslope_diffusivity=0.01,
                             nonlinear_hillslopes=True, critical_hillslope_slope=0.8, boundary_conditions="nnon")
    zb = lem.elevation
    st = lem.step()
    za = lem.elevation
    v = np.sort(np.abs(za - zb)[mask != 0].astype(np.float32))
    p = v[int(np.floor(0.999 * (v.size - 1)))]
    assert np.float32(st["p999_local_elevation_change"]) == p
    assert st["max_local_elevation_change"] == np.abs(za - zb)[mask != 0].max()
