"""GPU step outputs (through the C ABI) against the oracle: the six step
statistics, the per-pixel erosion-rate field, the resident elevation's
device address, Newton's restoration mode from a super-critical start, and
the fill workspace under concurrent callers.

Statistics (src/cpu/lsdtt_lem_cpu.cpp:614-722).  Tolerances:
  * hand-computed cases (dyadic inputs): equal;
  * fluvial-only steps: elevations agree to 1e-9 m, so the maxima and |dz|
    statistics agree to 1e-9 m; the means are sums over the grid in another
    order, 1e-9 m; rates are those divided by dt;
  * with the hillslope: each side solves to its residual tolerance, so
    elevations (and the statistics) agree within 4 dt tol (as in
    test_gpu_lem.py); the p99.9 is the exact order statistic of each
    side's own |dz|, so it moves by at most the elevation difference.
"""
import threading

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

STAT_KEYS = ("delta_max_elevation", "delta_mean_elevation", "max_local_elevation_change",
             "p999_local_elevation_change", "mean_erosion_rate", "max_erosion_rate")


@pytest.fixture(scope="module")
def gpu():
    import lsdtt3_amd
    assert lsdtt3_amd.gpu_available()
    return lsdtt3_amd


def test_gpu_uplift_only_statistics_hand_case(gpu):
    print("[gpu stats] uplift only", flush=True)
    z = (10.0 * np.arange(4)[:, None] + np.arange(5)[None, :]).astype(np.float64)
    lem = gpu.LandscapeModel(z, 30.0, timestep_years=8.0, uplift_rate=0.0625, hillslope_model="none",
                             fluvial_method="none", boundary_conditions="nnon", baselevel_fall_rate=0.03125,
                             compute_erosion_rate_field=1)
    st = lem.step()
    want = z.copy()
    want[:3] += 0.5
    want[3] -= 0.25
    assert np.array_equal(lem.elevation, want)
    assert st["n_valid"] == 20
    assert st["delta_max_elevation"] == 0.25
    assert st["delta_mean_elevation"] == 0.3125
    assert st["max_local_elevation_change"] == 0.5
    assert st["p999_local_elevation_change"] == 0.5
    assert st["mean_erosion_rate"] == 5 * 0.75 / 20 / 8.0
    assert st["max_erosion_rate"] == 0.75 / 8.0
    want_er = np.zeros(z.shape, np.float32)
    want_er[3] = 0.09375
    assert np.array_equal(lem.erosion_rate, want_er)


def test_gpu_single_eroding_cell_hand_case(gpu):
    print("[gpu stats] single eroding cell", flush=True)
    z = np.array([[0.0, 10.0]] * 3)
    lem = gpu.LandscapeModel(z, 4.0, timestep_years=16.0, erodibility=0.125, m_exponent=0.5, n_exponent=1.0,
                             uplift_rate=0.03125, hillslope_model="none", boundary_conditions="nnno",
                             compute_erosion_rate_field=1)
    st = lem.step()
    assert np.array_equal(lem.elevation, np.array([[0.0, 5.5]] * 3))
    assert st["cells_eroded"] == 3
    assert st["delta_max_elevation"] == 4.5
    assert st["delta_mean_elevation"] == 2.25
    assert st["max_local_elevation_change"] == 4.5
    assert st["p999_local_elevation_change"] == 4.5
    assert st["mean_erosion_rate"] == 3 * 5.5 / 6 / 16.0
    assert st["max_erosion_rate"] == 5.0 / 16.0
    assert np.array_equal(lem.erosion_rate, np.array([[0.03125, 0.3125]] * 3, np.float32))


CASES = {
    # D8 incision only, base-level fall, masked holes, n = 1.5
    "d8": dict(kw=dict(timestep_years=1000.0, erodibility=2e-5, m_exponent=0.5, n_exponent=1.5, uplift_rate=1e-3,
                       hillslope_model="none", boundary_conditions="nnon", baselevel_fall_rate=2e-4),
               okw=dict(dt=1000.0, K=2e-5, m=0.5, n=1.5, U=1e-3, D=0.0, bc="nnon", hillslope=False,
                        baselevel_fall=2e-4),
               tol=1e-9, holes=True),
    # D8 incision + Roering Newton (config-2 physics), two fixed edges
    "d8_roering": dict(kw=dict(timestep_years=500.0, erodibility=3e-5, m_exponent=0.5, n_exponent=1.5,
                               uplift_rate=5e-4, hillslope_diffusivity=0.005, nonlinear_hillslopes=True,
                               critical_hillslope_slope=0.8, boundary_conditions="onon",
                               baselevel_fall_rate=1e-4),
                       okw=dict(dt=500.0, K=3e-5, m=0.5, n=1.5, U=5e-4, D=0.005, Sc=0.8, nonlinear=True,
                                bc="onon", baselevel_fall=1e-4),
                       tol=None, holes=False),
    # D-infinity incision, no uplift
    "dinf": dict(kw=dict(timestep_years=1000.0, erodibility=2e-5, m_exponent=0.5, n_exponent=1.0,
                         uplift_rate=1e-3, enable_uplift=0, hillslope_model="none", fluvial_method="dinf",
                         boundary_conditions="oooo"),
                 okw=dict(dt=1000.0, K=2e-5, m=0.5, n=1.0, U=1e-3, D=0.0, bc="oooo", hillslope=False,
                          uplift=False, fluvial_dinf=True),
                 tol=1e-9, holes=False),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_gpu_step_statistics_match_oracle(gpu, case):
    """all six statistics and the erosion-rate field over three steps"""
    print(f"[gpu stats vs oracle] {case}", flush=True)
    from lsdtt3_amd.synthetic import lem_surface
    c = CASES[case]
    z = lem_surface(90, 104, dx=30.0, seed=61, max_slope=0.6)
    mask = np.ones(z.shape, np.uint8)
    if c["holes"]:
        mask[np.random.default_rng(5).random(z.shape) < 0.02] = 0
    lem = gpu.LandscapeModel(z, 30.0, mask=mask, compute_erosion_rate_field=1, **c["kw"])
    p = orc.lem_params(**c["okw"])
    zo = z.copy()
    er_o = np.empty(z.shape, np.float32)
    valid = mask != 0
    for step in range(3):
        sg = lem.step()
        zo, so, rc = orc.lem_step(zo, mask, 30.0, p, erosion_rate=er_o)
        assert rc == 0
        tol = c["tol"]
        if tol is None:
            assert sg["hillslope_converged"] == 1
            tol = max(4 * 500.0 * sg["hillslope_tolerance"], 1e-8)
        zg = lem.elevation
        assert np.abs(zg - zo)[valid].max() <= tol, (step, np.abs(zg - zo)[valid].max())
        assert sg["n_valid"] == so["n_valid"] == valid.sum()
        dt = c["okw"]["dt"]
        for k in STAT_KEYS:
            scale = 1.0 / dt if "rate" in k else 1.0
            # p99.9 is a float |dz|: allow its rounding (2 ulp) besides tol
            rel = 2.4e-7 if k.startswith("p999") else 1e-12
            assert abs(sg[k] - so[k]) <= 2 * tol * scale + rel * abs(so[k]), (step, k, sg[k], so[k])
        er_g = lem.erosion_rate
        assert np.all(er_g[~valid] == 0.0)
        assert np.abs(er_g.astype(np.float64) - er_o)[valid].max() <= 2 * tol / dt + 1e-6 * np.abs(er_o).max()
        # the field and the statistics are the same erosion
        pos = np.clip(er_g[valid].astype(np.float64), 0.0, None)
        assert abs(pos.sum() / valid.sum() - sg["mean_erosion_rate"]) <= 1e-6 * max(sg["mean_erosion_rate"], 1e-30)
        assert abs(pos.max() - sg["max_erosion_rate"]) <= 1e-6 * max(sg["max_erosion_rate"], 1e-30)


def test_gpu_erosion_rate_field_errors(gpu):
    from lsdtt3_amd._native import GpuError
    z = np.random.default_rng(0).random((20, 24)) * 10 + 100
    lem = gpu.LandscapeModel(z, 10.0, hillslope_model="none")
    lem.step()
    with pytest.raises(GpuError):
        lem.erosion_rate   # not enabled
    lem2 = gpu.LandscapeModel(z, 10.0, hillslope_model="none", compute_erosion_rate_field=1)
    with pytest.raises(GpuError):
        lem2.erosion_rate  # no step yet


def test_gpu_elevation_device_pointer_is_stable(gpu):
    """the device address of the resident elevation stays the state's across
    Roering steps (whose Newton iterates live in other buffers): a surface
    written through an address fetched before a step is what the next step
    reads"""
    print("[gpu lem] stable elevation pointer", flush=True)
    from lsdtt3_amd import _native as N
    from lsdtt3_amd.synthetic import lem_surface
    za = lem_surface(70, 86, dx=30.0, seed=71, max_slope=0.6)
    zb = lem_surface(70, 86, dx=30.0, seed=72, max_slope=0.6)
    kw = dict(timestep_years=500.0, erodibility=3e-5, m_exponent=0.5, n_exponent=1.5, uplift_rate=5e-4,
              hillslope_diffusivity=0.005, nonlinear_hillslopes=True, critical_hillslope_slope=0.8,
              boundary_conditions="nnon")
    lem = gpu.LandscapeModel(za, 30.0, **kw)
    ptr0 = lem.elevation_device_ptr()
    st = lem.step()
    assert st["newton_iterations"] >= 1 and st["hillslope_method"] == 0
    assert lem.elevation_device_ptr() == ptr0
    # the address holds the step's result
    got = np.empty(za.shape, np.float64)
    N.check(N.lib().lsdtt_gpu_memcpy(N.ptr(got), ptr0, got.nbytes, 2))
    assert np.array_equal(got, lem.elevation)
    # a surface written through the old address is the next step's input
    zb_c = np.ascontiguousarray(zb)
    N.check(N.lib().lsdtt_gpu_memcpy(ptr0, N.ptr(zb_c), zb_c.nbytes, 1))
    lem.step()
    ref = gpu.LandscapeModel(zb, 30.0, **kw)
    ref.step()
    assert np.array_equal(lem.elevation, ref.elevation)


def test_gpu_roering_restoration_mode_matches_oracle(gpu):
    """a super-critical start (largest centred slope above Sc): the first
    Newton iterations run the line search in restoration mode, accepting
    trials that lower the slope ratio (lsdtt_hillslope_newton_cpu.cpp:
    553-633); the converged step agrees with the oracle's within
    4 dt tol"""
    print("[gpu lem] restoration mode", flush=True)
    from lsdtt3_amd.synthetic import lem_surface, max_centred_slope
    z = lem_surface(64, 72, dx=30.0, seed=81, max_slope=1.0, n_pits=0)
    assert max_centred_slope(z, 30.0) > 0.8 * np.sqrt(0.9999)
    kw = dict(timestep_years=500.0, erodibility=3e-5, m_exponent=0.5, n_exponent=1.5, uplift_rate=5e-4,
              hillslope_diffusivity=0.005, nonlinear_hillslopes=True, critical_hillslope_slope=0.8,
              boundary_conditions="nnon")
    okw = dict(dt=500.0, K=3e-5, m=0.5, n=1.5, U=5e-4, D=0.005, Sc=0.8, nonlinear=True, bc="nnon")
    mask = np.ones(z.shape, np.uint8)
    lem = gpu.LandscapeModel(z, 30.0, **kw)
    sg = lem.step()
    zo, so, rc = orc.lem_step(z, mask, 30.0, orc.lem_params(**okw))
    assert rc == 0
    assert sg["newton_restoration_iterations"] >= 1
    assert sg["hillslope_method"] == so["hillslope_method"] == 0
    assert sg["hillslope_converged"] == 1
    assert sg["max_slope_ratio"] <= np.sqrt(0.9999)
    tol = max(4 * 500.0 * sg["hillslope_tolerance"], 1e-8)
    assert np.abs(lem.elevation - zo).max() <= tol, (np.abs(lem.elevation - zo).max(), tol)


def test_gpu_fill_device_concurrent_callers(gpu):
    """lsdtt_gpu_fill_device from two host threads at once on different
    grids: each result equals the oracle's fill (the workspace is shared
    one call at a time)"""
    print("[gpu fill] concurrent callers", flush=True)
    from lsdtt3_amd import _native as N
    from lsdtt3_amd.synthetic import routing_surface
    grids = [routing_surface(200, 240, seed=91), routing_surface(150, 330, seed=92)]
    want = [orc.fill(z, np.ones(z.shape, np.uint8)) for z in grids]
    errors = []

    def worker(k):
        try:
            z = grids[k]
            dz = N.DeviceArray.from_host(z)
            dm = N.DeviceArray.from_host(np.ones(z.shape, np.uint8))
            out = N.DeviceArray(z.shape, np.float32)
            for _ in range(3):
                N.check(N.lib().lsdtt_gpu_fill_device(dz.ptr, dm.ptr, z.shape[0], z.shape[1], -9999.0, b"oooo",
                                                      out.ptr, None))
                got = out.to_host()
                if not np.array_equal(got.view(np.uint32), want[k].view(np.uint32)):
                    errors.append(k)
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(repr(e))

    th = [threading.Thread(target=worker, args=(k,)) for k in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
