"""GPU timestep control (lsdtt_gpu_lem_median_interior_gradient, _diagnose,
_timestep_controller_*, _convergence_monitor_*) against the oracle
(oracle/orc_timestep.c): the device median of the gradients is exact (the
same FP64 expressions, an exact selection), so the controller's dt sequence
driven by real LEM steps equals the oracle's driven by the same statistics."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import lsdtt3_amd
    assert lsdtt3_amd.gpu_available()
    return lsdtt3_amd


@pytest.mark.parametrize("shape,holes", [((60, 60), 0), ((97, 131), 40), ((3, 9), 0), ((2, 50), 0), ((513, 257), 300)])
def test_gpu_median_interior_gradient_exact(gpu, shape, holes):
    print(f"[gpu timestep] median {shape}", flush=True)
    rng = np.random.default_rng(shape[0])
    z = rng.normal(size=shape).cumsum(0) * 3.0 + rng.normal(size=shape)
    mask = np.ones(shape, np.uint8)
    for _ in range(holes):
        mask[rng.integers(0, shape[0]), rng.integers(0, shape[1])] = 0
    lem = gpu.LandscapeModel(z, 7.5, mask=mask, hillslope_model="none")
    assert lem.median_interior_gradient() == orc.median_interior_gradient(z, mask, 7.5)


def test_gpu_median_of_a_ramp_and_a_flat(gpu):
    r = np.arange(60, dtype=np.float64)[:, None].repeat(60, 1)
    lem = gpu.LandscapeModel((0.6 * 5.0 * r).astype(np.float32).astype(np.float64), 5.0, hillslope_model="none")
    assert abs(lem.median_interior_gradient() - 0.6) < 1e-6
    d = lem.diagnose_timestep(4e-4, 20000.0)
    assert d["too_large"] and abs(d["ratio"] - 2.676) < 0.05 and "Suggested timestep" in d["warning"]
    assert not lem.diagnose_timestep(4e-4, d["suggested_dt_yr"])["too_large"]
    flat = gpu.LandscapeModel(np.zeros((60, 60)), 5.0, hillslope_model="none")
    f = flat.diagnose_timestep(4e-4, 1e6)
    assert f["reference_slope"] == 0.0 and not f["too_large"] and f["warning"] == ""


def test_gpu_adaptive_run_matches_oracle(gpu):
    """a coupled run whose dt follows the controller: the device controller
    (gradient refreshed from the resident surface every 3 steps) gives the
    same dt sequence as the oracle controller fed the same max|dz| and the
    same surfaces"""
    print("[gpu timestep] adaptive run", flush=True)
    from lsdtt3_amd.synthetic import lem_surface
    z = lem_surface(80, 90, dx=30.0, seed=8, max_slope=0.6)
    kw = dict(timestep_years=2000.0, erodibility=2e-5, m_exponent=0.5, n_exponent=1.0, uplift_rate=2e-3,
              hillslope_diffusivity=0.01, nonlinear_hillslopes=True, critical_hillslope_slope=0.8,
              boundary_conditions="nnon")
    lem = gpu.LandscapeModel(z, 30.0, **kw)
    tc = gpu.TimestepController(2000.0, slope_refresh_interval=3)
    oc = orc.TimestepController(2000.0, slope_refresh_interval=3)
    mon = gpu.ConvergenceMonitor(2000.0, dz_tolerance_m=1e-9)
    om = orc.ConvergenceMonitor(2000.0, dz_tolerance_m=1e-9)
    mask = np.ones(z.shape, np.uint8)
    dts = []
    for _ in range(12):
        dt_used = lem.timestep
        st = lem.step()
        zs = lem.elevation
        dt = tc.update(lem, st["max_local_elevation_change"])
        odt = oc.update(st["max_local_elevation_change"], zs, mask, 30.0)
        assert dt == odt
        assert mon.update(st["max_local_elevation_change"], dt_used) == \
            ("running", "converged", "stalled")[om.update(st["max_local_elevation_change"], dt_used)]
        lem.timestep = dt
        dts.append(dt)
    s = tc.state()
    assert s["n_shrinks"] == oc.c.n_shrinks and s["n_grows"] == oc.c.n_grows and s["n_changes"] == oc.c.n_changes
    assert s["reference_slope"] == oc.c.slope
    assert min(dts) < 2000.0  # U dt = 4 m per step overshoots ~one cell's relief


def test_gpu_controller_errors(gpu):
    from lsdtt3_amd._native import GpuError
    with pytest.raises(GpuError):
        gpu.TimestepController(0.0)
    lem = gpu.LandscapeModel(np.zeros((8, 8)), 1.0, hillslope_model="none")
    with pytest.raises(GpuError):
        lem.timestep = -1.0
