"""Tile shape of the count and stream-power sweeps (32 columns by 8, 16 or
32 rows) and its autotuner.  The sweeps compute every cell's value as the
same function of its receiver's (or donors') final values whatever the
tiling, so the results must be bit-identical for every height, as the
reference's results are for every thread count
(tests/cpp/test_thread_count_invariance.cpp)."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

HEIGHTS = (8, 16, 32)


@pytest.fixture(scope="module")
def gpu():
    import lsdtt3_amd
    assert lsdtt3_amd.gpu_available()
    return lsdtt3_amd


def roering_kw(bc="nnon"):
    return dict(timestep_years=500.0, erodibility=2e-5, m_exponent=0.5, n_exponent=1.5, uplift_rate=5e-4,
                hillslope_diffusivity=0.005, nonlinear_hillslopes=True, critical_hillslope_slope=0.8,
                boundary_conditions=bc)


@pytest.mark.parametrize("shape,bc", [((150, 170), "nnon"), ((33, 150), "popo"), ((97, 64), "oooo")])
def test_gpu_results_independent_of_tile_height(gpu, shape, bc):
    print(f"[gpu tiles] invariance {shape} {bc}", flush=True)
    from lsdtt3_amd.synthetic import lem_surface
    z = lem_surface(*shape, dx=30.0, seed=17, max_slope=0.6)
    out = {}
    for th in HEIGHTS:
        lem = gpu.LandscapeModel(z, 30.0, **roering_kw(bc))
        lem.tile_height = th
        assert lem.tile_height == th
        sts = [lem.step() for _ in range(3)]
        out[th] = (lem.elevation, sts)
    z16, s16 = out[16]
    for th in (8, 32):
        zt, st = out[th]
        assert np.array_equal(zt, z16), th
        for a, b in zip(st, s16):
            for key in ("cells_eroded", "newton_iterations", "krylov_iterations", "max_local_elevation_change",
                        "p999_local_elevation_change", "hillslope_residual"):
                assert a[key] == b[key], (th, key)


def test_gpu_long_channel_every_tile_height(gpu):
    """a 3000-cell channel whose tile crossings run against the flow needs
    many passes at every height; each height matches the oracle"""
    x = np.arange(3000, dtype=np.float64)
    z = np.empty((3, 3000))
    z[1] = 100.0 + 0.3 * x
    z[0] = z[2] = 5000.0 - 1.0 * x
    kw = dict(timestep_years=100.0, erodibility=1e-4, m_exponent=0.5, n_exponent=1.0, uplift_rate=1e-3,
              hillslope_model="none", boundary_conditions="nnno")
    mask = np.ones(z.shape, np.uint8)
    zo, so, rc = orc.lem_step(z, mask, 10.0, orc.lem_params(100.0, 1e-4, 0.5, 1.0, 1e-3, 0.0, bc="nnno",
                                                             hillslope=False))
    assert rc == 0
    for th in HEIGHTS:
        lem = gpu.LandscapeModel(z, 10.0, mask=mask, **kw)
        lem.tile_height = th
        sg = lem.step()
        assert np.abs(lem.elevation - zo).max() <= 1e-9, th
        assert sg["cells_eroded"] == so["cells_eroded"], th
        assert sg["stream_power_passes"] >= 2


def test_gpu_tile_height_switch_mid_run(gpu):
    """changing the height between steps (new tile buffers, counts kept from
    the previous step) changes nothing"""
    from lsdtt3_amd.synthetic import lem_surface
    z = lem_surface(120, 136, dx=30.0, seed=23, max_slope=0.6)
    a = gpu.LandscapeModel(z, 30.0, **roering_kw())
    b = gpu.LandscapeModel(z, 30.0, **roering_kw())
    for th in (32, 8, 16, 8):
        b.tile_height = th
        a.step()
        b.step()
        assert np.array_equal(a.elevation, b.elevation), th


def test_gpu_autotune_restores_surface(gpu):
    print("[gpu tiles] autotune", flush=True)
    from lsdtt3_amd.synthetic import lem_surface
    z = lem_surface(256, 256, dx=30.0, seed=29, max_slope=0.6)
    lem = gpu.LandscapeModel(z, 30.0, **roering_kw())
    ref = gpu.LandscapeModel(z, 30.0, **roering_kw())
    th, ms = lem.autotune_tiles(steps=2)
    assert th in HEIGHTS and lem.tile_height == th
    assert set(ms) == set(HEIGHTS) and all(v > 0 for v in ms.values())
    # the default height stays unless another is more than 3% faster
    assert th == 16 or ms[th] < 0.97 * ms[16]
    assert np.array_equal(lem.elevation, z)
    for _ in range(2):
        s1, s2 = lem.step(), ref.step()
        assert s1["cells_eroded"] == s2["cells_eroded"]
    assert np.array_equal(lem.elevation, ref.elevation)


def test_gpu_ensemble_autotune(gpu):
    print("[gpu tiles] ensemble autotune", flush=True)
    from lsdtt3_amd.synthetic import lem_surface
    shapes = [(40, 52), (160, 150), (64, 64), (300, 280), (90, 70)]
    ms = [dict(dem=lem_surface(nr, nc, dx=30.0, seed=70 + i, max_slope=0.6), cell_size=30.0, **roering_kw())
          for i, (nr, nc) in enumerate(shapes)]
    ens = gpu.Ensemble(ms)
    best, ms_per = ens.autotune_tiles(steps=1)
    assert best in HEIGHTS and ens.tile_height == best and all(ms_per[h] > 0 for h in HEIGHTS)
    for i, m in enumerate(ms):
        assert np.array_equal(ens.elevation(i), m["dem"]), i
    ens.run(2)
    # the tile height changes the speed only
    for th in HEIGHTS:
        ref = gpu.Ensemble(ms)
        ref.tile_height = th
        ref.run(2)
        for i in range(len(ms)):
            assert np.array_equal(ens.elevation(i), ref.elevation(i)), (th, i)


def test_gpu_tile_height_rejects_other_values(gpu):
    from lsdtt3_amd._native import GpuError
    from lsdtt3_amd.synthetic import lem_surface
    lem = gpu.LandscapeModel(lem_surface(32, 32, dx=30.0, seed=1), 30.0, **roering_kw())
    for bad in (0, 4, 12, 64):
        with pytest.raises(GpuError):
            lem.tile_height = bad
    assert lem.tile_height == 16


@pytest.mark.parametrize("shape,bc", [((150, 170), "nnon"), ((48, 96), "popo")])
def test_gpu_incremental_tile_order_matches_full_rebuild(gpu, shape, bc):
    """From the second step on, only the tiles where a receiver changed (or
    a changed receiver leads in or out) get a new level order, and the
    count passes start from them.  A model whose tile height alternates
    rebuilds every tile's order and runs every tile in every step; both must
    give the same surfaces bit for bit over steps whose receivers change (a
    rough surface under strong incision)"""
    print(f"[gpu tiles] incremental order {shape} {bc}", flush=True)
    from lsdtt3_amd.synthetic import lem_surface
    z = lem_surface(*shape, dx=30.0, seed=23, max_slope=0.6)
    kw = dict(roering_kw(bc), erodibility=3e-4, timestep_years=2000.0)
    inc = gpu.LandscapeModel(z, 30.0, **kw)
    full = gpu.LandscapeModel(z, 30.0, **kw)
    inc.tile_height = 16
    prev = None
    changed_steps = 0
    for k in range(12):
        full.tile_height = 8 if k % 2 else 16  # a height switch rebuilds every tile
        a, b = inc.step(), full.step()
        za, zb = inc.elevation, full.elevation
        assert np.array_equal(za, zb), k
        assert a["cells_eroded"] == b["cells_eroded"] and a["count_passes"] >= 1
        rc, _ = orc.d8_receivers(za, np.ones(shape, np.uint8), 30.0, bc=bc, allow_pits_as_outlets=True)
        if prev is not None and not np.array_equal(rc, prev):
            changed_steps += 1
        prev = rc
    assert changed_steps >= 3  # the receivers did change between steps
