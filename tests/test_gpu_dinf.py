"""GPU D-infinity flow graph (lsdtt_gpu_dinf_graph) and the LEM's
threshold-derived channel set, against the oracle (oracle/orc_dinf.c).

  * receivers, weights, slope and bearing: bit-exact against the oracle with
    the correctly rounded atan2 (the device's value; the reference's atan2f
    may sit one ulp beside it, see tests/test_oracle_dinf.py);
  * area: relative 1e-12 (the device adds a cell's donors in direction
    order, the reference in its Kahn order: rounding only), and the sinks
    collect n_valid dx^2;
  * LEM: the channel set of a step equals the oracle's set on the same
    post-incision surface, the step matches the oracle step (tolerance of
    tests/test_gpu_lem.py), channel pixels are Dirichlet, and the coupled
    regime ignores the threshold.
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


def surface(nr, nc, seed, holes=0, nodata_cells=0):
    rng = np.random.default_rng(seed)
    z = (rng.normal(size=(nr, nc)).cumsum(0).cumsum(1) * 0.3 + rng.normal(size=(nr, nc)) + 500.0)
    z = z.astype(np.float32)
    mask = np.ones((nr, nc), np.uint8)
    for _ in range(holes):
        mask[rng.integers(0, nr), rng.integers(0, nc)] = 0
    for _ in range(nodata_cells):
        z[rng.integers(0, nr), rng.integers(0, nc)] = -9999.0
    return z, mask


CASES = [((67, 45), "oooo", 3), ((130, 97), "nnnn", 0), ((64, 96), "pnpn", 5), ((41, 150), "bobo", 2),
         ((48, 40), "pppp", 0), ((3, 4), "pppp", 0), ((1, 9), "oooo", 0), ((257, 33), "nonb", 4)]


@pytest.mark.parametrize("shape,bc,holes", CASES)
def test_gpu_dinf_graph_matches_oracle(gpu, shape, bc, holes):
    print(f"[gpu dinf] graph {shape} {bc}", flush=True)
    z, mask = surface(*shape, seed=shape[0] * 7 + holes, holes=holes, nodata_cells=2)
    dx = 10.0
    fg = gpu.FlowGraph(z, mask, dx, boundary_conditions=bc, no_data_value=-9999.0)
    o = orc.dinf_graph(z, mask, dx, bc=bc, nodata=-9999.0)
    assert np.array_equal(fg.receivers, o["rec"])
    assert np.array_equal(fg.weights.view(np.uint32), o["w"].view(np.uint32))
    assert np.array_equal(fg.slope.view(np.uint32), o["slope"].view(np.uint32))
    assert np.array_equal(fg.bearing.view(np.uint32), o["bearing"].view(np.uint32))
    valid = (mask != 0) & np.isfinite(z) & (z != -9999.0)
    acc, _ = orc.dinf_accumulate(o, valid.astype(np.uint8), dx)
    area = fg.accumulate().ravel()
    assert np.all(area[~valid.ravel()] == 0.0)
    assert np.allclose(area, acc, rtol=1e-12, atol=0)
    sinks = valid.ravel() & (o["nrec"] == 0)
    total = valid.sum() * dx * dx
    assert abs(area[sinks].sum() - total) <= 1e-9 * total


def test_gpu_dinf_area_large(gpu):
    """1024^2 smooth surface: long flow paths cross many tiles"""
    print("[gpu dinf] 1024^2", flush=True)
    from lsdtt3_amd.synthetic import lem_surface
    z = lem_surface(1024, 1024, dx=30.0, seed=3, max_slope=0.6).astype(np.float32)
    mask = np.ones(z.shape, np.uint8)
    fg = gpu.FlowGraph(z, mask, 30.0, boundary_conditions="nnon")
    o = orc.dinf_graph(z, mask, 30.0, bc="nnon")
    assert np.array_equal(fg.receivers, o["rec"])
    acc, _ = orc.dinf_accumulate(o, mask, 30.0)
    assert np.allclose(fg.accumulate().ravel(), acc, rtol=1e-12, atol=0)


def roering(bc="nnon", **kw):
    d = dict(timestep_years=500.0, erodibility=2e-5, m_exponent=0.5, n_exponent=1.0, uplift_rate=5e-4,
             hillslope_diffusivity=0.02, nonlinear_hillslopes=True, critical_hillslope_slope=0.8,
             boundary_conditions=bc, incision_threshold_specific_area=600.0)
    d.update(kw)
    return d


@pytest.mark.parametrize("bc,nexp", [("nnon", 1.0), ("oooo", 1.5), ("popo", 1.0)])
def test_gpu_lem_threshold_channels_match_oracle(gpu, bc, nexp):
    print(f"[gpu dinf] lem threshold channels {bc} n={nexp}", flush=True)
    from lsdtt3_amd.synthetic import lem_surface
    z = lem_surface(96, 112, dx=30.0, seed=41, max_slope=0.6)
    mask = np.ones(z.shape, np.uint8)
    lem = gpu.LandscapeModel(z, 30.0, **roering(bc, n_exponent=nexp))
    zo = z.copy()
    p = orc.lem_params(500.0, 2e-5, 0.5, nexp, 5e-4, 0.02, Sc=0.8, nonlinear=True, bc=bc, threshold=600.0)
    pf = orc.lem_params(500.0, 2e-5, 0.5, nexp, 5e-4, 0.02, Sc=0.8, nonlinear=True, bc=bc, threshold=600.0,
                        hillslope=False, uplift=False)
    bound = 0.0
    for step in range(2):
        # the oracle's channel set on the post-incision surface of this step
        zi, _, _ = orc.lem_step(zo, mask, 30.0, pf, boundary=np.zeros(0, np.int64))
        ch_o, _ = orc.dinf_channel_mask(zi.astype(np.float32), mask, 30.0, 600.0, bc=bc)
        st = lem.step()
        zo, so, rc = orc.lem_step(zo, mask, 30.0, p)
        assert rc == 0
        ch = lem.channel_mask
        assert 0 < ch.sum() < ch.size // 4
        assert np.array_equal(ch, ch_o), step
        # both stop at ||F||inf <= 1e-8 ||F(z^n)||inf; each is within dt ||F|| / (1 - rho)
        # of the root, with the residual each actually reached (the GPU's adaptive
        # Krylov forcing stops nearer the tolerance than the oracle's exact solves);
        # a later step starts from states already that far apart
        bound += 4 * 500.0 * max(so["hillslope_residual"], st["hillslope_residual"], 1e-12)
        tol = max(bound, 1e-8)
        assert np.abs(lem.elevation - zo).max() <= max(tol, 1e-6), step
        assert st["cells_eroded"] == so["cells_eroded"]


def test_gpu_lem_derived_channels_are_dirichlet(gpu):
    """C1: with incision and uplift off, a step leaves the channel pixels
    exactly where they were; C3: the coupled regime is the run without a
    threshold"""
    r, c = np.mgrid[0:60, 0:41]
    z = 100.0 + 0.2 * (59 - r) + 3.0 * np.abs(c - 20) + 0.01 * np.sin(r * 1.3 + c)
    kw = roering("nnon", incision_threshold_specific_area=400.0, hillslope_diffusivity=0.5,
                 timestep_years=200.0)
    kw.update(fluvial_method="none", enable_uplift=0)
    lem = gpu.LandscapeModel(z, 10.0, **kw)
    lem.step()
    ch = lem.channel_mask
    assert 0 < ch.sum() < ch.size // 5
    zn = lem.elevation
    assert np.array_equal(zn[ch != 0], z[ch != 0])
    assert np.abs(zn - z)[ch == 0].max() > 1e-3
    kc = dict(kw, hillslope_coupling_regime=1)
    k0 = dict(kw, incision_threshold_specific_area=0.0)
    a = gpu.LandscapeModel(z, 10.0, **kc)
    b = gpu.LandscapeModel(z, 10.0, **k0)
    a.step()
    b.step()
    assert a.channel_mask.sum() == 0
    assert np.array_equal(a.elevation, b.elevation)


# ---- D-infinity incision in the LEM step ---------------------------------

@pytest.mark.parametrize("nexp,bc,kernel,kfield", [(1.0, "nnon", 1, False), (1.5, "oooo", 1, True),
                                                   (1.0, "popo", 3, False), (1.3, "nobn", 5, False)])
def test_gpu_dinf_fluvial_matches_oracle(gpu, nexp, bc, kernel, kfield):
    """fluvial only: |dz| <= 1e-9 m (the device and host pow may differ by
    an ulp; the area sums differ by rounding, see above)"""
    print(f"[gpu dinf] fluvial n={nexp} {bc} k={kernel}", flush=True)
    from lsdtt3_amd.synthetic import lem_surface
    z = lem_surface(130, 150, dx=30.0, seed=9, max_slope=0.6)
    mask = np.ones(z.shape, np.uint8)
    K = np.random.default_rng(2).uniform(1e-6, 1e-4, z.shape) if kfield else None
    kw = dict(timestep_years=1000.0, erodibility=1e-5, m_exponent=0.5, n_exponent=nexp, uplift_rate=1e-3,
              hillslope_model="none", boundary_conditions=bc, fluvial_method="dinf",
              drainage_area_kernel_size=kernel)
    lem = gpu.LandscapeModel(z, 30.0, mask=mask, erodibility_field=K, **kw)
    p = orc.lem_params(1000.0, 1e-5, 0.5, nexp, 1e-3, 0.0, bc=bc, hillslope=False, fluvial_dinf=True,
                       area_kernel=kernel)
    zo = z.copy()
    for step in range(3):
        sg = lem.step()
        zo, so, rc = orc.lem_step(zo, mask, 30.0, p, K_field=K)
        assert rc == 0
        assert np.abs(lem.elevation - zo).max() <= 1e-9, step
        assert sg["cells_eroded"] == so["cells_eroded"], step


def test_gpu_dinf_lem_full_step_matches_oracle(gpu):
    """D-inf incision, Roering hillslope with threshold-derived channels"""
    from lsdtt3_amd.synthetic import lem_surface
    z = lem_surface(96, 112, dx=30.0, seed=43, max_slope=0.6)
    mask = np.ones(z.shape, np.uint8)
    lem = gpu.LandscapeModel(z, 30.0, **roering("nnon", fluvial_method="dinf"))
    p = orc.lem_params(500.0, 2e-5, 0.5, 1.0, 5e-4, 0.02, Sc=0.8, nonlinear=True, bc="nnon", threshold=600.0,
                       fluvial_dinf=True)
    zo = z.copy()
    bound = 0.0
    for step in range(2):
        st = lem.step()
        zo, so, rc = orc.lem_step(zo, mask, 30.0, p)
        assert rc == 0
        # both stop at ||F||inf <= 1e-8 ||F(z^n)||inf; each is within dt ||F|| / (1 - rho)
        # of the root, with the residual each actually reached (the GPU's adaptive
        # Krylov forcing stops nearer the tolerance than the oracle's exact solves);
        # a later step starts from states already that far apart
        bound += 4 * 500.0 * max(so["hillslope_residual"], st["hillslope_residual"], 1e-12)
        tol = max(bound, 1e-8)
        assert np.abs(lem.elevation - zo).max() <= max(tol, 1e-6), step
        assert st["cells_eroded"] == so["cells_eroded"]


def test_gpu_dinf_fluvial_independent_of_tile_height(gpu):
    from lsdtt3_amd.synthetic import lem_surface
    z = lem_surface(97, 131, dx=30.0, seed=12, max_slope=0.6)
    out = []
    for th in (8, 16, 32):
        lem = gpu.LandscapeModel(z, 30.0, **roering("popo", fluvial_method="dinf", n_exponent=1.5))
        lem.tile_height = th
        lem.run(2)
        out.append((lem.elevation, lem.channel_mask))
    for zt, ct in out[1:]:
        assert np.array_equal(zt, out[0][0])
        assert np.array_equal(ct, out[0][1])
