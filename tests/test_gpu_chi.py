"""GPU chi (through the C ABI) against the oracle: bit-exact float chi for
the raster and the thresholded channel lists, on random DEMs (all boundary
codes, nodata holes, non-integer dx^2) and on a fill -> D8 -> chi pipeline
over a seeded routing surface with closed depressions and flat plateaus."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import lsdtt3_amd
    assert lsdtt3_amd.gpu_available()
    return lsdtt3_amd


def both(gpu, z, mask, dx, bc, theta, A0=1.0):
    fi = gpu.FlowInfo(z, mask, dx, bc, allow_pits_as_outlets=True)
    flow = orc.flow_build(z, mask, dx, bc, allow_pits_as_outlets=True)
    cg, mg = fi.get_chi_raster(theta, A0)
    co, mo = orc.chi_raster(flow, mask, dx, m_over_n=theta, A0=A0)
    return fi, flow, cg, mg, co, mo


@pytest.mark.parametrize("bc,holes,dx,theta", [("oooo", 0.0, 10.0, 0.45), ("nnon", 0.05, 30.0, 0.5),
                                               ("popo", 0.0, 30.137, 0.45), ("opop", 0.02, 10.0, 0.3),
                                               ("nnnn", 0.1, 10.0, 0.0), ("bnbn", 0.0, 7.5, 0.6)])
def test_gpu_chi_raster_bit_exact(gpu, bc, holes, dx, theta):
    print(f"[gpu chi raster] bc={bc} holes={holes} dx={dx} theta={theta}", flush=True)
    rng = np.random.default_rng(21)
    z = (rng.random((97, 131)) * 20.0).astype(np.float32)
    mask = np.ones(z.shape, np.uint8)
    if holes:
        mask[rng.random(z.shape) < holes] = 0
    fi, flow, cg, mg, co, mo = both(gpu, z, mask, dx, bc, theta)
    assert np.array_equal(mg, mo)
    assert np.array_equal(cg.view(np.uint32), co.view(np.uint32))


@pytest.mark.parametrize("mp", [1, 25, 1000])
def test_gpu_chi_threshold_matches_oracle(gpu, mp):
    from lsdtt3_amd.synthetic import routing_surface
    print(f"[gpu chi threshold] min_pixels={mp}", flush=True)
    z = routing_surface(160, 200, seed=5)
    mask = np.ones(z.shape, np.uint8)
    zf = gpu.fill_pits(z, mask)
    fi = gpu.FlowInfo(zf, mask, 10.0, "oooo", allow_pits_as_outlets=True)
    basins = fi.get_chi_from_all_outlets_threshold(mp, 0.45, 1.0)
    flow = orc.flow_build(zf, mask, 10.0, "oooo", allow_pits_as_outlets=True)
    t = orc.chi_threshold(flow, mask, 10.0, mp, m_over_n=0.45, A0=1.0)
    assert len(basins) == len(t["basin_start"])
    flats = np.concatenate([b["flats"] for b in basins]) if basins else np.zeros(0, np.int64)
    chi = np.concatenate([b["chi"] for b in basins]) if basins else np.zeros(0, np.float32)
    rec = np.concatenate([b["receiver_flats"] for b in basins]) if basins else np.zeros(0, np.int64)
    assert np.array_equal(flats, t["flats"])
    assert np.array_equal(chi.view(np.uint32), t["chi"].view(np.uint32))
    assert np.array_equal(rec, t["receivers"])
    assert [b["outlet_flat"] for b in basins] == list(t["flats"][t["basin_start"]])
    for b in basins:
        assert b["chi"][0] == 0.0 and b["receiver_flats"][0] == -2


def test_gpu_chi_routing_pipeline_1024(gpu):
    """fill -> D8 (pits as outlets) -> counts -> chi on a 1024^2 routing
    surface (closed bowls, exactly flat plateaus), dx = 10 m, theta 0.45"""
    from lsdtt3_amd.synthetic import routing_surface
    print("[gpu chi pipeline 1024]", flush=True)
    z = routing_surface(1024, 1024, seed=7)
    mask = np.ones(z.shape, np.uint8)
    zf = gpu.fill_pits(z, mask)
    zo = orc.fill(z, mask)
    assert np.array_equal(zf.view(np.uint32), zo.view(np.uint32))
    fi, flow, cg, mg, co, mo = both(gpu, zf, mask, 10.0, "oooo", 0.45)
    assert np.array_equal(cg.view(np.uint32), co.view(np.uint32))
    assert np.array_equal(fi.n_contributing_pixels(), flow["s_size"].astype(np.int64))


def test_gpu_routing_pipeline_4096_bit_exact(gpu):
    """the routing workload at its full size: 4096^2 float32 DEM (1 % closed
    bowls up to 20 m deep, 5 % exactly flat plateaus), dx = 10 m; fill ->
    D8 (pits as outlets) -> donors, outlets, Kahn order, s-vector, int64
    counts, float area -> chi raster and the thresholded channel lists,
    every array bit-exact against the oracle"""
    from lsdtt3_amd.synthetic import routing_surface
    print("[gpu routing pipeline 4096] generating", flush=True)
    z = routing_surface(4096, 4096, seed=7)
    mask = np.ones(z.shape, np.uint8)
    zf = gpu.fill_pits(z, mask)
    zo = orc.fill(z, mask)
    assert np.array_equal(zf.view(np.uint32), zo.view(np.uint32))
    assert (zo > z).sum() > 0.005 * z.size   # the bowls were filled
    if orc.ref_fill_available():     # the reference's own fill, built from its sources
        zr = orc.ref_fill(z, mask, 10.0, "oooo")
        assert np.array_equal(zf.view(np.uint32), zr.view(np.uint32))
        print("[gpu routing pipeline 4096] reference fill agrees", flush=True)
        del zr
    del zo
    print("[gpu routing pipeline 4096] fill ok", flush=True)
    fi = gpu.FlowInfo(zf, mask, 10.0, "oooo", allow_pits_as_outlets=True)
    flow = orc.flow_build(zf, mask, 10.0, "oooo", allow_pits_as_outlets=True)
    assert np.array_equal(fi.receivers, flow["receiver"])
    assert np.array_equal(fi.receiver_directions, flow["dir"])
    assert np.array_equal(fi.donor_counts, flow["donors"])
    assert np.array_equal(fi.outlets(), flow["outlets"])
    assert np.array_equal(fi.topological_order(), flow["topo"])
    assert np.array_equal(fi.s_vector, flow["s_vector"])
    assert np.array_equal(fi.s_vector_start, flow["s_start"])
    assert np.array_equal(fi.s_vector_size, flow["s_size"])
    counts = orc.upstream_counts(mask, flow["receiver"], flow["topo"])
    assert np.array_equal(fi.n_contributing_pixels(), counts)
    area, amask = fi.get_drainage_area_raster()
    assert np.array_equal(area.ravel(), (counts * 100).astype(np.float32))
    print("[gpu routing pipeline 4096] flow object ok", flush=True)
    cg, mg = fi.get_chi_raster(0.45, 1.0)
    co, mo = orc.chi_raster(flow, mask, 10.0, m_over_n=0.45, A0=1.0)
    assert np.array_equal(mg, mo)
    assert np.array_equal(cg.view(np.uint32), co.view(np.uint32))
    # The epsilon-free fill leaves its lakes exactly flat and every flat
    # cell is its own outlet (pits as outlets), so on this surface no basin
    # reaches 1000 pixels: the 1000-pixel lists are empty on both sides, and
    # a 30-pixel threshold gives non-empty lists to compare.
    for mp in (1000, 30):
        basins = fi.get_chi_from_all_outlets_threshold(mp, 0.45, 1.0)
        t = orc.chi_threshold(flow, mask, 10.0, mp, m_over_n=0.45, A0=1.0)
        assert len(basins) == len(t["basin_start"])
        if mp == 30:
            assert len(basins) > 100
        if not basins:
            continue
        assert np.array_equal(np.concatenate([b["flats"] for b in basins]), t["flats"])
        assert np.array_equal(np.concatenate([b["chi"] for b in basins]).view(np.uint32),
                              t["chi"].view(np.uint32))
        assert np.array_equal(np.concatenate([b["receiver_flats"] for b in basins]), t["receivers"])


def test_gpu_chi_errors(gpu):
    from lsdtt3_amd._native import GpuError
    z = np.random.default_rng(1).random((20, 30)).astype(np.float32)
    mask = np.ones(z.shape, np.uint8)
    fi = gpu.FlowInfo(z, mask, 10.0, allow_pits_as_outlets=True)
    with pytest.raises(GpuError) as e:
        fi.get_chi_raster(0.45, 0.0)
    assert e.value.status == 1 and "A_0" in str(e.value)
    with pytest.raises(GpuError) as e:
        fi.get_chi_raster(float("nan"), 1.0)
    assert e.value.status == 1
    with pytest.raises(GpuError) as e:
        fi.get_chi_from_all_outlets_threshold(0)
    assert e.value.status == 1
    fi2 = gpu.FlowInfo(z, mask, 10.0, allow_pits_as_outlets=True, build_s_vector=False)
    with pytest.raises(GpuError) as e:
        fi2.get_chi_raster()
    assert e.value.status == 2 and "s_vector" in str(e.value)
