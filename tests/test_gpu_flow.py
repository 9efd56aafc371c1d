"""GPU D8 flow object (through the C ABI) against the reference's golden
pins and against the oracle on seeded DEMs: bit-exact for every integer
output and for the float area raster."""
import os

import numpy as np
import pytest

from d8_fixtures import CONFIGS, COMPONENTS, SPECS, Fixture, hash_flow, load_pins, weights_for
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

PINS = load_pins(os.path.join(os.path.dirname(__file__), "golden", "d8_identity_pins.txt"))


@pytest.fixture(scope="module")
def gpu():
    import lsdtt3_amd
    assert lsdtt3_amd.gpu_available()
    return lsdtt3_amd


def gpu_row(gpu, fx, cfg):
    keep_topo, sv = CONFIGS[cfg]
    fi = gpu.FlowInfo(fx.dem, fx.mask, fx.dx, fx.bc, fx.nodata_is_outlet, fx.pits,
                      build_topological_order=keep_topo, build_s_vector=sv,
                      forced_outlet_flats=fx.forced)
    flow = dict(receiver=fi.receivers, dir=fi.receiver_directions, donors=fi.donor_counts,
                outlets=fi.outlets(), topo=fi.topological_order() if keep_topo else None,
                s_vector=fi.s_vector if sv else None, s_start=fi.s_vector_start if sv else None,
                s_size=fi.s_vector_size if sv else None)
    m = fx.mask.ravel()
    if sv:
        area, amask = fi.get_drainage_area_raster(fx.nodata)
    else:
        acc = fi.accumulate()
        area = np.where(m != 0, acc.astype(np.float32), np.float32(fx.nodata))
        amask = (m != 0).astype(np.uint8)
    weighted = fi.accumulate(weights_for(fx.nr * fx.nc)) if keep_topo else None
    return hash_flow(fx, cfg, flow, area.ravel(), amask.ravel(), weighted)


@pytest.mark.parametrize("spec", SPECS, ids=[s[0] for s in SPECS])
@pytest.mark.parametrize("cfg", list(CONFIGS))
def test_gpu_matches_reference_pins(gpu, spec, cfg):
    print(f"[gpu pins] {spec[0]}/{cfg}", flush=True)
    fx = Fixture(spec)
    got = gpu_row(gpu, fx, cfg)
    want = PINS[f"{spec[0]}/{cfg}"]
    bad = [f"{COMPONENTS[k]}: got {got[k]:#018x} want {want[k]:#018x}" for k in range(8) if got[k] != want[k]]
    assert not bad, "; ".join(bad)


def compare(gpu, z, mask, dx, bc, pits=True, nodata_outlet=True):
    fi = gpu.FlowInfo(z, mask, dx, bc, nodata_outlet, pits)
    ref = orc.flow_build(z, mask, dx, bc, nodata_outlet, pits)
    assert np.array_equal(fi.receivers, ref["receiver"])
    assert np.array_equal(fi.receiver_directions, ref["dir"])
    assert np.array_equal(fi.donor_counts, ref["donors"])
    assert np.array_equal(fi.outlets(), ref["outlets"])
    assert np.array_equal(fi.topological_order(), ref["topo"])
    assert np.array_equal(fi.s_vector, ref["s_vector"])
    assert np.array_equal(fi.s_vector_start, ref["s_start"])
    assert np.array_equal(fi.s_vector_size, ref["s_size"])
    m = np.asarray(mask).ravel()
    acc_ref = orc.accumulate(m, ref["receiver"], ref["topo"], dx)
    assert np.array_equal(fi.accumulate(), acc_ref)
    return fi


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("bc", ["oooo", "nnon", "popo", "bnbn"])
def test_gpu_flow_random_f32(gpu, seed, bc):
    print(f"[gpu flow f32] seed={seed} bc={bc}", flush=True)
    from lsdtt3_amd.synthetic import routing_surface
    z = routing_surface(193, 257, seed=seed)
    rng = np.random.default_rng(seed)
    mask = (rng.random(z.shape) > 0.02).astype(np.uint8)
    compare(gpu, z, mask, 10.0, bc)


@pytest.mark.parametrize("seed", [3, 4])
def test_gpu_flow_random_f64_lem_surface(gpu, seed):
    print(f"[gpu flow f64] seed={seed}", flush=True)
    from lsdtt3_amd.synthetic import lem_surface
    z = lem_surface(160, 224, dx=30.0, seed=seed)
    mask = np.ones(z.shape, np.uint8)
    compare(gpu, z, mask, 30.0, "nnon")


def near_tie_surface(nr, nc, seed, dtype):
    """few distinct heights (exact ties between drops) with a third of the
    cells moved by 1-3 ulps, so drops differ by less than the rounding of
    their quotients: the first-of-equal-slopes rule decides"""
    rng = np.random.default_rng(seed)
    z = (1000.0 + 0.25 * rng.integers(0, 6, (nr, nc))).astype(dtype)
    moved = rng.random((nr, nc)) < 0.33
    steps = rng.integers(1, 4, (nr, nc))
    up = rng.random((nr, nc)) < 0.5
    for k in range(1, 4):
        sel = moved & (steps >= k)
        z[sel] = np.nextafter(z[sel], np.where(up[sel], np.inf, -np.inf).astype(dtype))
    return z


def collapse_surface(nr, nc, seed):
    """a checkerboard of cells near 1000 and near 500, each a few ulps off:
    the drops from a high cell to its low neighbours differ by an ulp of
    themselves, so different drops can give the same rounded slope"""
    rng = np.random.default_rng(seed)
    r, c = np.mgrid[0:nr, 0:nc]
    high = (r + c) % 2 == 0
    z = np.where(high, 1000.0 + rng.integers(0, 3, (nr, nc)) * np.spacing(1000.0),
                 500.0 + rng.integers(0, 4, (nr, nc)) * np.spacing(500.0))
    return z.astype(np.float64)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("dx", [30.0, 7.0])
def test_gpu_flow_near_ties(gpu, dtype, dx):
    print(f"[gpu flow ties] {np.dtype(dtype).name} dx={dx}", flush=True)
    z = near_tie_surface(150, 170, int(dx), dtype)
    mask = np.ones(z.shape, np.uint8)
    compare(gpu, z, mask, dx, "oooo")


@pytest.mark.parametrize("dx", [30.0, 7.0, 3.0])
def test_gpu_flow_equal_rounded_slopes(gpu, dx):
    """different drops whose quotients by the step length round to the same
    double: the first direction of the largest rounded slope wins"""
    print(f"[gpu flow collapse] dx={dx}", flush=True)
    z = collapse_surface(120, 130, int(dx))
    mask = np.ones(z.shape, np.uint8)
    compare(gpu, z, mask, dx, "oooo")


def test_gpu_flow_long_chain(gpu):
    """a single 1-D channel of 20000 cells: deep levels run in the
    single-workgroup tail"""
    print("[gpu flow chain]", flush=True)
    z = np.arange(20000, dtype=np.float32)[None, :] * 0.5
    mask = np.ones(z.shape, np.uint8)
    fi = compare(gpu, z, mask, 10.0, "nnno", pits=False)
    assert fi.max_depth == 19999


def test_gpu_flow_long_chain_wide_keys(gpu):
    """a 1-D channel of 100000 cells: the Kahn keys need 17 + 17 > 32 bits,
    so the topological order sorts 64-bit keys; the depth pointer jumping
    needs more rounds than its first batch"""
    print("[gpu flow chain 100000]", flush=True)
    z = np.arange(100000, dtype=np.float32)[None, :] * 0.5
    mask = np.ones(z.shape, np.uint8)
    fi = compare(gpu, z, mask, 10.0, "nnno", pits=False)
    assert fi.max_depth == 99999


def test_gpu_flow_unresolved_raises(gpu):
    """a pit without allow_pits_as_outlets throws, naming the first cell"""
    import lsdtt3_amd._native as N
    r, c = np.indices((5, 5))
    z = (np.maximum(abs(r - 2), abs(c - 2)) + 1).astype(np.float32)
    z[2, 2] = 0.0
    z[2, 3] = 0.0      # ties: (2,3) drains nowhere lower either
    with pytest.raises(N.GpuError, match="row 2, col 2"):
        gpu.FlowInfo(z, np.ones_like(z, np.uint8), 10.0, "nnnn")


def test_gpu_flow_bad_forced_outlet(gpu):
    import lsdtt3_amd._native as N
    z = np.zeros((4, 4), np.float32)
    m = np.ones_like(z, np.uint8)
    m[0, 0] = 0
    with pytest.raises(N.GpuError, match="masked"):
        gpu.FlowInfo(z, m, 10.0, forced_outlet_flats=[0])
    with pytest.raises(N.GpuError, match="out of range"):
        gpu.FlowInfo(z, np.ones_like(z, np.uint8), 10.0, forced_outlet_flats=[99])


def test_gpu_flow_all_masked(gpu):
    z = np.zeros((3, 4), np.float32)
    fi = gpu.FlowInfo(z, np.zeros_like(z, np.uint8), 10.0)
    assert fi.n_valid == 0 and fi.n_outlets == 0
    assert np.all(fi.receivers == -1)
