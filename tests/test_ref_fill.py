"""The oracle's epsilon-free fill against the reference's OWN fill, compiled
from its sources (oracle/ref_fill.cpp, `make -C oracle ref`):
lsdtt::fill_pits(raster, 0, boundary_conditions, n_threads), the call that
run_lem_step makes (src/cpu/lsdtt_lem_cpu.cpp:826), and
fill_pits_priority_flood (lsdtt_raster_flow_preprocessing_cpu.cpp:893-920).
Bit-exact float32 on every valid cell, including exactly flat plateaus,
interior nodata, periodic edges and degenerate shapes.  Skipped when the
reference library was not built (the reference tree is absent)."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.skipif(not orc.ref_fill_available(),
                                reason="reference fill library not built (no reference tree)")


def same(a, b, valid):
    return np.array_equal(a[valid].view(np.uint32), b[valid].view(np.uint32))


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("bc", ["oooo", "nnnn", "popo", "opop", "onnn", "nnno"])
def test_oracle_fill_equals_reference_fill(seed, bc):
    from lsdtt3_amd.synthetic import routing_surface
    z = routing_surface(301, 203, seed=seed)
    mask = np.ones(z.shape, np.uint8)
    if seed >= 2:
        rng = np.random.default_rng(seed + 100)
        mask[rng.random(z.shape) < 0.01] = 0
        z[5:9, 5:9] = -9999.0
    valid = (mask == 1) & (z != -9999.0)
    a = orc.fill(z, mask, -9999.0, bc)
    b = orc.ref_fill(z, mask, 10.0, bc, nodata=-9999.0)
    assert same(a, b, valid)
    assert (a[valid] > z[valid]).sum() > 100      # depressions were raised
    # the plain priority-flood entry point agrees where no edge is periodic
    if "p" not in bc:
        assert same(orc.ref_fill(z, mask, 10.0, bc, nodata=-9999.0, mode=1), b, valid)


@pytest.mark.parametrize("shape", [(1, 1), (1, 9), (9, 1), (2, 2), (3, 3), (4, 7), (65, 3)])
@pytest.mark.parametrize("bc", ["oooo", "popo", "opop"])
def test_small_shapes_equal_reference(shape, bc):
    rng = np.random.default_rng(sum(shape) + len(bc))
    z = rng.integers(0, 6, shape).astype(np.float32)     # many exact ties
    mask = np.ones(shape, np.uint8)
    try:
        b = orc.ref_fill(z, mask, 1.0, bc)
    except orc.OracleError:
        with pytest.raises(orc.OracleError):
            orc.fill(z, mask, -9999.0, bc)
        return
    assert same(orc.fill(z, mask, -9999.0, bc), b, mask == 1)


@pytest.mark.parametrize("seed", range(8))
def test_random_integer_terrain_equals_reference(seed):
    """integer elevations on random masks: every fill level is a tie"""
    rng = np.random.default_rng(seed)
    nr, nc = (int(v) for v in rng.integers(5, 60, 2))
    z = rng.integers(0, 4, (nr, nc)).astype(np.float32)
    mask = (rng.random((nr, nc)) > 0.1).astype(np.uint8)
    bc = ["oooo", "popo", "opop", "nnnn"][seed % 4]
    b = orc.ref_fill(z, mask, 1.0, bc)
    assert same(orc.fill(z, mask, -9999.0, bc), b, mask == 1)


def test_both_periodic_refused_by_both():
    z = np.zeros((8, 8), np.float32)
    mask = np.ones(z.shape, np.uint8)
    with pytest.raises(orc.OracleError):
        orc.ref_fill(z, mask, 1.0, "pppp")
    with pytest.raises(orc.OracleError):
        orc.fill(z, mask, -9999.0, "pppp")


def test_routing_workload_1024_equals_reference():
    """the routing workload's surface at 1024^2 (1 % closed bowls, 5 % exact
    plateaus); the 4096^2 case runs beside the GPU fill"""
    from lsdtt3_amd.synthetic import routing_surface
    z = routing_surface(1024, 1024, seed=7)
    mask = np.ones(z.shape, np.uint8)
    assert same(orc.fill(z, mask), orc.ref_fill(z, mask, 10.0, "oooo"), mask == 1)
