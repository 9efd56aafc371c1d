"""GPU epsilon-free fill against the oracle: bit-exact float32."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import lsdtt3_amd
    assert lsdtt3_amd.gpu_available()
    return lsdtt3_amd


def same(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("bc", ["oooo", "popo", "opop", "nnnn"])
def test_gpu_fill_matches_oracle(gpu, seed, bc):
    print(f"[gpu fill] seed={seed} bc={bc}", flush=True)
    from lsdtt3_amd.synthetic import routing_surface
    z = routing_surface(301, 203, seed=seed)
    rng = np.random.default_rng(seed + 100)
    mask = np.ones(z.shape, np.uint8)
    if seed == 2:
        mask[rng.random(z.shape) < 0.01] = 0
        z[5:9, 5:9] = -9999.0          # nodata by value with mask set
    assert same(gpu.fill_pits(z, mask, -9999.0, bc), orc.fill(z, mask, -9999.0, bc))


@pytest.mark.parametrize("shape", [(1, 1), (1, 7), (7, 1), (2, 2), (3, 3), (65, 1), (130, 129)])
def test_gpu_fill_small_shapes(gpu, shape):
    rng = np.random.default_rng(sum(shape))
    z = rng.integers(0, 9, shape).astype(np.float32)
    mask = np.ones(shape, np.uint8)
    assert same(gpu.fill_pits(z, mask), orc.fill(z, mask))


def test_gpu_fill_nan_nodata(gpu):
    rng = np.random.default_rng(5)
    z = (rng.random((70, 90)) * 20).astype(np.float32)
    z[30:33, 40:44] = np.nan
    mask = np.ones(z.shape, np.uint8)
    got = gpu.fill_pits(z, mask, float("nan"))
    want = orc.fill(z, mask, float("nan"))
    assert np.array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    assert np.array_equal(got[ok], want[ok])


def test_gpu_fill_deep_maze(gpu):
    """a spiral trench: the spill path winds through many tiles"""
    n = 257
    z = np.full((n, n), 50.0, np.float32)
    r0, r1, c0, c1 = 2, n - 3, 2, n - 3
    k = 0
    while r0 < r1 and c0 < c1:
        z[r0, c0:c1 + 1] = 10.0 - 1e-3 * k
        z[r0:r1 + 1, c1] = 10.0 - 1e-3 * k
        z[r1, c0:c1 + 1] = 10.0 - 1e-3 * k
        z[r0 + 2:r1 + 1, c0] = 10.0 - 1e-3 * k
        r0 += 2; r1 -= 2; c0 += 2; c1 -= 2; k += 1
    z[0, 2] = 0.0
    mask = np.ones(z.shape, np.uint8)
    assert same(gpu.fill_pits(z, mask), orc.fill(z, mask))


@pytest.mark.parametrize("bc", ["oooo", "popo"])
def test_gpu_fill_plateaus_and_nested_pits(gpu, bc):
    """heights on a coarse lattice (wide exactly flat areas, ties between
    neighbours everywhere) with nested bowls and masked holes"""
    print(f"[gpu fill] plateaus {bc}", flush=True)
    rng = np.random.default_rng(17)
    n = 300
    r, c = np.mgrid[0:n, 0:n]
    z = np.floor(8.0 * np.sin(r / 23.0) * np.cos(c / 31.0) + rng.integers(0, 2, (n, n))).astype(np.float32)
    for k in range(12):  # nested bowls: a ring wall around a deeper pit
        cr, cc = rng.integers(20, n - 20, 2)
        d = np.hypot(r - cr, c - cc)
        z[(d > 8) & (d < 10)] = 30.0 + k
        z[d < 4] = -20.0 - k
    mask = np.ones(z.shape, np.uint8)
    mask[rng.random(z.shape) < 0.003] = 0
    assert same(gpu.fill_pits(z, mask, -9999.0, bc), orc.fill(z, mask, -9999.0, bc))


@pytest.mark.parametrize("bc", ["oooo", "popo"])
def test_gpu_fill_white_noise(gpu, bc):
    """uniform noise: about one cell in nine is a pit, the most basins and
    merge rounds per cell of any surface"""
    print(f"[gpu fill] white noise {bc}", flush=True)
    z = np.random.default_rng(23).random((511, 517)).astype(np.float32)
    mask = np.ones(z.shape, np.uint8)
    assert same(gpu.fill_pits(z, mask, -9999.0, bc), orc.fill(z, mask, -9999.0, bc))


def test_gpu_fill_pit_staircase(gpu):
    """10000 pits in a row, each spilling over a lower rim into the next:
    the basins merge in one chain 10000 deep"""
    print("[gpu fill] pit staircase", flush=True)
    nc = 20001
    z = np.full((3, nc), 100.0, np.float32)
    col = np.arange(1, nc - 1)
    z[1, 1:nc - 1] = np.where(col % 2 == 0, -1.0 - 1e-3 * (col % 7), 50.0 - 2e-3 * col).astype(np.float32)
    z[1, nc - 1] = 0.0
    mask = np.ones(z.shape, np.uint8)
    want = orc.fill(z, mask)
    # every pit fills to the lower of its two rims, the one downstream
    assert np.array_equal(want[1, 2:nc - 1:2], z[1, 3:nc:2][:want[1, 2:nc - 1:2].size])
    assert same(gpu.fill_pits(z, mask), want)


def test_gpu_fill_1024(gpu):
    print("[gpu fill] 1024^2", flush=True)
    from lsdtt3_amd.synthetic import routing_surface
    z = routing_surface(1024, 1024, seed=7)
    mask = np.ones(z.shape, np.uint8)
    assert same(gpu.fill_pits(z, mask, -9999.0, "oooo"), orc.fill(z, mask, -9999.0, "oooo"))


def test_gpu_fill_rejects_bad_bc(gpu):
    import lsdtt3_amd._native as N
    z = np.zeros((4, 4), np.float32)
    with pytest.raises(N.GpuError):
        gpu.fill_pits(z, np.ones_like(z, np.uint8), boundary_conditions="pppp")
