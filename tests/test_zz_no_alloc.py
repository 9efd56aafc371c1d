"""lsdtt_gpu_fill_device_noalloc (through the C ABI) against
lsdtt_gpu_fill_device: bit-identical fills on two calls in a row, the
second on a smaller grid (the fixed scratch keeps the first call's
contents), a third on another shape with nodata, and the size limit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def N():
    import lsdtt3_amd
    from lsdtt3_amd import _native
    assert lsdtt3_amd.gpu_available()
    return _native


def fill_both(N, z, mask, nodata=-9999.0, bc=b"oooo"):
    dz = N.DeviceArray.from_host(z)
    dm = N.DeviceArray.from_host(mask)
    a = N.DeviceArray(z.shape, np.float32)
    b = N.DeviceArray(z.shape, np.float32)
    nr, nc = z.shape
    N.check(N.lib().lsdtt_gpu_fill_device(dz.ptr, dm.ptr, nr, nc, nodata, bc, a.ptr, None))
    N.check(N.lib().lsdtt_gpu_fill_device_noalloc(dz.ptr, dm.ptr, nr, nc, nodata, bc, b.ptr, None))
    return a.to_host(), b.to_host()


def test_gpu_fill_noalloc_matches_fill_device(N):
    from lsdtt3_amd.synthetic import routing_surface
    print("[gpu fill noalloc] 1024^2 then 300x500", flush=True)
    assert N.lib().lsdtt_gpu_fill_noalloc_max_cells() == 1 << 22
    z1 = routing_surface(1024, 1024, seed=3)
    want, got = fill_both(N, z1, np.ones(z1.shape, np.uint8))
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert (got > z1).any()
    # second call, smaller grid: the scratch still holds the first call's keys
    z2 = routing_surface(300, 500, seed=4)
    want, got = fill_both(N, z2, np.ones(z2.shape, np.uint8))
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # another shape, periodic east-west, nodata holes
    print("[gpu fill noalloc] 257x129 nodata, npnp", flush=True)
    rng = np.random.default_rng(5)
    z3 = (rng.random((257, 129)) * 40.0).astype(np.float32)
    z3[rng.random(z3.shape) < 0.02] = -9999.0
    mask = np.ones(z3.shape, np.uint8)
    mask[100:104, 60:70] = 0
    want, got = fill_both(N, z3, mask, bc=b"npnp")
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_gpu_fill_noalloc_rejects_grids_beyond_the_scratch(N):
    from lsdtt3_amd._native import GpuError
    dz = N.DeviceArray((8,), np.float32)
    st = N.lib().lsdtt_gpu_fill_device_noalloc(dz.ptr, dz.ptr, 2049, 2048, -9999.0, b"oooo", dz.ptr, None)
    assert st == 1
    assert "4194304" in N.lib().lsdtt_gpu_last_error().decode()
    with pytest.raises(GpuError):
        N.check(N.lib().lsdtt_gpu_fill_device_noalloc(dz.ptr, dz.ptr, 1, (1 << 22) + 1, -9999.0, b"oooo", dz.ptr,
                                                      None))
