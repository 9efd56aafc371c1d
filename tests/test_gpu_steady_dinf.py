"""GPU D-infinity analytic steady state (lsdtt_gpu_dinf_steady_state)
against the oracle (orc_steady_state_dinf).

The device accumulates a cell's donors in direction order where the
reference follows its Kahn order, so the areas agree to rounding (1e-12
relative).  Given the device's areas, the oracle's chi sweep performs the
same float operations on every cell (correctly rounded powf), so the
elevations and scalars are compared bit for bit on those areas; with the
oracle's own areas they agree to 1e-5 relative."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import lsdtt3_amd
    assert lsdtt3_amd.gpu_available()
    return lsdtt3_amd


def surface(nr, nc, seed, holes=0, nodata_cells=0, rough=False):
    rng = np.random.default_rng(seed)
    r, c = np.mgrid[0:nr, 0:nc]
    if rough:
        z = rng.normal(size=(nr, nc)).cumsum(0).cumsum(1) * 0.3 + rng.normal(size=(nr, nc)) + 500.0
    else:
        z = 100.0 + 0.3 * r + 0.6 * np.abs(c - nc // 2) + 0.05 * rng.random((nr, nc))
    z = z.astype(np.float32)
    mask = np.ones((nr, nc), np.uint8)
    for _ in range(holes):
        mask[rng.integers(0, nr), rng.integers(0, nc)] = 0
    for _ in range(nodata_cells):
        z[rng.integers(0, nr), rng.integers(0, nc)] = -9999.0
    return z, mask


CASES = [
    dict(shape=(61, 47), bc="oooo", kw=dict(U=2e-4, K=3e-5, m=0.45, n=1.0)),
    dict(shape=(90, 130), bc="onoo", holes=20, nodata=3, kw=dict(U=1e-4, K=1e-4, m=0.5, n=1.5,
                                                                 divide_area_by_dx=False)),
    dict(shape=(40, 300), bc="onnn", kw=dict(U=5e-4, K=2e-5, m=0.4, n=1.2, relief_mode=True,
                                           target_relief=650.0, baselevel=3.0)),
    dict(shape=(77, 77), bc="oooo", rough=True, kw=dict(channel_threshold_area=300.0)),
    dict(shape=(64, 80), bc="onon", kfield=True, kw=dict(U=2e-4, K=3e-5, m=0.5, n=1.0)),
    dict(shape=(96, 64), bc="pnpn", rough=True, kw=dict(m=0.45, n=2.0)),
    dict(shape=(48, 40), bc="pppp", rough=True, kw=dict(m=0.45, n=1.0, relief_mode=True)),
    dict(shape=(257, 33), bc="nonb", holes=4, rough=True, kw=dict(m=0.3, n=1.0)),
    dict(shape=(1, 9), bc="oooo", kw=dict()),
    dict(shape=(9, 1), bc="oooo", kw=dict()),
    dict(shape=(3, 4), bc="pppp", rough=True, kw=dict()),
]


def run(gpu, z, mask, dx, bc, kw, Kf=None):
    kw = dict(kw)
    base = kw.pop("baselevel", 0.0)
    fg = gpu.FlowGraph(z, mask, dx, boundary_conditions=bc, no_data_value=-9999.0)
    zg, rg = fg.fluvial_steady_state(baselevel_elevation=base, erodibility_field=Kf, **kw)
    return fg, zg, rg


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['shape'][0]}x{c['shape'][1]}-{c['bc']}")
def test_gpu_dinf_steady_state_matches_oracle(gpu, case):
    print(f"[gpu dinf steady] {case['shape']} {case['bc']}", flush=True)
    z, mask = surface(*case["shape"], seed=case["shape"][0] + case["shape"][1], holes=case.get("holes", 0),
                      nodata_cells=case.get("nodata", 0), rough=case.get("rough", False))
    dx = 10.0
    Kf = None
    if case.get("kfield"):
        Kf = np.random.default_rng(9).uniform(1e-5, 1e-4, z.shape).astype(np.float32)
    fg, zg, rg = run(gpu, z, mask, dx, case["bc"], case["kw"], Kf)
    # the areas: to rounding of the donor order
    zo, ro = orc.steady_state_dinf(z, mask, dx, bc=case["bc"], K_field=Kf, **case["kw"])
    assert np.allclose(rg["area"].ravel(), ro["area"].ravel(), rtol=1e-12, atol=0)
    valid = ro["valid"].reshape(z.shape) != 0
    assert np.all(zg[~valid] == np.float32(-9999.0))
    assert np.allclose(zg[valid], zo[valid], rtol=1e-5, atol=1e-6 * max(1.0, float(np.abs(zo[valid]).max())))
    # on the device's areas: bit for bit
    ze, re = orc.steady_state_dinf(z, mask, dx, bc=case["bc"], K_field=Kf, area=rg["area"], **case["kw"])
    assert np.array_equal(zg.view(np.uint32), ze.view(np.uint32))
    for k in ("K_used", "max_unit_height", "computed_relief", "erodibility_field_scale",
              "max_unit_height_at_channel_head", "n_channel_pixels"):
        assert rg[k] == re[k], k
    assert rg["chi_c_reported"] == bool(re["chi_c_reported"])


def test_gpu_dinf_steady_state_large(gpu):
    """1024^2 surface of long flow paths across many tiles: bit for bit
    against the oracle on the device's areas, and the steady law holds on
    every single-receiver cell"""
    print("[gpu dinf steady] 1024^2", flush=True)
    from lsdtt3_amd.synthetic import lem_surface
    z = lem_surface(1024, 1024, dx=30.0, seed=3, max_slope=0.6).astype(np.float32)
    mask = np.ones(z.shape, np.uint8)
    fg, zg, rg = run(gpu, z, mask, 30.0, "nnon", dict(U=1e-4, K=2e-5, m=0.45, n=1.0))
    ze, re = orc.steady_state_dinf(z, mask, 30.0, bc="nnon", U=1e-4, K=2e-5, m=0.45, n=1.0, area=rg["area"])
    assert np.array_equal(zg.view(np.uint32), ze.view(np.uint32))
    assert rg["max_unit_height"] == re["max_unit_height"]
    g = re["graph"]
    one = (g["nrec"] == 1)
    rec = g["rec"][:, 0]
    i = np.nonzero(one)[0]
    r = rec[i]
    nc = z.shape[1]
    diag = (np.abs(r // nc - i // nc) == 1) & (np.abs(r % nc - i % nc) == 1)
    dxe = np.where(diag, 30.0 * np.float64(np.float32(np.sqrt(2.0))), 30.0)
    zf = zg.ravel().astype(np.float64)
    a = rg["area"].ravel()[i] / 30.0
    q = (zf[i] - zf[r]) / dxe * a ** np.float64(np.float32(0.45)) * (2e-5 / 1e-4)
    tol = 4.0 * float(np.spacing(np.float32(zg.max()))) / (zf[i] - zf[r]) + 1e-6
    assert i.size > 1000 and np.all(np.abs(q - 1.0) <= tol)


def test_gpu_dinf_steady_state_is_repeatable(gpu):
    """the chi sweep's passes reach one fixed point whatever order the
    tiles run in: repeated calls give the same bits"""
    z, mask = surface(200, 150, 11, rough=True)
    fg, z16, _ = run(gpu, z, mask, 10.0, "oooo", dict(m=0.45, n=1.0))
    for _ in range(2):
        _, again, _ = run(gpu, z, mask, 10.0, "oooo", dict(m=0.45, n=1.0))
        assert np.array_equal(again.view(np.uint32), z16.view(np.uint32))


def test_gpu_dinf_steady_state_rejects_bad_erodibility(gpu):
    from lsdtt3_amd._native import GpuError
    z, mask = surface(20, 20, 1)
    mask[12, 12] = 0
    fg = gpu.FlowGraph(z, mask, 10.0, boundary_conditions="oooo")
    for bad in (0.0, -1e-5, np.nan):
        Kf = np.full(z.shape, 1e-5, np.float32)
        Kf[5, 5] = bad
        with pytest.raises(GpuError, match=r"K must be > 0.*\(5, 5\)"):
            fg.fluvial_steady_state(erodibility_field=Kf)
        Kf[5, 5] = 1e-5
        Kf[12, 12] = bad
        fg.fluvial_steady_state(erodibility_field=Kf)
