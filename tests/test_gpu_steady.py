"""GPU analytic fluvial steady state (lsdtt_gpu_flow_steady_state) against
the oracle (oracle/orc_steady.c) with the correctly rounded powf: the same
float operations in the same order along every flow path, so the
elevations and the result scalars are bit-exact."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import lsdtt3_amd
    assert lsdtt3_amd.gpu_available()
    return lsdtt3_amd


def routable(nr, nc, seed, holes=0):
    rng = np.random.default_rng(seed)
    r, c = np.mgrid[0:nr, 0:nc]
    z = (100.0 + 0.5 * r + 0.3 * c + 0.1 * rng.random((nr, nc))).astype(np.float32)
    mask = np.ones((nr, nc), np.uint8)
    for _ in range(holes):
        mask[rng.integers(1, nr - 1), rng.integers(1, nc - 1)] = 0
    return z, mask


CASES = [
    dict(shape=(61, 47), bc="oooo", kw=dict(U=2e-4, K=3e-5, m=0.45, n=1.0)),
    dict(shape=(90, 130), bc="oooo", holes=20, kw=dict(U=1e-4, K=1e-4, m=0.5, n=1.5, divide_area_by_dx=False)),
    dict(shape=(40, 300), bc="onoo", kw=dict(U=5e-4, K=2e-5, m=0.4, n=1.2, relief_mode=True, target_relief=650.0,
                                           baselevel=3.0)),
    dict(shape=(77, 77), bc="oooo", kw=dict(channel_threshold_area=300.0)),
    dict(shape=(64, 80), bc="oooo", kfield=True, kw=dict(U=2e-4, K=3e-5, m=0.5, n=1.0)),
]


@pytest.mark.parametrize("case", CASES)
def test_gpu_steady_state_matches_oracle(gpu, case):
    print(f"[gpu steady] {case['shape']} {case['bc']}", flush=True)
    z, mask = routable(*case["shape"], seed=case["shape"][0], holes=case.get("holes", 0))
    kw = dict(case["kw"])
    Kf = None
    if case.get("kfield"):
        Kf = np.random.default_rng(9).uniform(1e-5, 1e-4, z.shape).astype(np.float32)
    zo, ro = orc.steady_state_d8(z, mask, 10.0, bc=case["bc"], K_field=Kf, **kw)
    base = kw.pop("baselevel", 0.0)
    fi = gpu.FlowInfo(z, mask, 10.0, boundary_conditions=case["bc"], build_s_vector=False)
    zg, rg = fi.fluvial_steady_state(baselevel_elevation=base, erodibility_field=Kf, **kw)
    assert np.array_equal(zg.view(np.uint32), zo.view(np.uint32))
    for k in ("K_used", "max_unit_height", "computed_relief", "erodibility_field_scale",
              "max_unit_height_at_channel_head", "n_channel_pixels"):
        assert rg[k] == ro[k], k
    assert rg["chi_c_reported"] == bool(ro["chi_c_reported"])


def test_gpu_steady_state_large(gpu):
    print("[gpu steady] 1024^2", flush=True)
    z, mask = routable(1024, 1024, 5)
    zo, ro = orc.steady_state_d8(z, mask, 30.0, bc="oooo", U=1e-4, K=2e-5, m=0.45, n=1.0)
    fi = gpu.FlowInfo(z, mask, 30.0, boundary_conditions="oooo", build_s_vector=False)
    zg, rg = fi.fluvial_steady_state(U=1e-4, K=2e-5, m=0.45, n=1.0)
    assert np.array_equal(zg.view(np.uint32), zo.view(np.uint32))
    assert rg["max_unit_height"] == ro["max_unit_height"]


def test_gpu_steady_state_rejects_bad_erodibility(gpu):
    from lsdtt3_amd._native import GpuError
    z, mask = routable(20, 20, 1)
    fi = gpu.FlowInfo(z, mask, 10.0, boundary_conditions="oooo", build_s_vector=False)
    Kf = np.full(z.shape, 1e-5, np.float32)
    Kf[5, 5] = -1.0
    with pytest.raises(GpuError):
        fi.fluvial_steady_state(erodibility_field=Kf)
