"""Batched ensemble (lsdtt_gpu_ensemble_*): instances of uneven sizes and
parameters stepped concurrently must give exactly what each instance gives
alone (same kernels, deterministic reductions), and match the oracle."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import lsdtt3_amd
    assert lsdtt3_amd.gpu_available()
    return lsdtt3_amd


def models():
    from lsdtt3_amd.synthetic import lem_surface
    shapes = [(40, 52), (120, 96), (64, 64), (33, 150), (90, 70), (150, 41)]
    out = []
    for i, (nr, nc) in enumerate(shapes):
        rng = np.random.default_rng(100 + i)
        nonlinear = i % 2 == 1
        out.append(dict(dem=lem_surface(nr, nc, dx=30.0, seed=60 + i, max_slope=0.6), cell_size=30.0,
                        timestep_years=500.0, erodibility=float(np.exp(rng.uniform(np.log(1e-6), np.log(1e-4)))),
                        m_exponent=float(rng.uniform(0.3, 0.6)), n_exponent=1.0 if i % 3 else 1.5,
                        uplift_rate=float(np.exp(rng.uniform(np.log(1e-4), np.log(1e-3)))),
                        hillslope_diffusivity=float(np.exp(rng.uniform(np.log(1e-3), np.log(1e-1)))),
                        nonlinear_hillslopes=nonlinear, critical_hillslope_slope=0.8,
                        boundary_conditions=["nnon", "oooo", "popo", "nnnn"][i % 4]))
    return out


def test_gpu_ensemble_equals_single_models(gpu):
    print("[gpu ensemble] 6 instances x 3 steps", flush=True)
    ms = models()
    ens = gpu.Ensemble(ms, workers=4)
    stats, sec = ens.run(3)
    assert sec > 0 and len(stats) == len(ms)
    for i, m in enumerate(ms):
        kw = dict(m)
        z = kw.pop("dem")
        dx = kw.pop("cell_size")
        lem = gpu.LandscapeModel(z, dx, **kw)
        for _ in range(3):
            st = lem.step()
        assert np.array_equal(ens.elevation(i), lem.elevation), i
        assert stats[i]["cells_eroded"] == st["cells_eroded"]
        assert stats[i]["newton_iterations"] == st["newton_iterations"]


def test_gpu_ensemble_instance_matches_oracle(gpu):
    ms = models()[:2]
    ens = gpu.Ensemble(ms, workers=2)
    ens.run(1)
    m = ms[0]
    mask = np.ones(m["dem"].shape, np.uint8)
    p = orc.lem_params(500.0, m["erodibility"], m["m_exponent"], m["n_exponent"], m["uplift_rate"],
                       m["hillslope_diffusivity"], Sc=0.8, nonlinear=m["nonlinear_hillslopes"],
                       bc=m["boundary_conditions"])
    zo, so, rc = orc.lem_step(m["dem"], mask, 30.0, p)
    assert rc == 0
    tol = max(4 * 500.0 * max(so["hillslope_residual"], 1e-12), 1e-8)
    assert np.abs(ens.elevation(0) - zo).max() <= max(tol, 1e-6)


def test_gpu_ensemble_rejects_bad_options(gpu):
    from lsdtt3_amd._native import GpuError
    ms = models()[:2]
    ms[1]["drainage_area_kernel_size"] = 2
    with pytest.raises(GpuError):
        gpu.Ensemble(ms)
