"""Oracle Picard hillslope fallback (oracle/orc_picard.c) against closed
forms and against the oracle Newton solver where both solve the same
equations.

* linear mode: one backward-Euler solve per sub-step, the same linear system
  as the Newton solver's -> equal to 1e-9 m;
* a Fourier mode on a periodic grid decays by the backward-Euler factor
  1 / (1 + dt D (4/h^2) sin^2(pi k / N)) per direction;
* no-flux domain, U = 0: mass is conserved (each Picard solve is
  conservative and the damped update is a convex combination);
* steady Roering 1-D profile (uplift, divide, fixed channel) converges to
  the closed-form slopes with refinement.
The reference's own Picard iterates need Eigen (absent): parity unpinned.
"""
import numpy as np
import pytest

from oracle import oracle as orc


def smooth(nr, nc, seed, amp=5.0):
    rng = np.random.default_rng(seed)
    z = rng.random((nr, nc))
    for _ in range(6):
        z = 0.25 * (np.roll(z, 1, 0) + np.roll(z, -1, 0) + np.roll(z, 1, 1) + np.roll(z, -1, 1))
    return 100.0 + amp * (z - z.mean()) / z.std()


@pytest.mark.parametrize("bc", ["nnon", "popo", "onon"])
def test_linear_picard_equals_linear_newton(bc):
    z = smooth(30, 36, 1)
    mask = np.ones(z.shape, np.uint8)
    mask[10:13, 14:17] = 0
    zp, rp = orc.picard_transient(z, mask, 10.0, bc, orc.picard_params(0.02, nonlinear=False), 500.0)
    prm = orc.hs_params(0.02, nonlinear=False, rel_tol=1e-14, abs_tol=1e-16)
    zn, rn = orc.hillslope_transient(z, mask, 10.0, bc, prm, 500.0, 500.0)
    assert rp["substeps"] == 1 and rp["picard_iters"] == 1
    assert np.abs(zp - zn)[mask != 0].max() <= 1e-9


@pytest.mark.parametrize("k", [1, 3])
def test_linear_picard_fourier_decay(k):
    n, h, D, dt = 32, 10.0, 0.05, 200.0
    x = np.arange(n)
    z = 100.0 + np.cos(2 * np.pi * k * x / n)[None, :].repeat(n, 0)
    mask = np.ones(z.shape, np.uint8)
    zp, _ = orc.picard_transient(z, mask, h, "pppp", orc.picard_params(D, nonlinear=False), dt)
    fac = 1.0 / (1.0 + dt * D * (4.0 / h ** 2) * np.sin(np.pi * k / n) ** 2)
    np.testing.assert_allclose(zp - 100.0, (z - 100.0) * fac, rtol=0, atol=1e-10)


@pytest.mark.parametrize("nonlinear", [False, True])
def test_picard_mass_conservation(nonlinear):
    z = smooth(24, 28, 3, amp=3.0)   # sub-critical: the limiter never fires
    mask = np.ones(z.shape, np.uint8)
    mask[5:8, 5:9] = 0
    prm = orc.picard_params(0.01, Sc=0.8, nonlinear=nonlinear)
    zp, r = orc.picard_transient(z, mask, 10.0, "nnnn", prm, 2000.0)
    m = mask != 0
    assert r["limiter_fired"] == 0
    assert abs(zp[m].sum() - z[m].sum()) <= 1e-9 * abs(z[m].sum())
    assert np.abs(zp - z)[m].max() > 1e-3   # something happened


def test_picard_super_critical_start_ends_sub_critical():
    """slopes ~3 Sc: steps are rejected down to the floor, where the
    projection lowers cells until every face slope is <= r_max Sc"""
    z = smooth(24, 28, 3, amp=8.0)
    mask = np.ones(z.shape, np.uint8)
    zp, r = orc.picard_transient(z, mask, 10.0, "nnnn", orc.picard_params(0.01, Sc=0.8, nonlinear=True), 2000.0)
    assert r["limiter_fired"] == 1 and r["rejected"] > 0
    assert r["max_slope_ratio"] < 1.0


def test_picard_roering_steady_profile():
    """1-D hillslope between a divide (no flux, west) and a fixed channel
    (east) under uplift: at steady state each face carries q = U x, and the
    Roering law q = D S / (1 - (S/Sc)^2) fixes the face slope in closed form"""
    D, U, Sc = 0.01, 3e-5, 0.8
    for n, dx in ((51, 2.0), (101, 1.0)):
        z = np.zeros((3, n))
        mask = np.ones(z.shape, np.uint8)
        prm = orc.picard_params(D, Sc=Sc, nonlinear=True, U=U, picard_tol=1e-9, picard_max_iter=200,
                                r_max=0.9999, dt_init=1e4)
        zz, r = orc.picard_transient(z, mask, dx, "nonn", prm, 1e8)
        row = zz[1]
        x_face = (np.arange(n - 1) + 1.0) * dx
        S = -(row[1:] - row[:-1]) / dx
        want = (Sc ** 2 / (2 * U * x_face)) * (np.sqrt(D ** 2 + 4 * U ** 2 * x_face ** 2 / Sc ** 2) - D)
        assert r["limiter_fired"] == 0
        assert np.abs(S - want).max() <= 1e-8 * want.max()
