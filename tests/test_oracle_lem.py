"""The oracle's LEM pieces against what the mathematics fixes.

Stream power (src/cpu/lsdtt_stream_power_cpu.cpp:159-326): steady state of
the implicit scheme under uplift is U = K A^m S^n exactly (per cell, with S
the slope to the receiver); the n != 1 per-node Newton leaves a root
residual below its stopping tolerance.

Hillslope (src/cpu/lsdtt_hillslope_newton_cpu.cpp): a Fourier mode decays by
1 / (1 + dt D lambda_k) per backward-Euler step under linear diffusion; the
steady divide curvature is -U/D (tests/cpp/test_hillslope_newton.cpp:
278-311); closed domains conserve mass; the Roering 1-D steady profile
S(x) = (Sc^2 / (2 U x)) (sqrt(D^2 + 4 U^2 x^2 / Sc^2) - D)
(tests/cpp/test_hillslope_implicit.cpp:38-39) is reached to O(h^2).
"""
import numpy as np
import pytest

from oracle import oracle as orc


def channel(n, dx, slope0=0.01):
    """a 1-row channel draining west to an open edge"""
    z = (np.arange(n, dtype=np.float64) * dx * slope0)[None, :] + 10.0
    return z, np.ones(z.shape, np.uint8)


@pytest.mark.parametrize("nexp", [1.0, 1.5])
def test_stream_power_steady_state_closed_form(nexp):
    n, dx = 40, 30.0
    K, m, U, dt = 1e-4, 0.5, 1e-3, 2000.0
    z, mask = channel(n, dx)
    p = orc.lem_params(dt, K, m, nexp, U, 0.0, bc="nnno", hillslope=False)
    bf = orc.boundary_flats(mask, "nnno")
    assert list(bf) == [0]
    for _ in range(3000):
        z, st, rc = orc.lem_step(z, mask, dx, p, bf)
        assert rc == 0
    # A = (cells upstream incl. self) * dx^2, a = A / dx; slope to receiver
    cells = np.arange(n, 0, -1)[1:]           # column c has n - c cells upstream
    a = cells * dx * dx / dx
    S = (z[0, 1:] - z[0, :-1]) / dx
    want = (U / (K * a ** m)) ** (1.0 / nexp)
    # the step erodes, then uplifts every cell but the fixed outlet: the
    # segment into the outlet carries the extra U*dt of that split
    want[0] += U * dt / dx
    np.testing.assert_allclose(S, want, rtol=2e-3)


def test_stream_power_newton_root_residual():
    rng = np.random.default_rng(4)
    z = 100.0 + np.cumsum(rng.uniform(0.1, 2.0, 50))[None, :]
    mask = np.ones(z.shape, np.uint8)
    dx, K, m, nexp, dt = 30.0, 1e-4, 0.5, 1.5, 500.0
    rcv, d = orc.d8_receivers(z, mask, dx, "nnno", True, True)
    fl = orc.flow_build(z, mask, dx, "nnno", True, True)
    cnt = orc.upstream_counts(mask, fl["receiver"], fl["topo"])
    area = cnt * dx * dx
    zn, _ = orc.stream_power(z, z, mask, fl["receiver"], fl["dir"], fl["topo"], area, K, m, nexp, dt, dx)
    zo, zn = z.ravel(), zn.ravel()
    tol = 1e-5 * dx
    for f in range(1, 50):
        r = fl["receiver"][f]
        a = area[f] / dx
        f_z = zn[f] - zo[f] + K * a ** m * dt * ((zn[f] - zn[r]) / dx) ** nexp
        fd = 1 + K * a ** m * dt * nexp / dx * ((zn[f] - zn[r]) / dx) ** (nexp - 1)
        assert abs(f_z / fd) <= tol


def fourier_grid(nr, nc, dx, k):
    x = np.arange(nc) * dx
    z = np.cos(2 * np.pi * k * x / (nc * dx))[None, :].repeat(nr, 0) + 100.0
    return z


@pytest.mark.parametrize("k", [1, 3, 7])
def test_linear_hillslope_fourier_decay(k):
    nr, nc, dx, D, dt = 4, 32, 10.0, 0.5, 50.0
    z = fourier_grid(nr, nc, dx, k)
    mask = np.ones(z.shape, np.uint8)
    prm = orc.hs_params(D, nonlinear=False, rel_tol=1e-12, abs_tol=1e-14)
    zz, r = orc.hillslope_newton(z, z, mask, dx, "pppp", prm, 1.0 / dt)
    assert r.converged
    lam = (2 - 2 * np.cos(2 * np.pi * k / nc)) / dx ** 2
    want = (z - 100.0) / (1 + dt * D * lam) + 100.0
    np.testing.assert_allclose(zz, want, rtol=0, atol=1e-11)


def test_linear_hillslope_divide_curvature():
    """steady state under uplift: D z'' = -U between two fixed edges"""
    nr, nc, dx, D, U = 3, 41, 10.0, 0.01, 1e-3
    z = np.zeros((nr, nc))
    mask = np.ones(z.shape, np.uint8)
    prm = orc.hs_params(D, nonlinear=False, U=U, rel_tol=1e-13, abs_tol=1e-15)
    zz, r = orc.hillslope_newton(z, z, mask, dx, "nono", prm, 0.0)
    assert r.converged
    mid = zz[1]
    curv = (mid[2:] - 2 * mid[1:-1] + mid[:-2]) / dx ** 2
    np.testing.assert_allclose(curv, -U / D, rtol=1e-8)


@pytest.mark.parametrize("nonlinear", [False, True])
def test_hillslope_mass_conservation_closed_domain(nonlinear):
    rng = np.random.default_rng(9)
    z = rng.random((20, 24)) * 3.0 + 50.0
    mask = np.ones(z.shape, np.uint8)
    prm = orc.hs_params(0.01, Sc=0.8, nonlinear=nonlinear, rel_tol=1e-12, abs_tol=1e-14)
    zz, r = orc.hillslope_transient(z, mask, 1.0, "nnnn", prm, 100.0)
    assert r.converged
    assert abs(zz.sum() - z.sum()) <= 1e-9 * abs(z.sum())
    assert np.abs(zz - z).max() > 1e-3   # something moved


def test_roering_steady_profile():
    """1-D hillslope between a divide (no flux, west) and a fixed channel
    (east): steady Roering slopes follow the closed form to O(h^2)"""
    D, U, Sc = 0.01, 1e-4, 0.8
    for n, dx in ((51, 2.0), (101, 1.0)):
        z = np.zeros((3, n))
        mask = np.ones(z.shape, np.uint8)
        prm = orc.hs_params(D, Sc=Sc, nonlinear=True, U=U, rel_tol=1e-13, abs_tol=1e-15)
        zz, r = orc.hillslope_transient(z, mask, dx, "nonn", prm, 1e7, 1e3)
        assert r.converged
        row = zz[1]
        x_face = (np.arange(n - 1) + 0.5) * dx + 0.5 * dx     # distance of face from the divide
        S = -(row[1:] - row[:-1]) / dx
        want = (Sc ** 2 / (2 * U * x_face)) * (np.sqrt(D ** 2 + 4 * U ** 2 * x_face ** 2 / Sc ** 2) - D)
        err = np.abs(S - want).max() / want.max()
        if n == 51:
            err51 = err
        else:
            assert err < 0.6 * err51 or err < 1e-6   # converges with refinement
    assert err51 < 0.05


def test_lem_step_boundary_and_uplift_only():
    rng = np.random.default_rng(2)
    z = rng.random((30, 40)) * 10 + 100
    mask = np.ones(z.shape, np.uint8)
    p = orc.lem_params(100.0, 1e-5, 0.5, 1.0, 1e-3, 0.01, bc="nnon", fluvial=False, hillslope=False,
                       baselevel_fall=2e-4)
    zz, st, rc = orc.lem_step(z, mask, 30.0, p)
    bf = orc.boundary_flats(mask, "nnon")
    interior = np.ones(z.size, bool)
    interior[bf] = False
    np.testing.assert_allclose(zz.ravel()[interior], z.ravel()[interior] + 0.1, rtol=0, atol=1e-12)
    np.testing.assert_allclose(zz.ravel()[bf], z.ravel()[bf] - 0.02, rtol=0, atol=1e-12)
    assert st["n_valid"] == z.size
