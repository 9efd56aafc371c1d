"""Oracle analytic fluvial steady state (oracle/orc_steady.c) pinned by the
property that defines it and by closed forms.

  * every pixel with a receiver satisfies U = K A^m S^n with
    S = (z - z_r) / dx_eff (the steady state of the stream-power law,
    src/cpu/lsdtt_steady_state_cpu.cpp:76-170), to float rounding;
  * a 1-D channel's chi equals the closed-form sum of dx (k dx)^(-m/n);
  * relief mode reaches target_relief and back-calculates K;
  * with an erodibility field, U = K_i A^m S^n per pixel;
  * chi_c counts the channel pixels and takes their largest chi.
"""
import numpy as np
import pytest

from oracle import oracle as orc


def surface(nr, nc, seed):
    rng = np.random.default_rng(seed)
    r, c = np.mgrid[0:nr, 0:nc]
    return (100.0 + 0.5 * r + 0.3 * c + 0.1 * rng.random((nr, nc))).astype(np.float32)


@pytest.mark.parametrize("n,divide", [(1.0, True), (1.5, True), (1.0, False)])
def test_steady_state_satisfies_stream_power_balance(n, divide):
    dx, U, K, m = 10.0, 2e-4, 3e-5, 0.45
    z = surface(40, 33, 1)
    mask = np.ones(z.shape, np.uint8)
    zs, res = orc.steady_state_d8(z, mask, dx, bc="oooo", U=U, K=K, m=m, n=n, divide_area_by_dx=divide)
    fl, area = res["flow"], res["area"].ravel()
    zf = zs.ravel().astype(np.float64)
    checked = 0
    for s in range(z.size):
        r = fl["receiver"][s]
        if r < 0:
            continue
        dxe = dx * np.sqrt(2.0) if fl["dir"][s] & 1 else dx
        A = area[s] / dx if divide else area[s]
        S = (zf[s] - zf[r]) / dxe
        assert S > 0
        assert abs(K * A ** m * S ** n - U) <= 2e-5 * U
        checked += 1
    assert checked > 1000
    assert res["K_used"] == np.float32(K)


def test_one_dimensional_channel_closed_form():
    """a single row draining west to an 'o' edge: chi(k cells from the
    outlet) = sum_j dx (A_j / dx)^(-m/n) with A_j = (ncols - j) dx^2"""
    nc, dx, m, nexp = 50, 10.0, 0.5, 1.0
    z = (np.arange(nc, dtype=np.float32) * 2.0 + 5.0)[None, :]
    mask = np.ones(z.shape, np.uint8)
    zs, res = orc.steady_state_d8(z, mask, dx, bc="nnno", U=1e-3, K=1e-4, m=m, n=nexp)
    f = (1e-3 / 1e-4) ** (1.0 / nexp)
    chi = 0.0
    for k in range(1, nc):
        A = (nc - k) * dx * dx / dx
        chi += dx * A ** (-m / nexp)
        assert abs(zs[0, k] - f * chi) <= 1e-5 * f * chi
    assert zs[0, 0] == 0.0


def test_relief_mode():
    z = surface(30, 30, 2)
    mask = np.ones(z.shape, np.uint8)
    zs, res = orc.steady_state_d8(z, mask, 10.0, bc="oooo", U=1e-4, m=0.45, n=1.3, relief_mode=True,
                                  target_relief=750.0, baselevel=12.0)
    assert abs(zs.max() - 12.0 - 750.0) <= 1e-3
    assert abs(res["computed_relief"] - 750.0) <= 1e-3
    # U = K_used * (max chi / relief)^n
    assert abs(res["K_used"] - 1e-4 * (res["max_unit_height"] / 750.0) ** 1.3) <= 1e-6 * res["K_used"]


def test_erodibility_field_balances_per_pixel():
    dx, U, K, m, nexp = 10.0, 2e-4, 3e-5, 0.5, 1.0
    z = surface(25, 28, 3)
    mask = np.ones(z.shape, np.uint8)
    Kf = np.random.default_rng(4).uniform(1e-5, 1e-4, z.shape).astype(np.float32)
    zs, res = orc.steady_state_d8(z, mask, dx, bc="oooo", U=U, K=K, m=m, n=nexp, K_field=Kf)
    fl, area = res["flow"], res["area"].ravel()
    zf = zs.ravel().astype(np.float64)
    for s in range(z.size):
        r = fl["receiver"][s]
        if r < 0:
            continue
        dxe = dx * np.sqrt(2.0) if fl["dir"][s] & 1 else dx
        S = (zf[s] - zf[r]) / dxe
        assert abs(float(Kf.ravel()[s]) * (area[s] / dx) ** m * S ** nexp - U) <= 3e-5 * U
    with pytest.raises(orc.OracleError):
        bad = Kf.copy()
        bad[3, 3] = 0.0
        orc.steady_state_d8(z, mask, dx, K_field=bad)


def test_channel_head_chi():
    z = surface(30, 30, 5)
    mask = np.ones(z.shape, np.uint8)
    zs, res = orc.steady_state_d8(z, mask, 10.0, bc="oooo", channel_threshold_area=200.0)
    area = res["area"]
    q = area / 10.0
    assert res["chi_c_reported"] and res["n_channel_pixels"] == int((q >= 200.0).sum())
    f = (1e-4 / 1e-4) ** 1.0
    assert res["max_unit_height_at_channel_head"] == np.float32(zs[q >= 200.0].max() / f)
    assert res["max_unit_height_at_channel_head"] <= res["max_unit_height"]


def test_libm_pow_within_an_ulp():
    z = surface(40, 40, 6)
    mask = np.ones(z.shape, np.uint8)
    a, _ = orc.steady_state_d8(z, mask, 10.0, bc="oooo", m=0.45, n=1.2, libm_pow=False)
    b, _ = orc.steady_state_d8(z, mask, 10.0, bc="oooo", m=0.45, n=1.2, libm_pow=True)
    assert np.allclose(a, b, rtol=4e-7, atol=0)
