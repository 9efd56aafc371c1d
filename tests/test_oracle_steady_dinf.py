"""Oracle D-infinity analytic steady state (orc_steady_state_dinf,
oracle/orc_steady.c) pinned by the reference's own checks and by closed
forms:

  * S1 (tests/cpp/test_lem_erodibility.cpp:774-810): a uniform erodibility
    field equal to K gives the scalar-K elevations bit for bit;
  * S4 (:866-925): relief mode with a field reaches the target relief within
    4 ulp, erodibility_field_scale == K_used / K, and K mode with
    scale * field reproduces the relief-mode surface;
  * S5 / S5b (:927-987): every single-receiver cell satisfies
    U = K_i a^m S^n with S = dz / dx_eff, and every two-receiver cell either
    the full-gradient identity |grad z| a^m = (U / K_i)^(1/n) or the
    cardinal fallback;
  * C7 (tests/cpp/test_chi_channel_head.cpp:250-272): chi_c is reported,
    with a channel count and below max chi;
  * closed forms: on a north-draining ramp every receiver is the cardinal
    one and chi(r) = sum_j dx ((nr - j) dx)^(-m/n); the same ramp and a
    diagonal plane give the D8 steady state (same float operations).
"""
import numpy as np
import pytest

from oracle import oracle as orc

DX = 30.0


def ulp(x):
    return float(np.spacing(np.float32(abs(x))))


def valley(nr=40, nc=30, seed=7, noise=0.05):
    """a V valley draining north with a little noise: single cardinal
    receivers along the floor, two-receiver cells on the walls"""
    rng = np.random.default_rng(seed)
    r, c = np.mgrid[0:nr, 0:nc]
    z = 100.0 + 0.01 * DX * r + 0.02 * DX * np.abs(c - nc // 2) + noise * rng.random((nr, nc))
    return z.astype(np.float32), np.ones((nr, nc), np.uint8)


def two_region(shape, K1, K2):
    k = np.full(shape, K2, np.float32)
    k[: shape[0] // 2] = K1
    return k


@pytest.mark.parametrize("relief", [False, True])
def test_s1_uniform_field_equals_scalar_K(relief):
    z, mask = valley()
    kw = dict(U=1e-4, K=2e-5, relief_mode=relief, target_relief=400.0)
    za, ra = orc.steady_state_dinf(z, mask, DX, **kw)
    zb, rb = orc.steady_state_dinf(z, mask, DX, K_field=np.full(z.shape, 2e-5, np.float32), **kw)
    assert np.array_equal(za.view(np.uint32), zb.view(np.uint32))
    assert ra["K_used"] == rb["K_used"]
    want = np.float32(rb["K_used"]) / np.float32(2e-5) if relief else 1.0
    assert np.float32(rb["erodibility_field_scale"]) == np.float32(want)
    assert ra["erodibility_field_scale"] == 1.0


def test_s4_relief_mode_with_a_field():
    z, mask = valley()
    field = two_region(z.shape, 8e-5, 2e-5)
    zr, rr = orc.steady_state_dinf(z, mask, DX, U=1e-4, K=2e-5, relief_mode=True, target_relief=400.0,
                                   K_field=field)
    assert abs(rr["computed_relief"] - 400.0) <= 4.0 * ulp(400.0)
    assert np.float32(rr["erodibility_field_scale"]) == np.float32(rr["K_used"]) / np.float32(2e-5)
    zk, _ = orc.steady_state_dinf(z, mask, DX, U=1e-4, K=2e-5,
                                  K_field=(field * np.float32(rr["erodibility_field_scale"])).astype(np.float32))
    assert np.max(np.abs(zr.astype(np.float64) - zk)) <= 200.0 * ulp(zr.max())


@pytest.mark.parametrize("nexp", [1.0, 2.0])
def test_s5_steady_law_cell_by_cell(nexp):
    z, mask = valley()
    U, K, m = 1e-4, 2e-5, 0.45
    field = two_region(z.shape, 8e-5, 2e-5)
    zs, res = orc.steady_state_dinf(z, mask, DX, U=U, K=K, m=m, n=nexp, K_field=field)
    g, area = res["graph"], res["area"].ravel()
    zf = zs.ravel().astype(np.float64)
    nc = z.shape[1]
    zmax = float(zs.max())
    rel_m = np.float64(np.float32(m)) / nexp
    one = one_bad = two_full = two_fb = two_bad = 0
    for i in range(z.size):
        k = int(g["nrec"][i])
        if k == 0:
            continue
        w0, w1 = g["w"][i]
        only_diag = k > 1 and w1 >= np.float32(1.0 - 1e-6)
        two_r = k > 1 and not (w0 >= np.float32(1.0 - 1e-6)) and not only_diag
        a = area[i] / DX
        kfac = (np.float64(field.ravel()[i]) / np.float64(np.float32(U))) ** (1.0 / nexp)
        if two_r:
            e1, e2 = g["rec"][i]
            ze1, ze2, zi = zf[e1], zf[e2], zf[i]
            dz = min(zi - ze1, abs(ze1 - ze2) + (zi - ze1))
            tol = 4.0 * ulp(zmax) / max(dz, 1e-12) + 1e-6
            grad = np.hypot((zi - ze1) / DX, (ze1 - ze2) / DX)
            q_full = grad ** nexp * a ** (rel_m * nexp) * kfac ** nexp
            q_card = ((zi - ze1) / DX) ** nexp * a ** (rel_m * nexp) * kfac ** nexp
            # the reference compares the n-th roots; the same up to the power
            q_full, q_card = q_full ** (1.0 / nexp), q_card ** (1.0 / nexp)
            if abs(q_full - 1.0) <= tol:
                two_full += 1
            elif abs(q_card - 1.0) <= tol:
                two_fb += 1
            else:
                two_bad += 1
        else:
            rr = g["rec"][i][1 if only_diag else 0]
            diag = abs(rr // nc - i // nc) == 1 and abs(rr % nc - i % nc) == 1
            dxe = DX * np.float64(np.float32(np.sqrt(2.0))) if diag else DX
            dz = zf[i] - zf[rr]
            tol = 4.0 * ulp(zmax) / dz + 1e-6
            q = dz / dxe * a ** rel_m * kfac
            one += 1
            one_bad += abs(q - 1.0) > tol
    assert one >= 50 and one_bad == 0, (one, one_bad)
    assert two_full >= 100 and two_bad == 0, (two_full, two_fb, two_bad)


def test_c7_chi_c_reported():
    z, mask = valley(60, 50, 3)
    zs, res = orc.steady_state_dinf(z, mask, DX, m=0.45, n=1.0, channel_threshold_area=1000.0)
    assert res["chi_c_reported"] and res["n_channel_pixels"] > 0
    assert res["max_unit_height_at_channel_head"] < res["max_unit_height"]
    # chi_c is the largest chi over the cells at or above the threshold
    chi = zs.astype(np.float64) / ((1e-4 / 1e-4) ** 1.0)
    chan = res["area"] / DX >= 1000.0
    assert res["n_channel_pixels"] == chan.sum()
    assert np.float32(res["max_unit_height_at_channel_head"]) == np.float32(chi[chan].max())


def test_ramp_closed_form_and_d8_equality():
    nr, nc, m, nexp, U, K = 35, 12, 0.5, 1.0, 1e-3, 1e-4
    z = (np.arange(nr, dtype=np.float32)[:, None] * 3.0 + 10.0).repeat(nc, 1)
    mask = np.ones(z.shape, np.uint8)
    zs, res = orc.steady_state_dinf(z, mask, DX, bc="onnn", U=U, K=K, m=m, n=nexp)
    g = res["graph"]
    assert np.all(g["nrec"][nc:] == 1) and np.all(g["nrec"][:nc] == 0)
    assert np.all(g["rec"][nc:, 0] == np.arange(nc * (nr - 1)))
    f = (U / K) ** (1.0 / nexp)
    for r in range(1, nr):
        chi = sum(DX * ((nr - j) * DX) ** (-m / nexp) for j in range(1, r + 1))
        assert np.allclose(zs[r], f * chi, rtol=1e-5, atol=0)
    assert np.all(zs[0] == 0.0)
    zd, _ = orc.steady_state_d8(z, mask, DX, bc="onnn", U=U, K=K, m=m, n=nexp)
    assert np.array_equal(zs.view(np.uint32), zd.view(np.uint32))


def test_diagonal_plane_uses_the_diagonal_length():
    """z = r + c drains north-west: every receiver is the diagonal cell, so
    chi steps by dx sqrt2 per cell along each diagonal"""
    nr = nc = 20
    r, c = np.mgrid[0:nr, 0:nc]
    z = (5.0 + r + c).astype(np.float32)
    mask = np.ones(z.shape, np.uint8)
    zs, res = orc.steady_state_dinf(z, mask, DX, bc="oonn", U=1e-4, K=1e-4, m=0.5, n=1.0)
    g = res["graph"]
    inner = (r >= 1) & (c >= 1)
    rec = np.where(g["w"][:, 1] >= np.float32(1 - 1e-6), g["rec"][:, 1], g["rec"][:, 0]).reshape(nr, nc)
    assert np.all(rec[inner] == ((r - 1) * nc + (c - 1))[inner])
    zd, rd = orc.steady_state_d8(z, mask, DX, bc="oonn", U=1e-4, K=1e-4, m=0.5, n=1.0)
    assert np.array_equal(rd["flow"]["receiver"].reshape(nr, nc)[inner], rec[inner])
    assert np.allclose(zs, zd, rtol=1e-5, atol=1e-6)
    # chi along the main diagonal: sum of dx sqrt2 (A / dx)^(-1/2)
    area = res["area"]
    chi = 0.0
    for k in range(1, nr):
        chi += DX * np.sqrt(2.0) * (area[k, k] / DX) ** -0.5
        assert abs(zs[k, k] - chi) <= 1e-5 * chi


def test_external_area_reproduces_own():
    z, mask = valley()
    za, ra = orc.steady_state_dinf(z, mask, DX, m=0.45, n=1.3)
    zb, _ = orc.steady_state_dinf(z, mask, DX, m=0.45, n=1.3, area=ra["area"])
    assert np.array_equal(za.view(np.uint32), zb.view(np.uint32))


def test_bad_erodibility_only_on_valid_cells():
    z, mask = valley()
    mask[20, 15] = 0
    for bad in (0.0, -1e-5, np.nan):
        k = np.full(z.shape, 2e-5, np.float32)
        k[10, 7] = bad
        with pytest.raises(orc.OracleError):
            orc.steady_state_dinf(z, mask, DX, K_field=k)
        k[10, 7] = 2e-5
        k[20, 15] = bad
        orc.steady_state_dinf(z, mask, DX, K_field=k)
