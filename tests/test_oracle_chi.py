"""Oracle chi (oracle/orc_chi.c) against what the mathematics fixes.

* theta = 0: (A0/A)^0 = 1, so each increment is exactly the step length and
  chi is the flow-path length accumulated in float from the outlet up.
* A 1-D channel (one row, or one diagonal) has contributing counts
  N - j, so chi is the trapezoid sum of (A0/(k dx^2))^theta, computed here in
  float64 (tolerance: float32 accumulation over <= 400 nodes, rtol 5e-5).
* Thresholded chi is the raster restricted to pixels with count >= the
  threshold, in s-vector order, one basin per qualifying outlet.
* A walk along the receivers of each cell (an independent traversal, no
  s-vector slices) reproduces every raster value exactly.
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc

K_SQRT2 = 1.4142135623730950488


def random_flow(nr, nc, bc, seed, holes=0.0, dx=10.0):
    rng = np.random.default_rng(seed)
    z = (rng.random((nr, nc)) * 10.0).astype(np.float32)
    mask = np.ones((nr, nc), np.uint8)
    if holes:
        mask[rng.random((nr, nc)) < holes] = 0
    flow = orc.flow_build(z, mask, dx, bc, allow_pits_as_outlets=True)
    return z, mask, flow


def path_steps(flow, f, dx):
    rcv, d = flow["receiver"], flow["dir"]
    cells = []
    while rcv[f] >= 0:
        cells.append(f)
        f = rcv[f]
    return cells  # upstream-most first


@pytest.mark.parametrize("bc", ["oooo", "nnon", "popo", "nnnn"])
def test_chi_theta_zero_is_float_path_length(bc):
    dx = 10.0
    z, mask, flow = random_flow(30, 40, bc, seed=3, holes=0.05 if bc == "nnnn" else 0.0, dx=dx)
    chi, cm = orc.chi_raster(flow, mask, dx, m_over_n=0.0, A0=1.0)
    d = flow["dir"]
    for f in np.flatnonzero(mask.ravel()):
        cells = path_steps(flow, f, dx)
        c = np.float32(0.0)
        for u in reversed(cells):
            step = dx if d[u] % 2 == 0 else dx * K_SQRT2
            c = np.float32(np.float64(c) + step)
        assert chi.ravel()[f] == c, (f, chi.ravel()[f], c)
    assert np.array_equal(cm, mask)
    assert np.all(chi[mask == 0] == np.float32(-9999.0))


@pytest.mark.parametrize("theta", [0.45, 0.5, 0.3])
def test_chi_row_channel_trapezoid(theta):
    n, dx, A0 = 300, 10.0, 1.0
    z = (np.arange(n, dtype=np.float32) * 0.5)[None, :] + 100.0
    mask = np.ones(z.shape, np.uint8)
    flow = orc.flow_build(z, mask, dx, "nnno", allow_pits_as_outlets=True)
    assert flow["receiver"][0] == -2 and np.all(flow["receiver"][1:] == np.arange(n - 1))
    chi, _ = orc.chi_raster(flow, mask, dx, m_over_n=theta, A0=A0)
    f = lambda k: (A0 / (k * dx * dx)) ** theta
    want = np.zeros(n)
    for j in range(1, n):
        want[j] = want[j - 1] + dx * 0.5 * (f(n - j) + f(n - j + 1))
    np.testing.assert_allclose(chi.ravel().astype(np.float64), want, rtol=5e-5, atol=0)


def test_chi_diagonal_channel_trapezoid():
    n, dx, A0, theta = 200, 30.0, 1.0, 0.45
    z = np.full((n, n), 1000.0, np.float32)
    mask = np.zeros((n, n), np.uint8)
    for k in range(n):
        z[k, k] = 10.0 + k
        mask[k, k] = 1
    flow = orc.flow_build(z, mask, dx, "oooo", allow_pits_as_outlets=True)
    chi, _ = orc.chi_raster(flow, mask, dx, m_over_n=theta, A0=A0)
    f = lambda k: (A0 / (k * dx * dx)) ** theta
    want = np.zeros(n)
    for j in range(1, n):
        want[j] = want[j - 1] + dx * K_SQRT2 * 0.5 * (f(n - j) + f(n - j + 1))
    got = np.array([chi[k, k] for k in range(n)], np.float64)
    np.testing.assert_allclose(got, want, rtol=5e-5, atol=0)


@pytest.mark.parametrize("theta", [0.45, 0.5])
def test_chi_receiver_walk_matches_raster(theta):
    dx, A0 = 30.137, 1.0
    z, mask, flow = random_flow(25, 33, "opop", seed=9, holes=0.03, dx=dx)
    chi, _ = orc.chi_raster(flow, mask, dx, m_over_n=theta, A0=A0)
    size, d = flow["s_size"], flow["dir"]
    pa = dx * dx
    fn = lambda k: math.pow(A0 / (float(k) * pa), theta)
    for f in np.flatnonzero(mask.ravel()):
        c = np.float32(0.0)
        for u in reversed(path_steps(flow, f, dx)):
            r = flow["receiver"][u]
            step = dx if d[u] % 2 == 0 else dx * K_SQRT2
            c = np.float32(np.float64(c) + step * 0.5 * (fn(size[u]) + fn(size[r])))
        assert chi.ravel()[f] == c


def test_chi_threshold_is_raster_restricted():
    dx, mp = 10.0, 20
    z, mask, flow = random_flow(60, 70, "oooo", seed=4, dx=dx)
    chi, _ = orc.chi_raster(flow, mask, dx)
    t = orc.chi_threshold(flow, mask, dx, mp)
    size, rcv = flow["s_size"], flow["receiver"]
    outlets = [o for o in np.flatnonzero(mask.ravel()) if rcv[o] == -2 and size[o] >= mp]
    assert len(t["basin_start"]) == len(outlets)
    want = []
    for o in outlets:
        st, sz = int(flow["s_start"][o]), int(size[o])
        want += [v for v in flow["s_vector"][st:st + sz] if size[v] >= mp]
    assert np.array_equal(t["flats"], np.array(want, np.int64))
    assert np.array_equal(t["chi"], chi.ravel()[t["flats"]])
    exp_rec = rcv[t["flats"]].copy()
    assert np.array_equal(t["receivers"], exp_rec)
    assert np.array_equal(t["flats"][t["basin_start"]], np.array(outlets, np.int64))
