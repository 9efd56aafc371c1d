"""Oracle D-infinity flow graph (oracle/orc_dinf.c) pinned by geometry and
linear algebra, and the threshold channel set of the LEM hillslope step.

  * planes: on a plane descending at angle theta the facet that holds theta
    receives the flow, split between its cardinal and diagonal cells in the
    ratio of the angle inside the facet (Tarboton 1997); slope = gradient,
    bearing = compass direction of theta.  Checked for every facet.
  * weights sum to 1, at most 2 receivers (tests/cpp/test_flow_properties.cpp
    P7), receivers strictly lower.
  * accumulation equals the dense solve of (I - W^T) a = dx^2 (valid cells);
    the sinks collect n_valid dx^2; the Kahn order is a topological order.
  * channel_threshold_quantity: tests/cpp/test_channel_threshold_units.cpp U1.
  * LEM: the derived channel pixels are Dirichlet (unchanged by a hillslope-
    only step, tests/cpp/test_lem_channel_coupling.cpp C1); the coupled
    regime ignores the threshold (C3).
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc


def plane(nr, nc, theta_deg, g=0.1, dx=10.0, z0=10.0):
    """z descending at angle theta (counter-clockwise from east, north up)"""
    t = math.radians(theta_deg)
    r, c = np.mgrid[0:nr, 0:nc]
    x, y = c * dx, -r * dx
    return (z0 - g * (math.cos(t) * x + math.sin(t) * y)).astype(np.float32)


# facet a: cardinal (dr, dc), diagonal (dr, dc), angle range start, sign
FACETS = [((0, 1), (-1, 1), 0, 1), ((-1, 0), (-1, 1), 90, -1), ((-1, 0), (-1, -1), 90, 1),
          ((0, -1), (-1, -1), 180, -1), ((0, -1), (1, -1), 180, 1), ((1, 0), (1, -1), 270, -1),
          ((1, 0), (1, 1), 270, 1), ((0, 1), (1, 1), 360, -1)]


@pytest.mark.parametrize("a", range(8))
@pytest.mark.parametrize("frac", [0.2, 0.55, 0.9])
def test_plane_facets(a, frac):
    card, diag, base, sign = FACETS[a]
    # angle inside the facet measured from the cardinal direction
    r_deg = 45.0 * frac
    theta = base + sign * r_deg
    z = plane(7, 7, theta)
    g = orc.dinf_graph(z, np.ones(z.shape, np.uint8), 10.0, bc="nnnn")
    f = 3 * 7 + 3
    e1 = (3 + card[0]) * 7 + 3 + card[1]
    e2 = (3 + diag[0]) * 7 + 3 + diag[1]
    assert g["nrec"][f] == 2
    assert list(g["rec"][f]) == [e1, e2]
    assert abs(g["w"][f][1] - frac) < 1e-5
    assert abs(g["w"][f][0] - (1.0 - frac)) < 1e-5
    assert abs(g["slope"][3, 3] - 0.1) < 1e-6 * 10
    expect = (90.0 - theta) % 360.0
    assert abs((g["bearing"][3, 3] - expect + 180.0) % 360.0 - 180.0) < 1e-3


@pytest.mark.parametrize("theta,drc", [(0.0, (0, 1)), (90.0, (-1, 0)), (180.0, (0, -1)), (270.0, (1, 0))])
def test_plane_cardinal_single_receiver(theta, drc):
    z = plane(5, 5, theta)
    g = orc.dinf_graph(z, np.ones(z.shape, np.uint8), 10.0, bc="nnnn")
    f = 2 * 5 + 2
    assert g["nrec"][f] == 1
    assert g["rec"][f][0] == (2 + drc[0]) * 5 + 2 + drc[1]
    assert g["w"][f][0] == 1.0


def random_surface(nr, nc, seed, holes=0):
    rng = np.random.default_rng(seed)
    z = (rng.normal(size=(nr, nc)).cumsum(0).cumsum(1) * 0.3 + rng.normal(size=(nr, nc))).astype(np.float32)
    mask = np.ones((nr, nc), np.uint8)
    for _ in range(holes):
        mask[rng.integers(1, nr - 1), rng.integers(1, nc - 1)] = 0
    return z, mask


@pytest.mark.parametrize("bc", ["nnnn", "oooo", "pnpn", "bobo", "pppp"])
def test_graph_properties(bc):
    z, mask = random_surface(31, 27, 3, holes=6)
    g = orc.dinf_graph(z, mask, 10.0, bc=bc)
    zf = z.ravel()
    for f in range(z.size):
        k = g["nrec"][f]
        assert k <= 2
        if not mask.ravel()[f]:
            assert k == 0
        if k:
            assert abs(float(g["w"][f][:k].astype(np.float64).sum()) - 1.0) <= 1e-6
            for j in range(k):
                assert zf[g["rec"][f][j]] < zf[f]
                assert mask.ravel()[g["rec"][f][j]]
    nr, nc = z.shape
    nrec = g["nrec"].reshape(nr, nc)
    # forced outlets ('b' edges: N, E, S, W) have no receivers
    for side, edge in zip(bc, (nrec[0, :], nrec[:, -1], nrec[-1, :], nrec[:, 0])):
        if side == "b":
            assert edge.max() == 0
    # interior nodata: cells touching a hole route nowhere
    for r in range(1, nr - 1):
        for c in range(1, nc - 1):
            if mask[r, c] and mask[r - 1:r + 2, c - 1:c + 2].min() == 0:
                assert g["nrec"][r * nc + c] == 0


@pytest.mark.parametrize("bc", ["nnnn", "oooo", "pppp"])
def test_accumulate_equals_linear_solve(bc):
    z, mask = random_surface(14, 17, 9, holes=3)
    dx = 10.0
    g = orc.dinf_graph(z, mask, dx, bc=bc)
    valid = mask.ravel().astype(bool)
    acc, topo = orc.dinf_accumulate(g, mask, dx)
    n = z.size
    W = np.zeros((n, n))
    for f in range(n):
        for j in range(g["nrec"][f]):
            W[f, g["rec"][f][j]] += float(g["w"][f][j])
    idx = np.flatnonzero(valid)
    A = np.eye(len(idx)) - W[np.ix_(idx, idx)].T
    a = np.linalg.solve(A, np.full(len(idx), dx * dx))
    assert np.allclose(acc[idx], a, rtol=1e-12, atol=0)
    assert np.all(acc[~valid] == 0.0)
    # sinks collect everything
    sinks = valid & (g["nrec"] == 0)
    assert abs(acc[sinks].sum() - valid.sum() * dx * dx) <= 1e-9 * valid.sum() * dx * dx
    # the order is topological and covers the valid cells once
    pos = np.full(n, -1)
    pos[topo] = np.arange(len(topo))
    assert sorted(topo) == list(idx)
    for f in idx:
        for j in range(g["nrec"][f]):
            assert pos[g["rec"][f][j]] > pos[f]


def test_threshold_quantity_units():
    # U1: A = 25000 m^2 at dx = 10 m is 2500 m specific area, or 25000 m^2
    assert orc.channel_threshold_quantity(25000.0, 10.0, True) == 2500.0
    assert orc.channel_threshold_quantity(25000.0, 10.0, False) == 25000.0
    assert orc.channel_threshold_quantity(1.0, 0.0, True) == 1.0 / 1e-12


def test_channel_mask_is_area_threshold():
    z, mask = random_surface(40, 36, 5)
    dx = 10.0
    thr = 300.0
    ch, area = orc.dinf_channel_mask(z, mask, dx, thr, bc="oooo")
    g = orc.dinf_graph(z, mask, dx, bc="oooo", nodata=float("nan"))
    acc, _ = orc.dinf_accumulate(g, mask, dx)
    assert np.array_equal(area.ravel(), acc)
    assert np.array_equal(ch.ravel().astype(bool), (acc / dx >= thr) & mask.ravel().astype(bool))
    assert 0 < ch.sum() < mask.sum()
    ch0, _ = orc.dinf_channel_mask(z, mask, dx, 0.0, bc="oooo")
    assert ch0.sum() == 0
    chA, _ = orc.dinf_channel_mask(z, mask, dx, thr * dx, bc="oooo", divide_area_by_dx=False)
    assert np.array_equal(chA, ch)


def test_libm_atan2_differs_by_at_most_ulps():
    """the reference's atan2f (this C library) against the correctly rounded
    value: same receivers, weights within a few float ulps, same channels"""
    z, mask = random_surface(60, 50, 12)
    a = orc.dinf_graph(z, mask, 10.0, bc="oooo", libm_atan2=False)
    b = orc.dinf_graph(z, mask, 10.0, bc="oooo", libm_atan2=True)
    assert np.array_equal(a["rec"], b["rec"])
    assert np.abs(a["w"] - b["w"]).max() <= 4 * np.finfo(np.float32).eps
    ca, _ = orc.dinf_channel_mask(z, mask, 10.0, 200.0, bc="oooo", libm_atan2=False)
    cb, _ = orc.dinf_channel_mask(z, mask, 10.0, 200.0, bc="oooo", libm_atan2=True)
    assert np.array_equal(ca, cb)


def valley(nr=40, nc=21, dx=10.0):
    """a V valley draining south to an 'o' edge (channel column 10)"""
    r, c = np.mgrid[0:nr, 0:nc]
    return 100.0 + 0.02 * dx * (nr - 1 - r) + 0.3 * dx * np.abs(c - nc // 2)


def test_lem_derived_channels_are_dirichlet():
    z = valley()
    mask = np.ones(z.shape, np.uint8)
    p = orc.lem_params(200.0, 0.0, 0.5, 1.0, 0.0, 0.5, Sc=0.8, nonlinear=True, bc="nnon", threshold=150.0,
                       fluvial=False, uplift=False)
    ch, _ = orc.dinf_channel_mask(z, mask, 10.0, 150.0, bc="nnon")
    assert 0 < ch.sum() < 0.2 * z.size
    zn, st, rc = orc.lem_step(z, mask, 10.0, p)
    assert rc == 0
    assert np.array_equal(zn[ch != 0], z[ch != 0])
    assert np.abs(zn - z)[ch == 0].max() > 1e-3
    # coupled: the threshold names no channels
    pc = orc.lem_params(200.0, 0.0, 0.5, 1.0, 0.0, 0.5, Sc=0.8, nonlinear=True, bc="nnon", threshold=150.0,
                        fluvial=False, uplift=False, regime=1)
    p0 = orc.lem_params(200.0, 0.0, 0.5, 1.0, 0.0, 0.5, Sc=0.8, nonlinear=True, bc="nnon", fluvial=False,
                        uplift=False)
    zc, _, _ = orc.lem_step(z, mask, 10.0, pc)
    z0, _, _ = orc.lem_step(z, mask, 10.0, p0)
    assert np.array_equal(zc, z0)


# ---- D-infinity stream power (lsdtt_stream_power_cpu.cpp:329-602) --------

def east_plane(nr=12, nc=30, dx=10.0):
    c = np.arange(nc, dtype=np.float64)
    return np.repeat((200.0 - 0.5 * dx * 0.1 * c)[None, :], nr, axis=0)


@pytest.mark.parametrize("nexp", [1.0, 1.5])
def test_dinf_incision_equals_d8_on_a_cardinal_plane(nexp):
    """every cell's single D-inf receiver is its D8 receiver (east, dx): the
    two incision laws are the same computation"""
    z = east_plane()
    mask = np.ones(z.shape, np.uint8)
    kw = dict(hillslope=False, uplift=False, bc="nonn")
    pd = orc.lem_params(1000.0, 1e-4, 0.5, nexp, 0.0, 0.0, fluvial_dinf=True, **kw)
    p8 = orc.lem_params(1000.0, 1e-4, 0.5, nexp, 0.0, 0.0, **kw)
    zd, sd, rd = orc.lem_step(z, mask, 10.0, pd)
    z8, s8, r8 = orc.lem_step(z, mask, 10.0, p8)
    assert rd == 0 and r8 == 0
    assert np.array_equal(zd, z8)
    assert sd["cells_eroded"] == s8["cells_eroded"] > 0
    assert (zd < z).any()


def test_dinf_single_diagonal_receiver_uses_the_cardinal_length():
    """a cell whose only lower neighbour is diagonal has one receiver, which
    the reference classes as 'cardinal' (one receiver): the n = 1 closed
    form runs with dx, not dx sqrt2 (lsdtt_stream_power_cpu.cpp:435-441,
    505-528)"""
    z = np.full((5, 5), 11.0)
    z[2, 2] = 10.0
    z[1, 3] = 9.0   # NE of the centre
    z[0, 4] = 8.0   # NE of that: the diagonal cell drains on to the corner
    mask = np.ones(z.shape, np.uint8)
    g = orc.dinf_graph(z.astype(np.float32), mask, 10.0, bc="nnnn")
    f = 2 * 5 + 2
    assert g["nrec"][f] == 1 and g["rec"][f][0] == 1 * 5 + 3
    area, _ = orc.dinf_accumulate(g, mask, 10.0)
    K, m, dt, dx = 1e-3, 0.5, 100.0, 10.0
    zn, _ = orc.stream_power_dinf(z, z, mask, g, area, K, m, 1.0, dt, dx)
    a = area[f] / dx
    spf = K * a ** m * (dt / dx)
    zr = zn.ravel()[1 * 5 + 3]
    expect = (10.0 + zr * spf) / (1.0 + spf)
    assert abs(zn[2, 2] - expect) <= 1e-12 * abs(expect)


def test_dinf_two_receiver_newton_root():
    """cells with two receivers end at a root of
    z - z_old + K A^m dt sbt^(n/2) = 0 (to the Newton stopping rule), unless
    the rise guard or the clamp fired"""
    rng = np.random.default_rng(4)
    r, c = np.mgrid[0:40, 0:40]
    z = 300.0 - 0.8 * r - 0.45 * c + rng.normal(0, 0.05, (40, 40))
    mask = np.ones(z.shape, np.uint8)
    dx, K, m, nexp, dt = 10.0, 2e-4, 0.5, 1.0, 200.0
    g = orc.dinf_graph(z.astype(np.float32), mask, dx, bc="oooo")
    area, _ = orc.dinf_accumulate(g, mask, dx)
    zn, _ = orc.stream_power_dinf(z, z, mask, g, area, K, m, nexp, dt, dx)
    zf, zo = zn.ravel(), z.ravel()
    checked = 0
    for f in range(z.size):
        if g["nrec"][f] != 2 or g["w"][f][0] >= 1 - 1e-6 or g["w"][f][1] >= 1 - 1e-6:
            continue
        e1, e2 = zf[g["rec"][f][0]], zf[g["rec"][f][1]]
        if zf[f] <= min(e1, e2) + 1e-5 * dx * 1.0001 or zf[f] == zo[f]:
            continue  # clamp or guard
        sbt = ((e1 - zf[f]) / dx) ** 2 + ((e2 - e1) / dx) ** 2
        kam = K * (area[f] / dx) ** m
        fz = zf[f] - zo[f] + kam * sbt ** (0.5 * nexp) * dt
        fd = 1.0 - kam * nexp * (e1 - zf[f]) / dx ** 2 * sbt ** (0.5 * nexp - 1.0) * dt
        assert abs(fz / fd) <= 1e-5 * dx
        checked += 1
    assert checked > 100
