"""The oracle's step statistics, erosion-rate field and boundary flats on
cases whose answers are written out by hand.

run_lem_step (src/cpu/lsdtt_lem_cpu.cpp:614-722): uplift U dt on every valid
cell, the boundary flats reset to their pre-step value minus
baselevel_fall_rate * dt, then per valid cell the statistics of
z_before -> z_after: |max after - max before|, |mean after - mean before|,
max |dz|, |dz| (float) at index floor(0.999 (n - 1)) of the sorted list,
and the erosion z_before - (z_after - U dt) -- uplift backed out on every
valid cell, the flats included -- whose positive part gives the mean
(over all valid cells) and maximum erosion rate, and whose signed value
over dt is the per-pixel field (0 off the mask).

build_boundary_flats (:735-808): valid cells with an 8-neighbour off a fixed
('o' or 'b') edge, or a masked 8-neighbour in the direction of a fixed edge
(the first matching branch of north, south, west, east decides an
off-grid neighbour).

Every input below is a dyadic rational, so each expected value is exact.
"""
import numpy as np
import pytest

from oracle import oracle as orc


def test_uplift_only_step_statistics():
    nr, nc = 4, 5
    z = (10.0 * np.arange(nr)[:, None] + np.arange(nc)[None, :]).astype(np.float64)
    mask = np.ones(z.shape, np.uint8)
    # U dt = 0.0625 * 8 = 0.5; base-level fall 0.03125 * 8 = 0.25
    p = orc.lem_params(8.0, 1e-5, 0.5, 1.0, 0.0625, 0.0, bc="nnon", fluvial=False, hillslope=False,
                       baselevel_fall=0.03125)
    er = np.empty(z.shape, np.float32)
    zz, st, rc = orc.lem_step(z, mask, 30.0, p, erosion_rate=er)
    assert rc == 0
    assert list(orc.boundary_flats(mask, "nnon")) == [15, 16, 17, 18, 19]
    want = z.copy()
    want[:3] += 0.5
    want[3] -= 0.25
    assert np.array_equal(zz, want)
    assert st["n_valid"] == 20
    assert st["delta_max_elevation"] == 0.25           # max 34 -> 33.75
    assert st["delta_mean_elevation"] == 0.3125        # (15 * 0.5 - 5 * 0.25) / 20
    assert st["max_local_elevation_change"] == 0.5
    assert st["p999_local_elevation_change"] == 0.5    # sorted |dz|: 5 x 0.25, 15 x 0.5; index 18
    # uplifted cells: zero erosion; flats: 0.25 fall + 0.5 uplift backed out
    assert st["mean_erosion_rate"] == 5 * 0.75 / 20 / 8.0
    assert st["max_erosion_rate"] == 0.75 / 8.0
    want_er = np.zeros(z.shape, np.float32)
    want_er[3] = 0.09375
    assert np.array_equal(er, want_er)


def single_cell_case():
    """three identical rows [0, 10] draining west to an 'o' edge: each east
    cell erodes with spf = K (A/dx)^m dt / dx = 0.125 * 2 * 16 / 4 = 1, so
    z = (10 + 0) / 2 = 5, then 5.5 after uplift"""
    z = np.array([[0.0, 10.0]] * 3)
    mask = np.ones(z.shape, np.uint8)
    p = orc.lem_params(16.0, 0.125, 0.5, 1.0, 0.03125, 0.0, bc="nnno", hillslope=False)
    return z, mask, p


def test_single_eroding_cell_statistics():
    z, mask, p = single_cell_case()
    er = np.empty(z.shape, np.float32)
    zz, st, rc = orc.lem_step(z, mask, 4.0, p, erosion_rate=er)
    assert rc == 0
    assert list(orc.boundary_flats(mask, "nnno")) == [0, 2, 4]
    assert np.array_equal(zz, np.array([[0.0, 5.5]] * 3))
    assert st["cells_eroded"] == 3
    assert st["delta_max_elevation"] == 4.5
    assert st["delta_mean_elevation"] == 2.25
    assert st["max_local_elevation_change"] == 4.5
    assert st["p999_local_elevation_change"] == 4.5    # sorted [0, 0, 0, 4.5, 4.5, 4.5], index 4
    # east cells erode 10 - (5.5 - 0.5) = 5; the outlets 0 - (0 - 0.5) = 0.5
    assert st["mean_erosion_rate"] == 3 * 5.5 / 6 / 16.0
    assert st["max_erosion_rate"] == 5.0 / 16.0
    assert np.array_equal(er, np.array([[0.03125, 0.3125]] * 3, np.float32))


def test_deposition_is_a_negative_rate_and_the_field_matches_the_sums():
    """a surface that rises by more than the uplift (a base level raised by
    a negative fall): the field keeps the sign, the statistics count only
    erosion"""
    rng = np.random.default_rng(3)
    z = np.round(rng.random((12, 14)) * 64.0) / 4.0 + 100.0
    mask = np.ones(z.shape, np.uint8)
    mask[5, 6] = 0
    p = orc.lem_params(4.0, 1e-3, 0.5, 1.0, 0.25, 0.0, bc="onon", hillslope=False, baselevel_fall=-0.5)
    er = np.empty(z.shape, np.float32)
    zz, st, rc = orc.lem_step(z, mask, 10.0, p, erosion_rate=er)
    assert rc == 0
    bf = orc.boundary_flats(mask, "onon")
    # flats: z_before - ((z_before + 2) - 1) = -1 per 4 yr
    assert np.all(er.ravel()[bf] == -0.25)
    assert er[5, 6] == 0.0
    valid = mask.ravel() != 0
    ero = (z - (zz - 1.0)).ravel()[valid]
    assert st["max_erosion_rate"] == max(ero.max(), 0.0) / 4.0
    np.testing.assert_allclose(st["mean_erosion_rate"], np.clip(ero, 0, None).sum() / valid.sum() / 4.0,
                               rtol=1e-15)
    np.testing.assert_allclose(er.ravel()[valid], ero / 4.0, rtol=1e-7)


@pytest.mark.parametrize("bc,holes,want", [
    # south fixed: the last row, and the valid cells whose southward
    # neighbours (S, SW, SE) include a hole
    ("nnon", [(1, 2), (2, 4), (4, 0)], [1, 2, 3, 9, 10, 11, 18, 19, 25, 26, 27, 28, 29]),
    # north 'b' fixed: the first row and the three cells below the hole
    ("bnnn", [(2, 2)], [0, 1, 2, 3, 4, 5, 19, 20, 21]),
    # west and east fixed: columns 0 and 5, and the cells beside each hole
    # (W/E neighbours, diagonals included); north/south holes do not count
    ("nono", [(2, 2)], [0, 5, 6, 7, 9, 11, 12, 13, 15, 17, 18, 19, 21, 23]),
    # nothing fixed: no flats, holes or not
    ("npnp", [(2, 2), (0, 0)], []),
])
def test_boundary_flats_hand_lists(bc, holes, want):
    nr, nc = (4, 6) if bc != "nnon" else (5, 6)
    mask = np.ones((nr, nc), np.uint8)
    for r, c in holes:
        mask[r, c] = 0
    assert list(orc.boundary_flats(mask, bc)) == want
