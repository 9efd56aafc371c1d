"""Oracle drainage-area moving average (orc_smooth_area) against a brute
box average over the valid cells of each window (periodic axes wrap).  With
integer-valued areas every partial sum is exact, so the rolling sums of the
reference order and the brute sums agree bit for bit."""
import numpy as np
import pytest

from oracle import oracle as orc


def brute(area, mask, k, bc):
    nr, nc = area.shape
    rad = k // 2
    out = np.zeros_like(area)
    pn, pe, ps, pw = (c == "p" for c in bc)
    for r in range(nr):
        for c in range(nc):
            if not mask[r, c]:
                continue
            s = 0.0
            cnt = 0
            for dr in range(-rad, rad + 1):
                rr = r + dr
                if rr < 0 and not pn or rr >= nr and not ps:
                    continue
                rr %= nr
                for dc in range(-rad, rad + 1):
                    cc = c + dc
                    if cc < 0 and not pw or cc >= nc and not pe:
                        continue
                    cc %= nc
                    if mask[rr, cc]:
                        s += area[rr, cc]
                        cnt += 1
            out[r, c] = s / cnt if cnt else 0.0
    return out


@pytest.mark.parametrize("k,bc", [(3, "oooo"), (5, "nnon"), (3, "popo"), (7, "opop"), (5, "pppp")])
def test_smoothing_is_window_mean(k, bc):
    rng = np.random.default_rng(k)
    area = rng.integers(1, 5000, (23, 31)).astype(np.float64) * 100.0
    mask = (rng.random(area.shape) > 0.1).astype(np.uint8)
    got = orc.smooth_area(area, mask, k, bc)
    assert np.array_equal(got, brute(area, mask, k, bc))


def test_smoothing_identity_and_errors():
    area = np.arange(12.0).reshape(3, 4)
    mask = np.ones(area.shape, np.uint8)
    assert np.array_equal(orc.smooth_area(area, mask, 1), area)
    for k in (0, 2, -3):
        with pytest.raises(orc.OracleError):
            orc.smooth_area(area, mask, k)


def test_smoothing_constant_field_is_fixed():
    area = np.full((20, 20), 1234.0)
    mask = np.ones(area.shape, np.uint8)
    mask[5:8, 5:9] = 0
    got = orc.smooth_area(area, mask, 5, "oooo")
    assert np.all(got[mask != 0] == 1234.0) and np.all(got[mask == 0] == 0.0)
