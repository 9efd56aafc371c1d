"""The oracle's epsilon-free fill against brute force and the reference's
fill properties (tests/cpp/test_flow_properties.cpp:21-31: P1 depression
free, P2 never lowers, P6 idempotent; exit rule of
lsdtt_raster_flow_preprocessing_cpu.cpp:518-589)."""
import numpy as np
import pytest

from oracle import oracle as orc

D8 = [(-1, 0), (-1, 1), (0, 1), (1, 1), (1, 0), (1, -1), (0, -1), (-1, -1)]


def brute_fill(z, valid, per_ns=False, per_ew=False):
    """minimax path height to an exit by Bellman-Ford over all cells"""
    nr, nc = z.shape
    interior_invalid = (~valid[1:-1, 1:-1]).any() if nr > 2 and nc > 2 else False
    INF = np.float32(np.inf)
    W = np.full(z.shape, INF, np.float32)
    exits = np.zeros(z.shape, bool)
    for r in range(nr):
        for c in range(nc):
            if not valid[r, c]:
                continue
            e = (not per_ns and r in (0, nr - 1)) or (not per_ew and c in (0, nc - 1))
            if not e and interior_invalid:
                for dr, dc in D8:
                    rr, cc = r + dr, c + dc
                    if 0 <= rr < nr and 0 <= cc < nc and not valid[rr, cc]:
                        e = True
            exits[r, c] = e
            if e:
                W[r, c] = z[r, c]
    changed = True
    while changed:
        changed = False
        for r in range(nr):
            for c in range(nc):
                if not valid[r, c] or exits[r, c]:
                    continue
                m = INF
                for dr, dc in D8:
                    rr, cc = r + dr, c + dc
                    if per_ns:
                        rr %= nr
                    if per_ew:
                        cc %= nc
                    if 0 <= rr < nr and 0 <= cc < nc and valid[rr, cc]:
                        m = min(m, W[rr, cc])
                v = max(z[r, c], m)
                if v < W[r, c]:
                    W[r, c] = v
                    changed = True
    return W


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("bc", ["oooo", "nnnn", "popo", "opop"])
def test_oracle_fill_matches_brute_force(seed, bc):
    rng = np.random.default_rng(seed)
    nr, nc = rng.integers(1, 12, 2)
    z = rng.integers(0, 6, (nr, nc)).astype(np.float32)
    mask = (rng.random((nr, nc)) > 0.15).astype(np.uint8)
    got = orc.fill(z, mask, bc=bc)
    valid = mask != 0
    want = brute_fill(z, valid, bc[0] == "p", bc[1] == "p")
    assert np.array_equal(got[valid], want[valid])
    assert np.all(got[~valid] == np.float32(-9999.0))


def test_oracle_fill_properties():
    rng = np.random.default_rng(3)
    z = (rng.random((40, 50)) * 30).astype(np.float32)
    mask = np.ones_like(z, np.uint8)
    f = orc.fill(z, mask)
    assert np.all(f >= z)                           # P2 never lowers
    assert np.array_equal(orc.fill(f, mask), f)     # P6 idempotent
    # P1: every interior cell has a neighbour no higher than itself
    for r in range(1, 39):
        for c in range(1, 49):
            assert min(f[r + dr, c + dc] for dr, dc in D8) <= f[r, c]


def test_oracle_fill_rejects_doubly_periodic():
    z = np.zeros((4, 4), np.float32)
    with pytest.raises(orc.OracleError):
        orc.fill(z, np.ones_like(z, np.uint8), bc="pppp")
    with pytest.raises(orc.OracleError):
        orc.fill(z, np.ones_like(z, np.uint8), bc="poop")
