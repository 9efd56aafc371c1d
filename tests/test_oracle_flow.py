"""The CPU oracle's D8 flow object against the reference's golden FNV-1a pins
(tests/cpp/test_flow_d8_identity.cpp:517-588): 17 integer-terrain fixtures x
4 option configurations, every component hashed exactly as the reference
test hashes it."""
import os

import numpy as np
import pytest

from d8_fixtures import CONFIGS, COMPONENTS, SPECS, Fixture, hash_flow, load_pins, weights_for
from oracle import oracle as orc

PINS = load_pins(os.path.join(os.path.dirname(__file__), "golden", "d8_identity_pins.txt"))
FIXTURES = {s[0]: Fixture(s) for s in SPECS}


def oracle_row(fx, cfg):
    keep_topo, sv = CONFIGS[cfg]
    fl = orc.flow_build(fx.dem, fx.mask, fx.dx, fx.bc, fx.nodata_is_outlet,
                       fx.pits, fx.forced, topo=True, svector=sv)
    n = fx.nr * fx.nc
    m = fx.mask.ravel()
    if sv:
        area = np.where(m != 0, (fl["s_size"].astype(np.float64) * (fx.dx * fx.dx)).astype(np.float32),
                        np.float32(fx.nodata))
    else:
        acc = orc.accumulate(m, fl["receiver"], fl["topo"], fx.dx)
        area = np.where(m != 0, acc.astype(np.float32), np.float32(fx.nodata))
    amask = (m != 0).astype(np.uint8)
    weighted = orc.accumulate(m, fl["receiver"], fl["topo"], fx.dx, weights_for(n)) if keep_topo else None
    return hash_flow(fx, cfg, fl, area, amask, weighted)


@pytest.mark.parametrize("name", [s[0] for s in SPECS])
@pytest.mark.parametrize("cfg", list(CONFIGS))
def test_oracle_matches_reference_pins(name, cfg):
    fx = FIXTURES[name]
    got = oracle_row(fx, cfg)
    want = PINS[f"{name}/{cfg}"]
    bad = [f"{COMPONENTS[k]}: got {got[k]:#018x} want {want[k]:#018x}"
           for k in range(8) if got[k] != want[k]]
    assert not bad, "; ".join(bad)


def test_fixtures_exercise_tie_break():
    """test_flow_d8_identity.cpp:484-510: fixtures >= 400 cells have tied
    steepest slopes, so the strict `>` first-wins rule is exercised."""
    for fx in FIXTURES.values():
        if fx.nr * fx.nc < 400:
            continue
        z = fx.dem.astype(np.float64)
        tied = 0
        for r in range(fx.nr):
            for c in range(fx.nc):
                if not fx.mask[r, c]:
                    continue
                best, nb = 0.0, 0
                for d in range(8):
                    k, rr, cc = fx._step(r, c, d)
                    if k != 1 or not fx.mask[rr, cc]:
                        continue
                    dz = z[r, c] - z[rr, cc]
                    if dz <= 0:
                        continue
                    s = dz / (fx.dx if d % 2 == 0 else fx.dx * 1.4142135623730950488)
                    if s > best:
                        best, nb = s, 1
                    elif s == best:
                        nb += 1
                if nb >= 2:
                    tied += 1
        assert tied > 0, fx.name
