"""Oracle timestep control (oracle/orc_timestep.c): the reference's own checks
D1-D6 (tests/cpp/test_lem_timestep.cpp, same 60 x 60 ramp of gradient 0.6
at 5 m and the flat surface), the median against numpy's, the ladder's
rungs and hysteresis by hand, and the convergence monitor's outcomes."""
import numpy as np
import pytest

from oracle import oracle as orc

NR = NC = 60
H = 5.0
U = 4.0e-4


def ramp(slope):
    r = np.arange(NR, dtype=np.float64)[:, None].repeat(NC, 1)
    return (slope * H * r).astype(np.float32).astype(np.float64)


MASK = np.ones((NR, NC), np.uint8)


def test_d1_fires_on_the_bad_case_only():
    dem = ramp(0.6)
    good = orc.diagnose_timestep(dem, MASK, H, U, 5000.0)
    bad = orc.diagnose_timestep(dem, MASK, H, U, 20000.0)
    assert not good["too_large"]
    assert bad["too_large"]
    assert abs(bad["ratio"] - 2.676) < 0.05
    assert abs(bad["ratio"] - U * 20000.0 / (H * 0.6)) < 1e-6


def test_d2_flat_surface_says_nothing():
    d = orc.diagnose_timestep(np.zeros((NR, NC)), MASK, H, U, 1e6)
    assert not d["too_large"] and d["reference_slope"] == 0.0


def test_d3_suggestion_clears_the_warning():
    dem = ramp(0.6)
    bad = orc.diagnose_timestep(dem, MASK, H, U, 20000.0)
    redo = orc.diagnose_timestep(dem, MASK, H, U, bad["suggested_dt_yr"])
    assert not redo["too_large"]


def test_d4_d5_d6_controller():
    dem = ramp(0.6)
    tc = orc.TimestepController(20000.0)
    dt1 = tc.update(8.0, dem, MASK, H)
    assert dt1 < 20000.0 and tc.c.last_ratio > 0.7
    c = 0.5 * 3.0 / 20000.0
    dt = dt1
    for _ in range(200):
        dt = tc.update(c * dt, dem, MASK, H)
    assert tc.c.n_shrinks >= 1 and dt > dt1 and tc.c.last_ratio <= 0.7
    assert tc.update(c * dt, dem, MASK, H) == dt
    a, b = orc.TimestepController(1000.0), orc.TimestepController(1000.0)
    a.update(8.0, dem, MASK, H)
    b.update(0.1 * 3.0, dem, MASK, H)
    assert 1000.0 - a.c.dt > b.c.dt - 1000.0
    f = orc.TimestepController(1000.0)
    for _ in range(100):
        f.update(1000.0, dem, MASK, H)
    assert f.c.dt >= 1000.0 / 100.0 - 1e-9


def test_median_matches_numpy():
    rng = np.random.default_rng(3)
    for nr, nc in ((17, 23), (40, 41)):
        z = rng.normal(size=(nr, nc)).cumsum(0) * 2.0
        mask = (rng.random((nr, nc)) > 0.1).astype(np.uint8)
        g = []
        for r in range(1, nr - 1):
            for c in range(1, nc - 1):
                if mask[r, c] and mask[r - 1, c] and mask[r + 1, c] and mask[r, c - 1] and mask[r, c + 1]:
                    gx = (z[r, c + 1] - z[r, c - 1]) / (2.0 * 7.5)
                    gy = (z[r + 1, c] - z[r - 1, c]) / (2.0 * 7.5)
                    g.append(np.sqrt(gx * gx + gy * gy))
        assert orc.median_interior_gradient(z, mask, 7.5) == np.median(g)
    assert orc.median_interior_gradient(np.zeros((2, 9)), np.ones((2, 9), np.uint8), 1.0) == 0.0


def test_ladder_rungs_and_hysteresis():
    """dt lives on dt0 * 2^k; a rung is climbed only after a full rung of
    continuous growth (1.15 per comfortable step: 5 steps for x2), and the
    first overshoot drops a rung at once"""
    dem = ramp(0.6)  # cell relief 3 m
    tc = orc.TimestepController(1000.0, ratio_smoothing=1.0)
    seq = [tc.update(0.1, dem, MASK, H) for _ in range(6)]
    assert seq[:4] == [1000.0] * 4 and seq[4] == 2000.0 and seq[5] == 2000.0
    assert tc.c.n_changes == 1
    d = tc.update(30.0, dem, MASK, H)  # ratio 10: overshoot
    assert d == 1000.0 and tc.c.n_shrinks == 1


def test_monitor_converges_and_respects_the_dt_guard():
    m = orc.ConvergenceMonitor(100.0)
    states = [m.update(5e-5, 100.0) for _ in range(5)]
    assert states == [0, 0, 0, 0, 1]
    g = orc.ConvergenceMonitor(100.0)
    assert [g.update(5e-5, 10.0) for _ in range(10)] == [0] * 10  # dt collapsed: not steady
    r = orc.ConvergenceMonitor(100.0, dz_rate_tolerance_m_per_yr=1e-7)
    assert [r.update(5e-5, 100.0) for _ in range(6)][-1] == 0  # 5e-7 m/yr > 1e-7


@pytest.mark.parametrize("alternate", [False, True])
def test_monitor_stall_and_oscillation(alternate):
    m = orc.ConvergenceMonitor(100.0, stall_window=50)
    st = 0
    for i in range(60):
        dz = 0.01 * (1.5 if (alternate and i % 2) else 1.0)
        st = m.update(dz, 100.0)
    assert st == 2
    assert bool(m.m.oscillating) == alternate
