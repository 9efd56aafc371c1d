"""Seeded synthetic DEMs for the LEM and routing workloads.

Spectral (power-law) noise surfaces, tilted planes, pits, closed bowls and
exactly flat plateaus, all generated from a numpy Generator seed so every
test and benchmark input is reproducible.  (The reference ships its own
generators in src/cpu/lsdtt_raster_synthetic_surfaces_cpu.*; inputs here are
shared between the GPU path and the oracle, so generator parity is not
needed.)
"""
from __future__ import annotations

import numpy as np


def spectral_noise(n_rows, n_cols, beta, std, seed):
    """Gaussian surface with radially averaged PSD ~ k^-beta, rescaled to
    the given standard deviation (float64)."""
    rng = np.random.default_rng(seed)
    white = rng.standard_normal((n_rows, n_cols))
    F = np.fft.rfft2(white)
    ky = np.fft.fftfreq(n_rows)[:, None]
    kx = np.fft.rfftfreq(n_cols)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    k[0, 0] = np.inf
    F *= k ** (-beta / 2.0)
    z = np.fft.irfft2(F, s=(n_rows, n_cols))
    z -= z.mean()
    s = z.std()
    if s > 0:
        z *= std / s
    return z


def max_centred_slope(z, dx):
    """largest centred-difference gradient magnitude over interior cells"""
    gx = (z[1:-1, 2:] - z[1:-1, :-2]) / (2 * dx)
    gy = (z[2:, 1:-1] - z[:-2, 1:-1]) / (2 * dx)
    return float(np.sqrt(gx * gx + gy * gy).max()) if gx.size else 0.0


def lem_surface(n_rows, n_cols, dx=30.0, tilt=0.01, beta=2.0, std=50.0, n_pits=200,
                seed=0, smooth_passes=0, max_slope=None):
    """Config-0-style initial surface: 0.01 * (distance north of the south
    edge) + spectral noise + single-cell pits 1-5 m deep (float64).

    `max_slope`: smooth the noise with 5-point averaging passes until the
    largest centred gradient is at most this value (a Roering hillslope with
    critical slope Sc needs a sub-critical start); the number of passes used
    is left in `lem_surface.last_passes`.
    """
    y = (n_rows - 1 - np.arange(n_rows, dtype=np.float64))[:, None] * dx
    noise = spectral_noise(n_rows, n_cols, beta, std, seed)
    passes = 0
    while True:
        z = tilt * y + noise
        if passes >= smooth_passes and (max_slope is None or max_centred_slope(z, dx) <= max_slope):
            break
        noise = 0.2 * (noise + np.roll(noise, 1, 0) + np.roll(noise, -1, 0) + np.roll(noise, 1, 1)
                       + np.roll(noise, -1, 1))
        passes += 1
        if passes > 10000:
            raise RuntimeError("could not reach the requested maximum slope")
    lem_surface.last_passes = passes
    z = np.broadcast_to(z, (n_rows, n_cols)).copy()
    rng = np.random.default_rng(seed + 1)
    if n_pits and n_rows > 2 and n_cols > 2:
        r = rng.integers(1, n_rows - 1, n_pits)
        c = rng.integers(1, n_cols - 1, n_pits)
        z[r, c] -= rng.uniform(1.0, 5.0, n_pits)
    z -= z.min()
    return z


def routing_surface(n_rows, n_cols, beta=2.2, std=100.0, depression_frac=0.01,
                    plateau_frac=0.05, max_depth=20.0, seed=0):
    """Config-1-style float32 DEM: spectral surface, closed bowls covering
    about `depression_frac` of the cells (depth U(0, max_depth]) and exactly
    flat plateaus covering about `plateau_frac` of the cells."""
    z = spectral_noise(n_rows, n_cols, beta, std, seed)
    z -= z.min()
    z += 100.0
    rng = np.random.default_rng(seed + 17)
    n = n_rows * n_cols
    yy, xx = np.ogrid[:n_rows, :n_cols]

    def stamp(frac, rad_lo, rad_hi, fn):
        covered = 0
        target = frac * n
        while covered < target:
            rad = int(rng.integers(rad_lo, rad_hi + 1))
            r = int(rng.integers(0, n_rows))
            c = int(rng.integers(0, n_cols))
            r0, r1 = max(0, r - rad), min(n_rows, r + rad + 1)
            c0, c1 = max(0, c - rad), min(n_cols, c + rad + 1)
            d2 = (yy[r0:r1] - r) ** 2 + (xx[:, c0:c1] - c) ** 2
            inside = d2 <= rad * rad
            fn(z[r0:r1, c0:c1], inside, d2, rad)
            covered += int(inside.sum())

    def bowl(sub, inside, d2, rad):
        depth = rng.uniform(0.0, max_depth)
        depth = max_depth if depth == 0.0 else depth
        sub[inside] -= depth * (1.0 - d2[inside] / float(rad * rad + 1))

    def plateau(sub, inside, d2, rad):
        sub[inside] = np.float32(sub[inside].mean())

    rmax = max(2, min(n_rows, n_cols) // 64)
    stamp(depression_frac, 2, max(2, rmax), bowl)
    stamp(plateau_frac, 2, max(3, 2 * rmax), plateau)
    return z.astype(np.float32)
