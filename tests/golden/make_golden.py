"""Write the golden trajectories under tests/golden/ from the CPU oracle.

Each fixture holds the final state of a seeded run computed by the oracle's
NumPy FP64 restatement (oracle.models: bve_tendency / swe_tendency and the
RK4 combination of run/bve/runner.py:211-216 and run/engine.py:406-431),
stored as the l >= m triangle, plus a checksum of the initial state so a test
can confirm it regenerated the same input.  Initial states come from the
seeded generators of aeolus_amd.run.initial_conditions (NumPy only).

    python tests/golden/make_golden.py            # all fixtures (a few minutes)
    python tests/golden/make_golden.py c2 bve85   # selected ones
"""
from __future__ import annotations

import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from aeolus_amd.run.initial_conditions import spectrum_l15, swe_spectra  # noqa: E402
from oracle.models import (Planet, SWEParams, bve_tendency, mountain, rk4_step,  # noqa: E402
                           rk4_step_array, swe_tendency)

R = 6.371e6
OMEGA = 2.0 * math.pi / 86400.0
G = 9.80616


def triangle(a: np.ndarray) -> np.ndarray:
    """l >= m entries of (..., n, n) in row-major (l, m) order."""
    n = a.shape[-1]
    li, mi = np.tril_indices(n)
    return a[..., li, mi]


def checksum(a: np.ndarray) -> float:
    return float(np.sum(np.abs(a)))


def bve_run(planet, z0, dt, nsteps):
    y = z0.copy()
    for _ in range(nsteps):
        y = rk4_step(lambda s: bve_tendency(planet, s), y, dt)
    return y


CASES = {}


def case(name):
    def deco(f):
        CASES[name] = f
        return f
    return deco


@case("c2")
def c2():
    """BVE, geodesic resolution 4 (2562 points) / product resolution 5, L = 21,
    voronoi weights, seeded l^-1.5 vorticity rms 5e-5, 96 x 900 s."""
    L = 21
    p = Planet(l_max=L, radius=R, omega=OMEGA, grid_type="geodesic", resolution=4)
    z0 = spectrum_l15(L, np.random.default_rng(21), rms=5e-5)
    return dict(z0_checksum=checksum(z0), zT=triangle(bve_run(p, z0, 900.0, 96)),
                l_max=L, dt=900.0, nsteps=96, seed=21)


@case("bve85")
def bve85():
    """BVE, Gauss 128 x 256, L = 85, seeded l^-1.5 vorticity rms 5e-5, 288 x 300 s."""
    L = 85
    p = Planet(l_max=L, radius=R, omega=OMEGA, grid_type="latlon", nlat=128, nlon=256)
    z0 = spectrum_l15(L, np.random.default_rng(85), rms=5e-5)
    return dict(z0_checksum=checksum(z0), zT=triangle(bve_run(p, z0, 300.0, 288)),
                l_max=L, dt=300.0, nsteps=288, seed=85)


@case("bve170")
def bve170():
    """BVE, Gauss 256 x 512, L = 170, seeded l^-1.5 vorticity rms 5e-5, 288 x 300 s
    (the L = 170 member class of the uneven-L ensemble; one simulated day)."""
    L = 170
    p = Planet(l_max=L, radius=R, omega=OMEGA, grid_type="latlon", nlat=256, nlon=512)
    z0 = spectrum_l15(L, np.random.default_rng(2026), rms=5e-5)
    return dict(z0_checksum=checksum(z0), zT=triangle(bve_run(p, z0, 300.0, 288)),
                l_max=L, dt=300.0, nsteps=288, seed=2026)


@case("c3")
def c3():
    """Shallow water, Gauss 128 x 256, L = 85, g = 9.80616, H = 10 km, 2 km
    Gaussian mountain (1000 km e-folding) at a seeded centre, seeded spectra
    (rms u 20 m/s, divergence 0.1 x vorticity, thickness 100 m), 288 x 300 s."""
    L = 85
    p = Planet(l_max=L, radius=R, omega=OMEGA, grid_type="latlon", nlat=128, nlon=256)
    rng = np.random.default_rng(85)
    lat0 = math.degrees(math.asin(rng.uniform(-1, 1)))
    lon0 = math.degrees(rng.uniform(0, 2 * math.pi))
    width = math.degrees(1e6 / R)
    elev = mountain(p, height_m=2000.0, lat_deg=lat0, lon_deg=lon0, width_deg=width)
    prm = SWEParams(p, gravity=G, mean_depth=1e4, elevation_lm=elev)
    y0 = swe_spectra(L, R, G, seed=85)
    y = y0.copy()
    for _ in range(288):
        y = rk4_step_array(lambda s: swe_tendency(prm, s), y, 300.0)
    return dict(y0_checksum=checksum(y0), yT=triangle(y), l_max=L, dt=300.0, nsteps=288,
                seed=85, mountain=np.array([2000.0, lat0, lon0, width]))


def main(names):
    for name in names or list(CASES):
        t0 = time.perf_counter()
        data = CASES[name]()
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **{k: np.asarray(v) for k, v in data.items()})
        print(f"{name}: {path} ({os.path.getsize(path) / 1024:.0f} KiB, "
              f"{time.perf_counter() - t0:.1f} s)", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
