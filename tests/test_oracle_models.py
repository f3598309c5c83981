"""Pins of the oracle's models, RK4 and diagnostics (CPU).

Known answers of the reference tests: f_lm (tests/test_r5_spectral_eta.py:61),
Rossby-mode tendency (:186), RH4 one-day conservation
(tests/test_backend_parity.py:180-203), Williamson-2 initial tendencies and
one-day steadiness (tests/test_williamson2.py:100,176), rest state
(tests/test_shallow_water.py:56), SWE->BVE limit (:99), monopoles exact
(:296), lake at rest (tests/test_swe_topography.py:329), RK4 convergence
order (:487-519), diagnostics closed forms (tests/test_diagnostics.py:63-108),
planetary radius formula (pla/planetary_parameters.py:12-19).
"""
import math

import numpy as np
import pytest

from oracle.diagnostics import spectral_diagnostics
from oracle.models import (Planet, SWEParams, bve_tendency, mountain, planet_constants,
                           rk4_step, rk4_step_array, swe_tendency)

R_E = 6.371e6
OM = 7.292e-5


def test_planet_constants():
    R, om = planet_constants(24.0)
    assert abs(R - 6367354.69) < 0.01
    assert om == 2 * math.pi / 86400.0


def test_f_lm_is_single_mode():
    p = Planet(l_max=8, radius=R_E, omega=OM, grid_type="latlon", nlat=12, nlon=24)
    f = p.f_lm.copy()
    assert f[1, 0] == 2 * OM * math.sqrt(4 * math.pi / 3)
    f[1, 0] = 0
    assert np.all(f == 0)


@pytest.fixture(scope="module")
def latlon21():
    return Planet(l_max=21, radius=R_E, omega=OM, grid_type="latlon", nlat=32, nlon=64)


@pytest.mark.parametrize("l,m", [(4, 2), (6, 3)])
def test_rossby_mode(latlon21, l, m):
    z = np.zeros((22, 22), dtype=np.complex128)
    z[l, m] = 1e-5
    dz = bve_tendency(latlon21, z)
    expected = 1j * 2 * OM * m / (l * (l + 1)) * 1e-5
    assert abs(dz[l, m] / expected - 1) < 2e-2


def _rh4_coeffs(p):
    lat, lon = p.grid.point_latitudes, p.grid.point_longitudes
    nu = K = 7.848e-6
    z = 2 * nu * np.sin(lat) - 30 * K * np.sin(lat) * np.cos(lat) ** 4 * np.sin(4 * lon)
    return p.sh.transform(z)


def test_rh4_one_day_conservation(latlon21):
    p = latlon21
    z0 = _rh4_coeffs(p)
    z = z0.copy()
    for _ in range(48):
        z = rk4_step(lambda y: bve_tendency(p, y), z, 1800.0)
    d0, d1 = spectral_diagnostics(z0, R_E, OM), spectral_diagnostics(z, R_E, OM)
    assert abs(d1["energy"] - d0["energy"]) / abs(d0["energy"]) < 3e-4
    assert abs(d1["enstrophy_abs"] - d0["enstrophy_abs"]) / abs(d0["enstrophy_abs"]) < 3e-3
    assert abs(d1["circulation"] - d0["circulation"]) / math.sqrt(2 * d0["energy"]) < 1e-12


def _w2(p):
    a, om = p.R, p.Omega
    u0 = 2 * math.pi * a / (12 * 86400)
    C = a * om * u0 + 0.5 * u0 * u0
    g = 9.80616
    H = (2.94e4 - C / 3) / g
    n = p.l_max + 1
    y = np.zeros((3, n, n), dtype=np.complex128)
    y[0, 1, 0] = (2 * u0 / a) * math.sqrt(4 * math.pi / 3)
    y[2, 2, 0] = -(4 * C / 3) * math.sqrt(math.pi / 5)
    return SWEParams(p, gravity=g, mean_depth=H), y, u0, a, om


def test_williamson2_steady():
    p = Planet(l_max=15, radius=R_E, omega=7.29212e-5, grid_type="latlon", nlat=32, nlon=64)
    prm, y0, u0, a, om = _w2(p)
    dot = swe_tendency(prm, y0)
    eta = 2 * u0 / a + 2 * om
    assert np.abs(dot[0]).max() <= 1e-13 * (2 * u0 / a) * (u0 / a)
    assert np.abs(dot[1]).max() <= 1e-12 * 2 * u0 * eta / a
    assert np.abs(dot[2]).max() <= 1e-13 * prm.phi0 * 2 * u0 / a
    y = y0.copy()
    for _ in range(60):
        y = rk4_step_array(lambda s: swe_tendency(prm, s), y, 1440.0)
    assert np.linalg.norm(y[0] - y0[0]) / np.linalg.norm(y0[0]) <= 1e-12
    assert np.linalg.norm(y[2] - y0[2]) / np.linalg.norm(y0[2]) <= 1e-12
    assert np.linalg.norm(y[1]) / np.linalg.norm(y0[0]) <= 1e-12
    assert y[2, 0, 0] == 0


def test_rest_state_and_bve_limit(latlon21):
    p = latlon21
    prm = SWEParams(p, mean_depth=1e4)
    assert np.all(swe_tendency(prm, np.zeros((3, 22, 22), dtype=np.complex128)) == 0)
    rng = np.random.default_rng(5)
    z = np.zeros((22, 22), dtype=np.complex128)
    for l in range(1, 11):
        z[l, 0] = 1e-5 * rng.standard_normal()
        z[l, 1:l + 1] = 1e-5 * (rng.standard_normal(l) + 1j * rng.standard_normal(l))
    y = np.stack([z, np.zeros_like(z), np.zeros_like(z)])
    swe = swe_tendency(prm, y)[0]
    bve = bve_tendency(p, z)
    assert np.abs(swe - bve).max() <= 1e-12 * np.abs(bve).max()


def test_lake_at_rest_and_monopoles(latlon21):
    p = latlon21
    elev = mountain(p, height_m=2000.0, lat_deg=30.0, lon_deg=45.0, width_deg=20.0)
    prm = SWEParams(p, mean_depth=1e4, elevation_lm=elev)
    y = np.zeros((3, 22, 22), dtype=np.complex128)
    y[2] = -prm.phi_s_lm
    y[2, 0, 0] = 0.0
    dot = swe_tendency(prm, y)
    assert np.all(dot == 0)
    # mountain flow: monopoles stay exactly zero
    from aeolus_amd.run.initial_conditions import swe_spectra
    y = swe_spectra(21, R_E, 9.80616, seed=9)
    for _ in range(5):
        y = rk4_step_array(lambda s: swe_tendency(prm, s), y, 300.0)
    assert np.all(y[:, 0, 0] == 0)


def test_rk4_fourth_order(latlon21):
    p = latlon21
    from aeolus_amd.run.initial_conditions import spectrum_l15
    z0 = spectrum_l15(21, np.random.default_rng(1), rms=1e-4)
    T = 21600.0

    def run(n):
        z = z0.copy()
        for _ in range(n):
            z = rk4_step(lambda y: bve_tendency(p, y), z, T / n)
        return z

    ref = run(96)
    e1 = np.linalg.norm(run(6) - ref)
    e2 = np.linalg.norm(run(12) - ref)
    assert e1 / e2 > 8.0


def test_diagnostics_closed_forms():
    L = 5
    z = np.zeros((L + 1, L + 1), dtype=np.complex128)
    z[2, 1] = 0.3 + 0.4j
    d = spectral_diagnostics(z, R_E, 0.0)
    P = 2 * 0.25
    assert d["enstrophy_rel"] == pytest.approx(0.5 * R_E ** 2 * P, rel=1e-14)
    assert d["energy"] == pytest.approx(0.5 * R_E ** 4 * P / 6, rel=1e-14)
    z = np.zeros_like(z)
    z[0, 0] = 2.0
    assert spectral_diagnostics(z, R_E, 0.0)["circulation"] == pytest.approx(
        R_E ** 2 * math.sqrt(4 * math.pi) * 2.0, rel=1e-15)
    z = np.zeros_like(z)
    d = spectral_diagnostics(z, R_E, OM)
    f10 = 2 * OM * math.sqrt(4 * math.pi / 3)
    assert d["enstrophy_abs"] == pytest.approx(0.5 * R_E ** 2 * f10 ** 2, rel=1e-14)
    z[0, 3] = 9.0   # padding is ignored
    assert spectral_diagnostics(z, R_E, OM)["enstrophy_rel"] == 0.0
