"""Pins of the oracle's diagnostics rows (CPU): one row of
``DiagnosticsRecorder.record`` (run/bve/diagnostics.py:162-211) and of
``SWEDiagnosticsRecorder.record`` (run/swe/diagnostics.py:113-173) against
closed forms: solid-body rotation for the BVE row, the resting layer for the
SWE row."""
import math

import numpy as np

from oracle.diagnostics import bve_record_row, swe_record_row
from oracle.models import Planet, SWEParams

R = 6.371e6
OMEGA = 2.0 * math.pi / 86400.0


def _rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


def test_bve_row_solid_body_rotation():
    # zeta = (2U/R) sin(lat) <=> u = U cos(lat), v = 0
    U, L = 20.0, 10
    p = Planet(l_max=L, radius=R, omega=OMEGA, grid_type="latlon", nlat=16, nlon=32)
    z = np.zeros((L + 1, L + 1), dtype=np.complex128)
    z[1, 0] = (2.0 * U / R) * math.sqrt(4.0 * math.pi / 3.0)
    dt, ls = 600.0, 1.0e5
    row = bve_record_row(p, z, dt, ls)
    lat = p.grid.point_latitudes
    smax = U * float(np.max(np.cos(lat)))
    assert _rel(row["max_speed_ms"], smax) < 1e-12
    assert _rel(row["cfl"], smax * dt / ls) < 1e-12
    assert _rel(row["energy"], 4.0 * math.pi / 3.0 * U * U * R * R) < 1e-13
    assert _rel(row["energy_l1"], row["energy"]) < 1e-15
    assert _rel(row["enstrophy_rel"], 8.0 * math.pi / 3.0 * U * U) < 1e-13
    assert abs(row["circulation"]) == 0.0
    assert _rel(row["zeta_max"], 2.0 * U / R * float(np.max(np.abs(np.sin(lat))))) < 1e-12
    # Gauss quadrature integrates sin^2 exactly: rms = (2U/R)/sqrt(3)
    assert _rel(row["zeta_rms"], 2.0 * U / R / math.sqrt(3.0)) < 1e-12
    assert row["high_l_enstrophy_frac"] == 0.0
    assert row["roundtrip_residual"] < 1e-13


def test_bve_row_absolute_enstrophy_and_padding():
    L = 8
    p = Planet(l_max=L, radius=R, omega=OMEGA, grid_type="latlon", nlat=12, nlon=24)
    z = np.zeros((L + 1, L + 1), dtype=np.complex128)
    z[0, 3] = 5.0   # padding (m > l) is ignored by every spectral number
    row = bve_record_row(p, z, 0.0, 1.0)
    f10 = 2.0 * OMEGA * math.sqrt(4.0 * math.pi / 3.0)
    assert _rel(row["enstrophy_abs"], 0.5 * R * R * f10 * f10) < 1e-14
    assert row["enstrophy_rel"] == 0.0 and row["energy"] == 0.0
    assert math.isnan(row["cfl"])   # dt = 0


def test_swe_row_resting_layer():
    L, H, g = 10, 1.0e4, 9.80616
    p = Planet(l_max=L, radius=R, omega=OMEGA, grid_type="latlon", nlat=16, nlon=32)
    sp = SWEParams(p, gravity=g, mean_depth=H)
    y = np.zeros((3, L + 1, L + 1), dtype=np.complex128)
    row = swe_record_row(sp, y, 300.0, 1.0e5)
    phi0 = g * H
    assert row["max_wind_ms"] == 0.0
    assert row["phi_total_min"] == phi0 and row["phi_total_max"] == phi0
    assert _rel(row["max_char_speed_ms"], math.sqrt(phi0)) < 1e-15
    assert _rel(row["total_mass"], 4.0 * math.pi * R * R * phi0) < 1e-15
    assert _rel(row["total_energy"], 0.5 * phi0 * phi0 * 4.0 * math.pi * R * R) < 1e-13
    assert _rel(row["h_min_m"], H) < 1e-15
    assert row["eta_min_m"] == 0.0 and row["eta_max_m"] == 0.0
    assert row["zeta_l2"] == 0.0 and row["delta_l2"] == 0.0 and row["phi_l2"] == 0.0


def test_swe_row_geopotential_mode():
    # phi' = A sqrt(4 pi) Y_00-free degree-1 mode: phi = c sin(lat); its mass
    # contribution vanishes, its extrema follow the grid, its L2 norm is exact
    L, H, g = 6, 5.0e3, 9.80616
    p = Planet(l_max=L, radius=R, omega=OMEGA, grid_type="latlon", nlat=12, nlon=24)
    sp = SWEParams(p, gravity=g, mean_depth=H)
    y = np.zeros((3, L + 1, L + 1), dtype=np.complex128)
    c = 100.0
    y[2, 1, 0] = c * math.sqrt(4.0 * math.pi / 3.0)
    row = swe_record_row(sp, y, 0.0, None)
    phi0 = g * H
    s = np.sin(p.grid.point_latitudes)
    assert _rel(row["phi_total_max"], phi0 + c * float(s.max())) < 1e-13
    assert _rel(row["eta_max_m"], c * float(s.max()) / g) < 1e-12
    assert _rel(row["total_mass"], 4.0 * math.pi * R * R * phi0) < 1e-15
    # integral (c sin)^2 dA = R^2 c^2 4 pi / 3
    assert _rel(row["phi_l2"], math.sqrt(R * R * c * c * 4.0 * math.pi / 3.0)) < 1e-14
