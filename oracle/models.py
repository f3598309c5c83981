"""Model restatements (oracle, test infrastructure only).

* :class:`Planet` bundles state transform, product transform, product
  ``coslat`` and operators the way ``pla/planet.py:192-206`` wires
  ``Planet.generate`` (geometry -> transform -> backend -> operators with
  ``product_quadrature="fine"``); planetary constants follow
  ``pla/planetary_parameters.py:12-55``.
* :func:`bve_tendency` follows ``phy/barotropic.py:60-167``.
* :func:`swe_tendency` follows ``phy/shallow_water.py:172-418``.
* :func:`rk4_step` follows ``run/bve/runner.py:211-216`` and
  :func:`rk4_step_array` follows ``run/engine.py:406-431``.
* :func:`mountain` follows ``phy/topography.py:207-290``.

Pinned by (tests/test_oracle_models.py): Rossby-Haurwitz and Rossby-mode
known answers (``tests/test_r5_spectral_eta.py:61,186``), RH4 one-day
conservation bands (``tests/test_backend_parity.py:180-203``), the
Williamson-2 steady state (``tests/test_williamson2.py:100,176``), rest state,
lake at rest and SWE-to-BVE limit (``tests/test_shallow_water.py:56,99``,
``tests/test_swe_topography.py:329``), exact monopole conservation
(``tests/test_shallow_water.py:296``) and RK4 fourth-order convergence.
"""
from __future__ import annotations

import math

import numpy as np

from .grids import COSLAT_FLOOR, GaussGrid, GeodesicGrid, product_grid_dims
from .operators import SpectralOps
from .transforms import DenseTransform, GaussTransform

__all__ = [
    "G_NEWTON",
    "planet_constants",
    "Planet",
    "bve_tendency",
    "swe_tendency",
    "rk4_step",
    "rk4_step_array",
    "mountain",
    "WILLIAMSON_GRAVITY",
]

G_NEWTON = 6.67430e-11            # scipy.constants.G (CODATA 2018)
EARTH_MASS = 5.972e24             # pla/planetary_parameters.py:45
EARTH_RADIUS = 6.371e6            # pla/planetary_parameters.py:46
WILLIAMSON_GRAVITY = 9.80616


def planet_constants(day_hours: float = 24.0, mass_earth_units: float = 1.0,
                     radius_earth_units: float = 1.0) -> tuple[float, float]:
    """``(radius, angular_velocity)`` per ``pla/planetary_parameters.py:12-19``.

    The radius is the volumetric-mean radius of the first-order oblate
    spheroid, ``cbrt(R_eq^3 (1 - f))`` with ``f = Omega^2 R_eq^3/(2 G M)``
    clipped to [0, 0.5].
    """
    M = mass_earth_units * EARTH_MASS
    Req = radius_earth_units * EARTH_RADIUS
    omega = 2.0 * np.pi / (day_hours * 3600.0)
    f = float(np.clip(omega ** 2 * Req ** 3 / (2.0 * G_NEWTON * M), 0.0, 0.5))
    volume = 4.0 / 3.0 * np.pi * Req ** 3 * (1.0 - f)
    radius = float(np.cbrt(volume / (4.0 / 3.0 * np.pi)))
    return radius, float(omega)


class Planet:
    """State transform + product transform + operators (oracle wiring)."""

    def __init__(self, *, l_max: int, radius: float, omega: float,
                 grid_type: str = "latlon", nlat: int = 0, nlon: int = 0,
                 resolution: int = 0, product_quadrature: str = "fine",
                 dense: bool = False, basis: str = "stable"):
        self.l_max = int(l_max)
        self.R = float(radius)
        self.Omega = float(omega)
        if grid_type == "latlon":
            g = GaussGrid(nlat, nlon, radius)
            self.grid = g
            self.sh = (DenseTransform(g.point_latitudes, g.point_longitudes, l_max,
                                      g.weights, basis=basis, grid_shape=g.grid_shape)
                       if dense else GaussTransform(g, l_max))
            if product_quadrature == "fine":
                nf, mf = product_grid_dims(l_max, nlat, nlon)
                gp = GaussGrid(nf, mf, radius)
                self.pgrid = gp
                self.psh = (DenseTransform(gp.point_latitudes, gp.point_longitudes,
                                           l_max, gp.weights, basis=basis)
                            if dense else GaussTransform(gp, l_max))
            else:
                self.pgrid, self.psh = g, self.sh
        elif grid_type == "geodesic":
            g = GeodesicGrid(resolution, radius)
            self.grid = g
            self.sh = DenseTransform(g.latitudes, g.longitudes, l_max,
                                     g.voronoi_weights, basis=basis)
            if product_quadrature == "fine":
                gp = GeodesicGrid(resolution + 1, radius)
                self.pgrid = gp
                self.psh = DenseTransform(gp.latitudes, gp.longitudes, l_max,
                                          gp.voronoi_weights, basis=basis)
            else:
                self.pgrid, self.psh = g, self.sh
        else:
            raise ValueError(f"grid_type {grid_type!r}")
        self.pcoslat = np.maximum(np.cos(self.pgrid.point_latitudes), COSLAT_FLOOR)
        self.scoslat = np.maximum(np.cos(self.grid.point_latitudes), COSLAT_FLOOR)
        self.so = SpectralOps(l_max, radius)
        self.f_lm = np.zeros((l_max + 1, l_max + 1), dtype=np.complex128)
        self.f_lm[1, 0] = 2.0 * self.Omega * math.sqrt(4.0 * math.pi / 3.0)


# ---------------------------------------------------------------------------
# Barotropic vorticity (phy/barotropic.py)
# ---------------------------------------------------------------------------

def bve_tendency(planet: Planet, zeta: np.ndarray, nu: float = 0.0,
                 forcing: np.ndarray | None = None) -> np.ndarray:
    """``d zeta/dt`` per ``phy/barotropic.py:120-167``."""
    R = planet.R
    l = np.arange(planet.l_max + 1, dtype=np.float64)
    lap = -l * (l + 1.0) / R ** 2
    lap[0] = 1.0 / R ** 2                                  # :66
    psi = np.zeros_like(zeta)
    psi[1:, :] = zeta[1:, :] / lap[1:, None]                # :91
    eta = zeta + planet.f_lm                                # :147
    adv = -planet.so.jacobian(psi, eta, planet.psh, planet.pcoslat)   # :153
    diff = nu * lap[:, None] * zeta                         # :158
    out = adv + diff
    if forcing is not None:
        out = out + forcing
    out[0, :] = 0.0                                         # :165
    return out


def rk4_step(tendency, y: np.ndarray, dt: float) -> np.ndarray:
    """``run/bve/runner.py:211-216`` (same arithmetic order)."""
    k1 = tendency(y)
    k2 = tendency(y + 0.5 * dt * k1)
    k3 = tendency(y + 0.5 * dt * k2)
    k4 = tendency(y + dt * k3)
    return y + (dt / 6.0) * (k1 + 2 * k2 + 2 * k3 + k4)


def rk4_step_array(tendency, y: np.ndarray, dt: float, stage_validator=None) -> np.ndarray:
    """``run/engine.py:406-431`` (same arithmetic order)."""
    def stage(ys):
        if stage_validator is not None:
            stage_validator(ys)
        return ys
    k1 = tendency(y)
    k2 = tendency(stage(y + 0.5 * dt * k1))
    k3 = tendency(stage(y + 0.5 * dt * k2))
    k4 = tendency(stage(y + dt * k3))
    return y + (dt / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4)


# ---------------------------------------------------------------------------
# Shallow water (phy/shallow_water.py)
# ---------------------------------------------------------------------------

class SWEParams:
    def __init__(self, planet: Planet, *, gravity: float = WILLIAMSON_GRAVITY,
                 mean_depth: float, nu4: float = 0.0, elevation_lm=None):
        self.planet = planet
        self.gravity = float(gravity)
        self.mean_depth = float(mean_depth)
        self.phi0 = self.gravity * self.mean_depth
        self.nu4 = float(nu4)
        L = planet.l_max
        l = np.arange(L + 1, dtype=np.float64)
        self.lap_eigs = -l * (l + 1.0) / planet.R ** 2
        if elevation_lm is not None and np.any(elevation_lm != 0):
            self.phi_s_lm = self.gravity * elevation_lm
            anom = self.phi_s_lm.copy()
            anom[0, 0] = 0.0
            self.phi_s_anom = anom
            self.lap_phi_s = self.lap_eigs[:, None] * self.phi_s_lm
            self.nu4_lap2_phi_s = (self.nu4 * self.lap_eigs[:, None] ** 2 * anom
                                   if self.nu4 > 0 else None)
        else:
            self.phi_s_lm = None
            self.lap_phi_s = None
            self.nu4_lap2_phi_s = None

    def inv_lap(self, a):
        out = np.zeros_like(a)
        out[1:, :] = a[1:, :] / self.lap_eigs[1:, None]
        return out


def swe_tendency(p: SWEParams, y: np.ndarray) -> np.ndarray:
    """``d/dt`` of the ``(3, L+1, L+1)`` stack per ``phy/shallow_water.py:347-418``."""
    pl = p.planet
    so, sh, c, R = pl.so, pl.psh, pl.pcoslat, pl.R
    zeta, delta, phi = y[0], y[1], y[2]
    psi = p.inv_lap(zeta)
    chi = p.inv_lap(delta)
    eta = zeta + pl.f_lm

    def deriv(a):
        return (sh.inv_transform(so.d_lambda(a)),
                sh.inv_transform(so.sin_theta_d_theta(a)) / R)

    psi_lam, psi_snt = deriv(psi)
    chi_lam, chi_snt = deriv(chi)
    u = (psi_snt + chi_lam) / c
    v = (psi_lam - chi_snt) / c
    eta_lam, eta_snt = deriv(eta)
    phi_lam, phi_snt = deriv(phi)
    eta_g = sh.inv_transform(eta)
    zeta_g = sh.inv_transform(zeta)
    delta_g = sh.inv_transform(delta)
    phi_g = sh.inv_transform(phi)

    adv_eta = (u * eta_lam - v * eta_snt) / c
    zeta_dot_g = -(adv_eta + eta_g * delta_g)
    curl_g = eta_g * zeta_g + (v * eta_lam + u * eta_snt) / c
    kin_g = 0.5 * (u * u + v * v)
    adv_phi = (u * phi_lam - v * phi_snt) / c
    phi_dot_g = -(adv_phi + phi_g * delta_g)

    zeta_dot = so.truncate(sh.transform(zeta_dot_g))
    curl_c = so.truncate(sh.transform(curl_g))
    kin_c = so.truncate(sh.transform(kin_g))
    phi_dot = so.truncate(sh.transform(phi_dot_g)) - p.phi0 * delta

    delta_dot = curl_c - p.lap_eigs[:, None] * (kin_c + phi)
    if p.lap_phi_s is not None:
        delta_dot = delta_dot - p.lap_phi_s
    out = np.stack([zeta_dot, delta_dot, phi_dot])
    if p.nu4 > 0.0:
        out = out - p.nu4 * (p.lap_eigs[None, :, None] ** 2) * y
        if p.nu4_lap2_phi_s is not None:
            out[2] = out[2] - p.nu4_lap2_phi_s
    out[:, 0, :] = 0.0
    return out


def mountain(planet: Planet, *, height_m: float, lat_deg: float, lon_deg: float,
             width_deg: float, l_cut: int | None = None,
             return_residual: bool = False):
    """Band-limited Gaussian mountain (``phy/topography.py:207-290``).

    ``l_cut`` zeroes every coefficient of degree or order above it before the
    0.2 projection-residual gate (``:263-266``); with ``return_residual`` the
    relative residual is returned beside the coefficients.
    """
    lat = planet.grid.point_latitudes
    lon = planet.grid.point_longitudes
    lat0, lon0, sig = map(math.radians, (lat_deg, lon_deg, width_deg))
    cosd = math.sin(lat0) * np.sin(lat) + math.cos(lat0) * np.cos(lat) * np.cos(lon - lon0)
    d = np.arccos(np.clip(cosd, -1.0, 1.0))
    analytic = height_m * np.exp(-(d / sig) ** 2)
    coeffs = planet.sh.transform(analytic)
    if l_cut is not None:
        coeffs[int(l_cut) + 1:, :] = 0.0
        coeffs[:, int(l_cut) + 1:] = 0.0
    synth = planet.sh.inv_transform(coeffs)
    w = planet.sh.weights
    rel = math.sqrt(np.sum(w * (synth - analytic) ** 2)) / math.sqrt(np.sum(w * analytic ** 2))
    if not rel <= 0.2:
        raise ValueError(f"mountain projection residual {rel:.3f} exceeds 0.2")
    return (coeffs, rel) if return_residual else coeffs
