"""Planet wiring: geometry -> transform -> backend -> operators.

Mirrors ``PlanetaryParameters`` (``pla/planetary_parameters.py:6-62``) and
the part of ``Planet.generate`` (``pla/planet.py:155-206``) that builds the
grid, the transform, the backend and ``SpectralOperators``.  The decorative
random terrain the reference also draws never enters the dynamics and is
not built here.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .numerics.geodesic_grid import GeodesicGridGeometry
from .numerics.geodesic_sh import GeodesicSphericalHarmonics
from .numerics.latlon_grid import GaussLatLonGridGeometry, GaussLatLonSphericalHarmonics
from .numerics.spectral_operators import SpectralOperators
from .numerics.spherical_backend import GeodesicBackend, LatLonBackend

__all__ = ["PlanetaryParameters", "Planet", "G_NEWTON", "EARTH_MASS", "EARTH_RADIUS"]

G_NEWTON = 6.6743e-11      # scipy.constants.G (CODATA 2018), as the reference imports it
EARTH_MASS = 5.972e24
EARTH_RADIUS = 6.371e6


@dataclass
class PlanetaryParameters:
    mass: float
    equatorial_radius: float
    sidereal_day: float
    angular_velocity: float = field(init=False)
    oblateness: float = field(init=False)
    polar_radius: float = field(init=False)
    volume: float = field(init=False)
    radius: float = field(init=False)
    density: float = field(init=False)

    def __post_init__(self):
        self.angular_velocity = 2 * np.pi / self.sidereal_day
        R, M, w = self.equatorial_radius, self.mass, self.angular_velocity
        self.oblateness = float(np.clip((w ** 2 * R ** 3) / (2 * G_NEWTON * M), 0.0, 0.5))
        self.polar_radius = self.equatorial_radius * (1 - self.oblateness)
        self.volume = 4 / 3 * np.pi * self.equatorial_radius ** 3 * (1 - self.oblateness)
        self.radius = float(np.cbrt(self.volume / (4 / 3 * np.pi)))
        self.density = self.mass / self.volume

    @classmethod
    def from_earth_like(cls, mass_earth_units: float = 1.0, radius_earth_units: float = 1.0,
                        day_hours: float = 24.0) -> "PlanetaryParameters":
        return cls(mass=mass_earth_units * EARTH_MASS,
                   equatorial_radius=radius_earth_units * EARTH_RADIUS,
                   sidereal_day=day_hours * 3600.0)

    @classmethod
    def from_si(cls, mass_kg, eqt_radius_m, day_s) -> "PlanetaryParameters":
        return cls(mass=mass_kg, equatorial_radius=eqt_radius_m, sidereal_day=day_s)


class _ExplicitParams:
    """Parameters with an explicitly chosen dynamics radius and rotation rate."""

    def __init__(self, radius: float, angular_velocity: float):
        self.radius = float(radius)
        self.angular_velocity = float(angular_velocity)


class Planet:
    def __init__(self, params, grid, sh, so):
        self.params = params
        self.grid = grid
        self.sh = sh
        self.so = so

    @classmethod
    def generate(cls, params, grid_resolution: int = 3, l_max: int = 15,
                 product_quadrature: str = "fine", grid_type: str = "geodesic",
                 nlat: int = 128, nlon: int = 256, device=None, **_ignored) -> "Planet":
        if grid_type == "geodesic":
            grid = GeodesicGridGeometry(grid_resolution, params.radius, device=device)
            sh = GeodesicSphericalHarmonics(grid, l_max)
            backend = GeodesicBackend(grid, sh)
        elif grid_type == "latlon":
            grid = GaussLatLonGridGeometry(nlat, nlon, params.radius, device=device)
            sh = GaussLatLonSphericalHarmonics(grid, l_max)
            backend = LatLonBackend(grid, sh)
        else:
            raise ValueError(f"grid_type must be 'geodesic' or 'latlon', got {grid_type!r} "
                             "(no silent fallback)")
        so = SpectralOperators(sh, params.radius, grid, product_quadrature=product_quadrature,
                               backend=backend)
        return cls(params, grid, sh, so)

    @classmethod
    def with_radius(cls, radius: float, angular_velocity: float, **kw) -> "Planet":
        """``generate`` with an explicit dynamics radius (for example 6.371e6 m)."""
        return cls.generate(_ExplicitParams(radius, angular_velocity), **kw)
