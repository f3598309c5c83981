"""Grids, transforms, the backend seam and spectral operators."""
from .geodesic_grid import GeodesicGridGeometry
from .geodesic_sh import GeodesicSphericalHarmonics
from .latlon_grid import GaussLatLonGridGeometry, GaussLatLonSphericalHarmonics
from .point_set_sh import PointSetSphericalHarmonics
from .spectral_operators import SpectralOperators
from .spherical_backend import (GeodesicBackend, LatLonBackend, PointSetBackend, ProductSpace,
                                SphericalGridBackend, make_backend)

__all__ = ["GeodesicGridGeometry", "GeodesicSphericalHarmonics", "GaussLatLonGridGeometry",
           "GaussLatLonSphericalHarmonics", "PointSetSphericalHarmonics", "SpectralOperators",
           "GeodesicBackend", "LatLonBackend", "PointSetBackend", "ProductSpace",
           "SphericalGridBackend", "make_backend"]
