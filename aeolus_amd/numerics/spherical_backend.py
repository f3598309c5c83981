"""Backend seam: geometry + transform + product-space policy.

Mirrors ``num/spherical_backend.py:43-241``: ``ProductSpace(sh, coslat,
geometry, label)``; ``coarse`` (mandatory, the state sampling) and ``fine``
product modes built once and cached; unsupported modes raise ``ValueError``
with "no silent fallback"; ``coslat`` clamped at 1e-8.  The lat-lon ``fine``
space is the 3/2-rule Gauss grid ``max(3L//2+1, nlat) x max(3L+1, nlon)``;
the geodesic one is the resolution+1 grid with Voronoi weights.
"""
from __future__ import annotations

from abc import ABC, abstractmethod
from dataclasses import dataclass
from typing import Any

import torch

from .geodesic_grid import GeodesicGridGeometry
from .geodesic_sh import GeodesicSphericalHarmonics
from .latlon_grid import GaussLatLonGridGeometry, GaussLatLonSphericalHarmonics

__all__ = ["ProductSpace", "SphericalGridBackend", "GeodesicBackend", "LatLonBackend",
           "PointSetBackend", "make_backend"]

_COSLAT_FLOOR = 1e-8


@dataclass(frozen=True)
class ProductSpace:
    sh: Any
    coslat: torch.Tensor
    geometry: Any
    label: str


class SphericalGridBackend(ABC):
    def __init__(self, geometry, sh):
        self.geometry = geometry
        self.sh = sh
        self.l_max = sh.l_max
        self._product_spaces: dict[str, ProductSpace] = {}

    @abstractmethod
    def supported_product_quadratures(self) -> tuple[str, ...]:
        """Modes accepted by product_space(); must include 'coarse'."""

    def product_space(self, mode: str) -> ProductSpace:
        if mode not in self.supported_product_quadratures():
            raise ValueError(
                f"product_quadrature must be one of {self.supported_product_quadratures()}, "
                f"got {mode!r} (backend {type(self).__name__}; no silent fallback)")
        if mode not in self._product_spaces:
            self._product_spaces[mode] = self._build_product_space(mode)
        return self._product_spaces[mode]

    @abstractmethod
    def _build_product_space(self, mode: str) -> ProductSpace:
        """Construct the product space for a supported mode (called once)."""

    def describe(self, product_quadrature: str) -> dict:
        return {
            "backend": type(self).__name__,
            "grid": type(self.geometry).__name__,
            "state_sampling": self.product_space("coarse").label,
            "product_quadrature": product_quadrature,
            "product_sampling": self.product_space(product_quadrature).label,
            "transform": type(self.sh).__name__,
            "l_max": int(self.l_max),
        }

    def _coarse_coslat(self) -> torch.Tensor:
        coslat = getattr(self.geometry, "coslat", None)
        if coslat is None:
            coslat = torch.cos(torch.as_tensor(self.geometry.point_latitudes))
        return torch.clamp(torch.as_tensor(coslat), min=_COSLAT_FLOOR)


class GeodesicBackend(SphericalGridBackend):
    def __init__(self, geometry, sh):
        if not isinstance(geometry, GeodesicGridGeometry):
            raise ValueError(f"GeodesicBackend requires a GeodesicGridGeometry, got "
                             f"{type(geometry).__name__}")
        super().__init__(geometry, sh)

    def supported_product_quadratures(self) -> tuple[str, ...]:
        return ("coarse", "fine")

    def _build_product_space(self, mode: str) -> ProductSpace:
        if mode == "coarse":
            return ProductSpace(self.sh, self._coarse_coslat(), None,
                                f"geodesic-res{self.geometry.resolution}-state")
        fine = GeodesicGridGeometry(self.geometry.resolution + 1, self.geometry.radius,
                                    device=self.geometry.device)
        fine_sh = GeodesicSphericalHarmonics(fine, self.l_max, weights="voronoi")
        return ProductSpace(fine_sh, torch.clamp(fine.coslat, min=_COSLAT_FLOOR), fine,
                            f"geodesic-res{fine.resolution}-voronoi")


class LatLonBackend(SphericalGridBackend):
    def __init__(self, geometry, sh):
        if not isinstance(geometry, GaussLatLonGridGeometry):
            raise ValueError(f"LatLonBackend requires a GaussLatLonGridGeometry, got "
                             f"{type(geometry).__name__}")
        super().__init__(geometry, sh)

    def supported_product_quadratures(self) -> tuple[str, ...]:
        return ("coarse", "fine")

    def _build_product_space(self, mode: str) -> ProductSpace:
        g = self.geometry
        if mode == "coarse":
            return ProductSpace(self.sh, self._coarse_coslat(), None,
                                f"latlon-gauss-{g.nlat}x{g.nlon}-state")
        nlat_f = max((3 * self.l_max) // 2 + 1, g.nlat)
        nlon_f = max(3 * self.l_max + 1, g.nlon)
        if (nlat_f, nlon_f) == (g.nlat, g.nlon):
            # Same dimensions: the state transform already is the product
            # transform (numerically identical to building a second one).
            fine, fine_sh = g, self.sh
        else:
            fine = GaussLatLonGridGeometry(nlat_f, nlon_f, g.radius, device=g.device)
            fine_sh = GaussLatLonSphericalHarmonics(fine, self.l_max)
        return ProductSpace(fine_sh, torch.clamp(fine.coslat, min=_COSLAT_FLOOR), fine,
                            f"latlon-gauss-{nlat_f}x{nlon_f}-3/2rule")


class PointSetBackend(SphericalGridBackend):
    def supported_product_quadratures(self) -> tuple[str, ...]:
        return ("coarse",)

    def _build_product_space(self, mode: str) -> ProductSpace:
        return ProductSpace(self.sh, self._coarse_coslat(), None, "point-set-state")


def make_backend(geometry, sh) -> SphericalGridBackend:
    if isinstance(geometry, GeodesicGridGeometry):
        return GeodesicBackend(geometry, sh)
    if isinstance(geometry, GaussLatLonGridGeometry):
        return LatLonBackend(geometry, sh)
    return PointSetBackend(geometry, sh)
