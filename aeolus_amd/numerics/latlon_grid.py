"""Gauss-Legendre x uniform-longitude grid and its HIP transform.

Mirrors ``num/latlon_grid.py:43-199``: latitude rows north to south from
``numpy.polynomial.legendre.leggauss`` (the same library call), longitudes
``2 pi k / nlon``, tensor-product weights ``w_j 2 pi / nlon``, fields flat
``(nlat*nlon,)`` in lat-major order.  The transform is the device FFT +
per-m Legendre contraction of ``libaeolus_hip`` (plan created once here).
"""
from __future__ import annotations

import ctypes
import warnings

import numpy as np
import torch

from .. import _native as nat
from .point_set_sh import _TransformBase

__all__ = ["GaussLatLonGridGeometry", "GaussLatLonSphericalHarmonics"]


class GaussLatLonGridGeometry:
    """Structured Gauss-Legendre x uniform-longitude grid (``num/latlon_grid.py:43-148``)."""

    def __init__(self, nlat: int, nlon: int, radius: float = 1.0, device=None):
        if nlat < 2 or nlon < 4:
            raise ValueError(
                f"GaussLatLonGridGeometry needs nlat >= 2 and nlon >= 4, got ({nlat}, {nlon})")
        self.nlat = int(nlat)
        self.nlon = int(nlon)
        self.radius = float(radius)
        self.device = device if device is not None else nat.default_device()
        x, w_gl = np.polynomial.legendre.leggauss(self.nlat)
        order = np.argsort(-x)
        x, w_gl = x[order], w_gl[order]
        colat = np.arccos(x)
        self._x_np = x
        self._latitudes_np = np.pi / 2.0 - colat
        self._longitudes_np = np.linspace(0.0, 2.0 * np.pi, self.nlon, endpoint=False)
        self._gl_weights = w_gl
        self._row_weights = w_gl * (2.0 * np.pi / self.nlon)
        self._coslat_rows = np.cos(self._latitudes_np)
        dev = self.device
        t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
        self._latitudes = t(self._latitudes_np)
        self._longitudes = t(self._longitudes_np)
        self._point_latitudes = t(np.repeat(self._latitudes_np, self.nlon))
        self._point_longitudes = t(np.tile(self._longitudes_np, self.nlat))
        self.coslat = torch.cos(self._point_latitudes)
        self.sinlat = torch.sin(self._point_latitudes)
        self.solid_angle_weights = t(np.repeat(w_gl, self.nlon) * (2.0 * np.pi / self.nlon))

    @property
    def latitudes(self) -> torch.Tensor:
        return self._latitudes

    @property
    def longitudes(self) -> torch.Tensor:
        return self._longitudes

    @property
    def point_latitudes(self) -> torch.Tensor:
        return self._point_latitudes

    @property
    def point_longitudes(self) -> torch.Tensor:
        return self._point_longitudes

    @property
    def n_points(self) -> int:
        return self.nlat * self.nlon

    @property
    def grid_shape(self) -> tuple[int, int]:
        return (self.nlat, self.nlon)

    def points_latlon(self) -> np.ndarray:
        return np.column_stack([np.tile(self._longitudes_np, self.nlat),
                                np.repeat(self._latitudes_np, self.nlon)])

    @property
    def cell_areas(self) -> torch.Tensor:
        return self.solid_angle_weights * self.radius ** 2

    @property
    def cfl_length_scale(self) -> float:
        colat = np.pi / 2.0 - self._latitudes_np
        return self.radius * float(np.min(np.abs(np.diff(colat))))


class GaussLatLonSphericalHarmonics(_TransformBase):
    """SH analysis/synthesis on a Gauss grid (``num/latlon_grid.py:151-199``).

    ``transform`` accepts flat ``(N,)`` or ``(nlat, nlon)`` fields (and stacks
    ``(F, N)`` / ``(F, nlat, nlon)``); ``inv_transform`` returns flat fields.
    """

    def __init__(self, grid: GaussLatLonGridGeometry, l_max: int, max_fields: int = 16):
        if not isinstance(grid, GaussLatLonGridGeometry):
            raise ValueError(
                f"GaussLatLonSphericalHarmonics requires a GaussLatLonGridGeometry, "
                f"got {type(grid).__name__}")
        self.grid = grid
        l_max = int(l_max)
        if grid.nlat < l_max + 1 or grid.nlon < 2 * l_max + 1:
            warnings.warn(
                f"Gauss lat-lon SH is under-resolved: ({grid.nlat}, {grid.nlon}) grid for "
                f"l_max={l_max}. Analysis is exact only for nlat >= {l_max + 1} and "
                f"nlon >= {2 * l_max + 1}; below that, round trips leak energy across modes.",
                stacklevel=2)
        x = np.ascontiguousarray(grid._x_np, dtype=np.float64)
        w = np.ascontiguousarray(grid._row_weights, dtype=np.float64)
        c = np.ascontiguousarray(grid._coslat_rows, dtype=np.float64)
        handle = ctypes.c_void_p()
        nat.check(nat.lib().aeo_plan_create_gauss(grid.nlat, grid.nlon, l_max, nat.host_ptr(x),
                                                  nat.host_ptr(w), nat.host_ptr(c),
                                                  int(max_fields), ctypes.byref(handle)),
                  "GaussLatLonSphericalHarmonics")
        super().__init__(handle, l_max, grid.n_points, grid.device)
        self.weights = grid.solid_angle_weights
        self.latitudes = grid.point_latitudes
        self.longitudes = grid.point_longitudes
        self.grid_shape = (grid.n_points,)
        self.coslat_rows = torch.as_tensor(c, device=grid.device)

    def enable_split(self, nslices: int) -> None:
        """Build the INT8 digit tables for split-precision contractions."""
        nat.check(nat.lib().aeo_plan_enable_split(self._plan, int(nslices)), "enable_split")
        self.nslices = int(nslices)
