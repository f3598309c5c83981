"""Grid geometry restatements (oracle, test infrastructure only).

* :func:`gauss_latlon` follows ``num/latlon_grid.py:50-83`` (NumPy
  ``leggauss`` nodes sorted north to south, uniform longitudes
  ``2 pi k / nlon``, tensor weights ``w_j 2 pi / nlon``, lat-major flattening).
* :func:`product_grid_dims` follows ``num/spherical_backend.py:401-402``.
* :func:`geodesic` follows ``num/geodesic_grid.py:15-266`` and
  ``num/cartesian_to_spherical.py:3-33``: unit icosahedron, 0.01 rad rotation
  about x (closed-form rotation matrix instead of SciPy ``Rotation``),
  midpoint subdivision with a shared-edge dictionary, ``atan2`` longitude and
  latitude, L'Huilier spherical excess split one third per vertex; the
  Voronoi weights are ``area / R^2`` renormalised to sum to 4 pi
  (``num/optimized_geodesic_sh.py:127-133``).

Pinned by (tests/test_oracle_grids.py): sum of Gauss weights = 4 pi to 1e-13
(``tests/test_latlon_transform.py:51``), 3/2-rule sizes
(``tests/test_latlon_backend.py:90-101``), geodesic point count
``10*4^res + 2``, unit-norm points, areas summing to 4 pi R^2, and the
rotation checked against SciPy's ``Rotation.from_rotvec`` (the library call
the reference makes).
"""
from __future__ import annotations

import math

import numpy as np

__all__ = [
    "GaussGrid",
    "gauss_latlon",
    "product_grid_dims",
    "GeodesicGrid",
    "geodesic",
    "COSLAT_FLOOR",
]

COSLAT_FLOOR = 1e-8  # num/spherical_backend.py:72


class GaussGrid:
    """Gauss-Legendre x uniform-longitude grid (``num/latlon_grid.py:43-148``)."""

    def __init__(self, nlat: int, nlon: int, radius: float = 1.0):
        if nlat < 2 or nlon < 4:
            raise ValueError("GaussLatLonGridGeometry needs nlat >= 2 and nlon >= 4")
        self.nlat = int(nlat)
        self.nlon = int(nlon)
        self.radius = float(radius)
        x, w = np.polynomial.legendre.leggauss(self.nlat)   # :60
        order = np.argsort(-x)                                # :61
        x, w = x[order], w[order]
        colat = np.arccos(x)
        self.x = x                                            # cos(colat)
        self.gl_weights = w
        self.latitudes = np.pi / 2.0 - colat                  # :64
        self.longitudes = np.linspace(0.0, 2.0 * np.pi, self.nlon, endpoint=False)
        self.point_latitudes = np.repeat(self.latitudes, self.nlon)   # :73
        self.point_longitudes = np.tile(self.longitudes, self.nlat)   # :75
        self.coslat = np.cos(self.point_latitudes)
        self.sinlat = np.sin(self.point_latitudes)
        self.weights = np.repeat(w, self.nlon) * (2.0 * np.pi / self.nlon)  # :82
        self.n_points = self.nlat * self.nlon
        self.grid_shape = (self.nlat, self.nlon)

    @property
    def cfl_length_scale(self) -> float:
        colat = np.pi / 2.0 - self.latitudes
        return self.radius * float(np.min(np.abs(np.diff(colat))))


def gauss_latlon(nlat: int, nlon: int, radius: float = 1.0) -> GaussGrid:
    return GaussGrid(nlat, nlon, radius)


def product_grid_dims(l_max: int, nlat: int, nlon: int) -> tuple[int, int]:
    """3/2-rule product grid (``num/spherical_backend.py:401-402``)."""
    return max((3 * l_max) // 2 + 1, nlat), max(3 * l_max + 1, nlon)


# Unit icosahedron vertex and face tables.  The table order fixes the point
# numbering of the subdivided grid, so it is the same standard icosahedron in
# the same order as num/geodesic_grid.py:15-28.
_PHI = (1.0 + math.sqrt(5.0)) / 2.0
_ICO_V = np.array([
    [-1, _PHI, 0], [1, _PHI, 0], [-1, -_PHI, 0], [1, -_PHI, 0],
    [0, -1, _PHI], [0, 1, _PHI], [0, -1, -_PHI], [0, 1, -_PHI],
    [_PHI, 0, -1], [_PHI, 0, 1], [-_PHI, 0, -1], [-_PHI, 0, 1],
], dtype=np.float64) / math.hypot(1.0, _PHI)
_ICO_F = [
    [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
    [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
    [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
    [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1],
]


def _rotation_x(angle: float) -> np.ndarray:
    """Rotation matrix about +x by ``angle`` (right-handed)."""
    c, s = math.cos(angle), math.sin(angle)
    return np.array([[1.0, 0.0, 0.0], [0.0, c, -s], [0.0, s, c]])


class GeodesicGrid:
    """Icosahedral geodesic grid (``num/geodesic_grid.py:31-266``)."""

    def __init__(self, resolution: int, radius: float = 1.0):
        self.resolution = int(resolution)
        self.radius = float(radius)
        verts = (_rotation_x(1e-2) @ _ICO_V.T).T          # :37-38
        vlist = [v for v in verts]
        faces = [list(f) for f in _ICO_F]
        for _ in range(self.resolution):                   # :110-134
            cache: dict[tuple[int, int], int] = {}
            new_faces = []

            def mid(a: int, b: int) -> int:
                key = (a, b) if a < b else (b, a)
                if key in cache:
                    return cache[key]
                p = (vlist[a] + vlist[b]) / 2.0
                vlist.append(p / np.linalg.norm(p))
                cache[key] = len(vlist) - 1
                return cache[key]

            for v1, v2, v3 in faces:
                a = mid(v1, v2)
                b = mid(v2, v3)
                c = mid(v3, v1)
                new_faces.extend([[v1, a, c], [v2, b, a], [v3, c, b], [a, b, c]])
            faces = new_faces
        self.unit_points = np.asarray(vlist, dtype=np.float64)
        self.points = self.radius * self.unit_points
        self.faces = np.asarray(faces, dtype=np.int64)
        self.n_points = self.points.shape[0]
        x, y, z = self.points[:, 0], self.points[:, 1], self.points[:, 2]
        xy = np.hypot(x, y)
        self.longitudes = np.arctan2(y, x)                 # cartesian_to_spherical.py:22
        self.latitudes = np.arctan2(z, xy)                 # :26
        self.point_latitudes = self.latitudes
        self.point_longitudes = self.longitudes
        self.coslat = np.cos(self.latitudes)
        self.sinlat = np.sin(self.latitudes)
        self.grid_shape = (self.n_points,)
        self._areas = None
        # Minimum edge length (num/geodesic_grid.py:159-185).
        f = self.faces
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        d = np.linalg.norm(self.points[e[:, 0]] - self.points[e[:, 1]], axis=1)
        self.min_edge_length = float(d[d > 0].min())

    @property
    def cell_areas(self) -> np.ndarray:
        """L'Huilier areas, one third per vertex (``num/geodesic_grid.py:212-266``)."""
        if self._areas is None:
            p = self.points / np.linalg.norm(self.points, axis=1)[:, None]
            f = self.faces
            a, b, c = p[f[:, 0]], p[f[:, 1]], p[f[:, 2]]
            sa = np.arccos(np.clip(np.sum(b * c, axis=1), -1.0, 1.0))
            sb = np.arccos(np.clip(np.sum(a * c, axis=1), -1.0, 1.0))
            sc = np.arccos(np.clip(np.sum(a * b, axis=1), -1.0, 1.0))
            s = (sa + sb + sc) / 2.0
            t = np.sqrt(np.tan(s / 2.0) * np.tan((s - sa) / 2.0)
                        * np.tan((s - sb) / 2.0) * np.tan((s - sc) / 2.0))
            tri = 4.0 * np.arctan(t) * self.radius ** 2
            areas = np.zeros(self.n_points)
            for k in range(3):
                np.add.at(areas, f[:, k], tri / 3.0)
            self._areas = areas
        return self._areas

    @property
    def voronoi_weights(self) -> np.ndarray:
        """``area / R^2`` renormalised to 4 pi (``num/optimized_geodesic_sh.py:127-133``)."""
        w = self.cell_areas / (self.radius ** 2)
        return w * (4.0 * np.pi) / np.sum(w)

    @property
    def cfl_length_scale(self) -> float:
        return self.min_edge_length


def geodesic(resolution: int, radius: float = 1.0) -> GeodesicGrid:
    return GeodesicGrid(resolution, radius)
