"""Icosahedral geodesic grid geometry (host-side initialisation).

Mirrors ``GeodesicGridGeometry`` (``num/geodesic_grid.py:11-266``) and
``cartesian_to_spherical`` (``num/cartesian_to_spherical.py:3-33``): the unit
icosahedron (same vertex and face tables, which fix the point numbering),
rotated 0.01 rad about x with SciPy's ``Rotation.from_rotvec`` (the same
library call), subdivided ``resolution`` times with midpoints numbered in
order of first use, ``atan2`` longitude/latitude, and L'Huilier spherical
triangle areas shared one third per vertex.  The subdivision is vectorised
per level but assigns exactly the reference's vertex numbers.
"""
from __future__ import annotations

import math

import numpy as np
import torch
from scipy.spatial.transform import Rotation

from .. import _native as nat

__all__ = ["GeodesicGridGeometry"]

_PHI = (1.0 + math.sqrt(5.0)) / 2.0
ICO_VERTICES = (np.array([
    [-1, _PHI, 0], [1, _PHI, 0], [-1, -_PHI, 0], [1, -_PHI, 0],
    [0, -1, _PHI], [0, 1, _PHI], [0, -1, -_PHI], [0, 1, -_PHI],
    [_PHI, 0, -1], [_PHI, 0, 1], [-_PHI, 0, -1], [-_PHI, 0, 1],
]) / np.hypot(1, _PHI))
ICO_FACES = np.array([
    [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
    [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
    [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
    [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1],
], dtype=np.int64)


def subdivide(vertices: np.ndarray, faces: np.ndarray, levels: int):
    """Midpoint subdivision; new vertices numbered in order of first use."""
    v = vertices
    f = faces
    for _ in range(levels):
        nv = v.shape[0]
        # edges in encounter order: per face (v1,v2), (v2,v3), (v3,v1)
        e = np.stack([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], axis=1).reshape(-1, 2)
        lo = np.minimum(e[:, 0], e[:, 1])
        hi = np.maximum(e[:, 0], e[:, 1])
        key = lo * nv + hi
        uniq, first, inv = np.unique(key, return_index=True, return_inverse=True)
        order = np.argsort(first, kind="stable")
        rank = np.empty_like(order)
        rank[order] = np.arange(order.size)
        new_ids = nv + rank[inv]
        a_lo = lo[first[order]]
        a_hi = hi[first[order]]
        mid = (v[a_lo] + v[a_hi]) / 2.0
        mid = mid / np.linalg.norm(mid, axis=-1)[..., None]
        v = np.concatenate([v, mid], axis=0)
        m = new_ids.reshape(-1, 3)           # m12, m23, m31 per face
        v1, v2, v3 = f[:, 0], f[:, 1], f[:, 2]
        a, b, c = m[:, 0], m[:, 1], m[:, 2]
        f = np.stack([np.stack([v1, a, c], 1), np.stack([v2, b, a], 1),
                      np.stack([v3, c, b], 1), np.stack([a, b, c], 1)], axis=1).reshape(-1, 3)
    return v, f


class GeodesicGridGeometry:
    """Geodesic grid with Voronoi (dual-mesh) cell areas."""

    def __init__(self, resolution: int = 0, radius: float = 1.0, device=None):
        self.resolution = int(resolution)
        self.radius = float(radius)
        self.device = device if device is not None else nat.default_device()
        verts = Rotation.from_rotvec(np.array([1e-2, 0, 0])).apply(ICO_VERTICES)
        mesh_v, mesh_f = subdivide(verts, ICO_FACES, self.resolution)
        self.points = self.radius * mesh_v
        self.faces = mesh_f.astype(np.int32)
        self.n_points = self.points.shape[0]
        self.n_faces = self.faces.shape[0]
        x, y, z = self.points[:, 0], self.points[:, 1], self.points[:, 2]
        xy = np.hypot(x, y)
        self._longitudes_np = np.arctan2(y, x)
        self._latitudes_np = np.arctan2(z, xy)
        t = lambda arr: torch.as_tensor(np.ascontiguousarray(arr), dtype=torch.float64,
                                        device=self.device)
        self.longitudes = t(self._longitudes_np)
        self.latitudes = t(self._latitudes_np)
        self.sinlat = torch.sin(self.latitudes)
        self.coslat = torch.cos(self.latitudes)
        self.grid_shape = (self.n_points,)
        f = self.faces
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        d = np.linalg.norm(self.points[e[:, 0]] - self.points[e[:, 1]], axis=1)
        self.min_edge_length = float(d[d > 0].min()) if np.any(d > 0) else 0.0
        self._cell_areas = None

    @property
    def point_latitudes(self) -> torch.Tensor:
        return self.latitudes

    @property
    def point_longitudes(self) -> torch.Tensor:
        return self.longitudes

    def points_latlon(self) -> np.ndarray:
        return np.column_stack([self._longitudes_np, self._latitudes_np])

    @property
    def cfl_length_scale(self) -> float:
        return self.min_edge_length

    @property
    def cell_areas_np(self) -> np.ndarray:
        """L'Huilier triangle areas, one third to each vertex (``num/geodesic_grid.py:212-266``)."""
        if self._cell_areas is None:
            p = self.points / np.linalg.norm(self.points, axis=1)[:, None]
            f = self.faces
            a, b, c = p[f[:, 0]], p[f[:, 1]], p[f[:, 2]]
            sa = np.arccos(np.clip(np.sum(b * c, axis=1), -1.0, 1.0))
            sb = np.arccos(np.clip(np.sum(a * c, axis=1), -1.0, 1.0))
            sc = np.arccos(np.clip(np.sum(a * b, axis=1), -1.0, 1.0))
            s = (sa + sb + sc) / 2.0
            t4 = np.sqrt(np.tan(s / 2.0) * np.tan((s - sa) / 2.0)
                         * np.tan((s - sb) / 2.0) * np.tan((s - sc) / 2.0))
            tri = 4.0 * np.arctan(t4) * (self.radius ** 2)
            areas = np.zeros(self.n_points, dtype=np.float64)
            # np.add.at accumulates in index order: face by face, corner by corner.
            np.add.at(areas, f.ravel(), np.repeat(tri / 3.0, 3))
            self._cell_areas = areas
        return self._cell_areas

    @property
    def cell_areas(self) -> torch.Tensor:
        return torch.as_tensor(self.cell_areas_np, device=self.device)
