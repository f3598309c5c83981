"""Geodesic-grid transform wrapper (``num/optimized_geodesic_sh.py:42-257``).

Weight modes: ``"voronoi"`` (default; cell areas / R^2 renormalised so the
weights sum to 4 pi, ``:127-133``), ``"uniform"`` (4 pi / N) or an explicit
array.  ``"optimize"`` (BVLS weight fitting with a pickle cache) is not
part of the time-step path and raises.
"""
from __future__ import annotations

import warnings

import numpy as np

from .point_set_sh import PointSetSphericalHarmonics

__all__ = ["GeodesicSphericalHarmonics", "MIN_POINTS_PER_BASIS"]

MIN_POINTS_PER_BASIS = 6.0


def _warn_if_underresolved(n_points: int, l_max: int) -> None:
    n_basis = (l_max + 1) * (l_max + 2) // 2
    ratio = n_points / n_basis
    if ratio < MIN_POINTS_PER_BASIS:
        l_safe = int((np.sqrt(1.0 + 8.0 * n_points / MIN_POINTS_PER_BASIS) - 3.0) / 2.0)
        warnings.warn(
            f"Geodesic SH is under-resolved: {n_points} points for l_max={l_max} "
            f"({n_basis} basis functions, {ratio:.1f} pts/basis < {MIN_POINTS_PER_BASIS:.0f}). "
            f"Use l_max <= {l_safe} at this resolution, or a finer grid.", stacklevel=3)


class GeodesicSphericalHarmonics(PointSetSphericalHarmonics):
    def __init__(self, grid, l_max=None, cache_dir=None, weights="voronoi", latitudes=None,
                 longitudes=None, max_fields: int = 16):
        if l_max is None:
            raise ValueError("l_max is required.")
        self.grid = grid
        if grid is None:
            if latitudes is None or longitudes is None:
                raise ValueError("Provide either grid or latitudes/longitudes.")
            _warn_if_underresolved(len(latitudes), l_max)
            if isinstance(weights, str) and weights in ("optimize", "voronoi"):
                raise ValueError("weights='optimize' or 'voronoi' requires a grid.")
            w = None if (weights is None or weights == "uniform") else weights
            super().__init__(latitudes, longitudes, l_max, weights=w, max_fields=max_fields)
            return
        if latitudes is not None or longitudes is not None:
            raise ValueError("Provide either grid or latitudes/longitudes, not both.")
        _warn_if_underresolved(grid.n_points, l_max)
        if isinstance(weights, str):
            mode = weights.lower()
            if mode == "voronoi":
                w = grid.cell_areas_np / (grid.radius ** 2)
                w = w * (4 * np.pi) / np.sum(w)
            elif mode == "uniform":
                w = None
            elif mode == "optimize":
                raise ValueError("weights='optimize' is not provided by aeolus_amd "
                                 "(no silent fallback); use 'voronoi' or an explicit array")
            else:
                raise ValueError(f"Unknown weights mode: {weights}")
        else:
            w = None if weights is None else np.asarray(weights, dtype=np.float64)
        super().__init__(grid._latitudes_np, grid._longitudes_np, l_max, weights=w,
                         max_fields=max_fields, device=grid.device)
