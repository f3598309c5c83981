"""Spherical-harmonic transform restatements (oracle, test infrastructure only).

* :class:`DenseTransform` restates ``PointSetSphericalHarmonics``
  (``num/fast_geodesic_sh.py:15-107``): analysis ``a = Y^H (w * f)`` scattered
  into the dense ``(L+1, L+1)`` layout with zeros for m > l (``:68-87``);
  synthesis ``f = 2 Re(Y a) - Re(Y a_{m=0})`` (``:89-107``); uniform default
  weights ``4 pi / N`` (``:35-37``).  ``basis="reference"`` builds ``Y`` with
  the reference formula (``cu/sh_matrix.cu``), ``basis="stable"`` with the
  stable recurrence.
* :class:`GaussTransform` computes the same sums on a Gauss grid by a
  longitude DFT (``numpy.fft``) and a per-m Legendre contraction with the
  stable recurrence; aliasing of zonal wavenumbers m >= nlon is folded
  exactly as the dense sum would see it.

Pinned by (tests/test_oracle_transforms.py): Gauss round trip to 1e-13,
single-mode delta to 1e-11, ``a00 = sqrt(4 pi)`` for f = 1 to 1e-13
(``tests/test_latlon_transform.py:99-121``); GaussTransform equals the dense
reference-formula transform to 1e-13 on grids up to L = 42; geodesic
round-trip envelope 5e-2 at res 4 / L = 21 (``tests/test_backend_parity.py:81``).
"""
from __future__ import annotations

import numpy as np

from .basis import lm_indices, n_basis, normalized_legendre, reference_basis, stable_basis

__all__ = ["DenseTransform", "GaussTransform", "scatter_dense", "gather_flat"]


def scatter_dense(flat: np.ndarray, l_max: int) -> np.ndarray:
    """Flat ``(..., B)`` -> dense ``(..., L+1, L+1)`` with m > l zero."""
    li, mi = lm_indices(l_max)
    out = np.zeros(flat.shape[:-1] + (l_max + 1, l_max + 1), dtype=np.complex128)
    out[..., li, mi] = flat
    return out


def gather_flat(coeffs: np.ndarray, l_max: int) -> np.ndarray:
    li, mi = lm_indices(l_max)
    return coeffs[..., li, mi]


class DenseTransform:
    """Dense point-set transform (``num/fast_geodesic_sh.py:8-111``)."""

    def __init__(self, lat, lon, l_max: int, weights=None, basis: str = "reference",
                 grid_shape=None):
        self.l_max = int(l_max)
        self.latitudes = np.asarray(lat, dtype=np.float64)
        self.longitudes = np.asarray(lon, dtype=np.float64)
        self.n_points = self.latitudes.shape[0]
        if weights is None:
            self.weights = np.full(self.n_points, 4.0 * np.pi / self.n_points)
        else:
            self.weights = np.asarray(weights, dtype=np.float64)
        if basis == "reference":
            self.Y = reference_basis(self.latitudes, self.longitudes, self.l_max)
        elif basis == "stable":
            self.Y = stable_basis(self.latitudes, self.longitudes, self.l_max)
        else:
            raise ValueError(f"unknown basis {basis!r}")
        _, self.m_idx = lm_indices(self.l_max)
        self.grid_shape = grid_shape

    def transform(self, values: np.ndarray) -> np.ndarray:
        f = np.asarray(values, dtype=np.float64).ravel()
        flat = self.Y.conj().T @ (f * self.weights)       # :80-81
        return scatter_dense(flat, self.l_max)

    def inv_transform(self, coeffs: np.ndarray) -> np.ndarray:
        flat = gather_flat(np.asarray(coeffs), self.l_max)  # :97
        term_sum = self.Y @ flat                          # :98
        term_m0 = self.Y @ (flat * (self.m_idx == 0))      # :99-101
        return 2.0 * term_sum.real - term_m0.real         # :102


class GaussTransform:
    """Gauss-grid transform as longitude DFT + per-m Legendre contraction.

    Produces the same sums as :class:`DenseTransform` on the tensor-product
    Gauss grid (same weights, same point order); only the summation order
    differs.
    """

    def __init__(self, grid, l_max: int):
        self.grid = grid
        self.l_max = int(l_max)
        self.nlat, self.nlon = grid.nlat, grid.nlon
        L = self.l_max
        P = normalized_legendre(grid.x, L)               # (B, nlat)
        self.P = P
        # Per-m blocks P_m[l - m, j].
        self.Pm = []
        for m in range(L + 1):
            ks = [l * (l + 1) // 2 + m for l in range(m, L + 1)]
            self.Pm.append(P[ks, :])
        self.wlat = grid.gl_weights * (2.0 * np.pi / grid.nlon)
        self.weights = grid.weights
        self.n_points = grid.n_points

    def transform(self, values: np.ndarray) -> np.ndarray:
        L = self.l_max
        f = np.asarray(values, dtype=np.float64).reshape(self.nlat, self.nlon)
        F = np.fft.fft(f, axis=1)                       # sum_k f e^{-2 pi i m k/n}
        out = np.zeros((L + 1, L + 1), dtype=np.complex128)
        for m in range(L + 1):
            Fm = F[:, m % self.nlon] * self.wlat
            out[m:, m] = self.Pm[m] @ Fm
        return out

    def inv_transform(self, coeffs: np.ndarray) -> np.ndarray:
        L = self.l_max
        a = np.asarray(coeffs, dtype=np.complex128)
        spec = np.zeros((self.nlat, self.nlon), dtype=np.complex128)
        for m in range(L + 1):
            col = a[m:, m]
            if m == 0:
                G = self.Pm[0].T @ col.real
                spec[:, 0] += G
            else:
                G = self.Pm[m].T @ col
                spec[:, m % self.nlon] += 2.0 * G
        f = np.fft.ifft(spec, axis=1) * self.nlon        # sum_m S_m e^{+2 pi i m k/n}
        return f.real.ravel()
