"""Point-set spherical-harmonic transform on the HIP device.

Mirrors ``PointSetSphericalHarmonics`` (``num/fast_geodesic_sh.py:8-111``):
complex orthonormal basis with m >= 0, analysis ``a = Y^H (w f)`` into the
dense ``(L+1, L+1)`` layout (m > l written as 0), synthesis
``f = 2 Re(Y a) - Re(Y a_{m=0})``, uniform default weights ``4 pi / N``.
The dense basis is built on the device by the library's FP64 stable
recurrence when the plan is created.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from .. import _native as nat

__all__ = ["PointSetSphericalHarmonics"]


class _TransformBase:
    """Owns a library plan; shared by the Gauss and point-set transforms."""

    def __init__(self, handle: ctypes.c_void_p, l_max: int, n_points: int, device):
        self._plan = handle
        self.l_max = int(l_max)
        self.n_points = int(n_points)
        self.device = device
        self.n_basis = (self.l_max + 1) * (self.l_max + 2) // 2
        ls, ms = [], []
        for l in range(self.l_max + 1):
            for m in range(l + 1):
                ls.append(l)
                ms.append(m)
        self.l_indices = torch.tensor(ls, dtype=torch.int32, device=device)
        self.m_indices = torch.tensor(ms, dtype=torch.int32, device=device)

    @property
    def plan(self) -> int:
        return self._plan.value

    def __del__(self):
        h = getattr(self, "_plan", None)
        if h is not None and h.value:
            try:
                nat.lib().aeo_plan_destroy(h)
            except Exception:  # interpreter shutdown
                pass
            self._plan = ctypes.c_void_p()

    # -- transforms --------------------------------------------------------
    def _as_grid(self, values) -> tuple[torch.Tensor, bool]:
        v = torch.as_tensor(values, dtype=torch.float64, device=self.device)
        if v.numel() == self.n_points:
            return v.reshape(1, self.n_points).contiguous(), False
        if v.numel() % self.n_points:
            raise ValueError(f"field of {v.numel()} values does not match {self.n_points} points")
        return v.reshape(-1, self.n_points).contiguous(), True

    def transform(self, values) -> torch.Tensor:
        """Analysis: field(s) -> ``(L+1, L+1)`` (or ``(F, L+1, L+1)``) complex128."""
        f, batched = self._as_grid(values)
        n = self.l_max + 1
        out = torch.empty((f.shape[0], n, n), dtype=torch.complex128, device=self.device)
        nat.check(nat.lib().aeo_analysis(self._plan, f.data_ptr(), f.shape[0], out.data_ptr(),
                                         nat.stream()), "transform")
        return out if batched else out[0]

    def inv_transform(self, coeffs) -> torch.Tensor:
        """Synthesis: coefficients -> flat field(s) ``(N,)`` (``(F, N)`` for stacks)."""
        a = torch.as_tensor(coeffs, device=self.device).to(torch.complex128)
        n = self.l_max + 1
        if a.shape[-2:] != (n, n):
            raise ValueError(f"coefficients must have trailing shape ({n}, {n}), got {tuple(a.shape)}")
        batched = a.dim() == 3
        a = a.reshape(-1, n, n).contiguous()
        out = torch.empty((a.shape[0], self.n_points), dtype=torch.float64, device=self.device)
        nat.check(nat.lib().aeo_synthesis(self._plan, a.data_ptr(), a.shape[0], out.data_ptr(),
                                          nat.stream()), "inv_transform")
        return out if batched else out[0]

    def inverse_transform(self, coeffs):
        """Alias for inv_transform."""
        return self.inv_transform(coeffs)


class PointSetSphericalHarmonics(_TransformBase):
    """Dense-basis transform on any point set (``num/fast_geodesic_sh.py:8-111``)."""

    def __init__(self, latitudes, longitudes, l_max: int, weights=None, max_fields: int = 16,
                 device=None):
        device = device if device is not None else nat.default_device()
        lat = np.ascontiguousarray(_to_numpy(latitudes), dtype=np.float64)
        lon = np.ascontiguousarray(_to_numpy(longitudes), dtype=np.float64)
        if lat.shape != lon.shape:
            raise ValueError("Latitude and Longitude arrays must have the same shape.")
        if lat.ndim != 1:
            raise ValueError("Input lat/lon must be 1D arrays.")
        n = lat.shape[0]
        w = None if weights is None else np.ascontiguousarray(_to_numpy(weights), dtype=np.float64)
        coslat = np.ascontiguousarray(np.cos(lat))
        handle = ctypes.c_void_p()
        nat.check(nat.lib().aeo_plan_create_points(
            n, int(l_max), nat.host_ptr(lat), nat.host_ptr(lon),
            None if w is None else nat.host_ptr(w), nat.host_ptr(coslat), int(max_fields),
            ctypes.byref(handle)), "PointSetSphericalHarmonics")
        super().__init__(handle, l_max, n, device)
        self.latitudes = torch.as_tensor(lat, device=device)
        self.longitudes = torch.as_tensor(lon, device=device)
        self.grid_shape = self.latitudes.shape
        self.phi_grid_gpu = np.pi / 2.0 - self.latitudes
        self.lambda_grid_gpu = self.longitudes
        self.weights = (torch.full((n,), 4.0 * np.pi / n, dtype=torch.float64, device=device)
                        if w is None else torch.as_tensor(w, device=device))
        self.coslat_points = torch.as_tensor(coslat, device=device)
        self._Y = None

    @property
    def Y_matrix(self) -> torch.Tensor:
        """Dense basis ``(N, B)`` complex128 (``num/fast_geodesic_sh.py:48-56``)."""
        if self._Y is None:
            Y = torch.empty((self.n_points, self.n_basis), dtype=torch.complex128,
                            device=self.device)
            nat.check(nat.lib().aeo_plan_basis(self._plan, Y.data_ptr(), nat.stream()),
                      "Y_matrix")
            self._Y = Y
        return self._Y

    @property
    def Y_matrix_real(self) -> torch.Tensor:
        return self.Y_matrix.real

    def enable_split(self, nslices: int) -> None:
        """Build the INT8 digit tables of the dense basis for split-precision
        model stages (``set_precision`` of the models on this product space)."""
        nat.check(nat.lib().aeo_plan_enable_split(self._plan, int(nslices)), "enable_split")
        self.nslices = int(nslices)


def _to_numpy(x) -> np.ndarray:
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    return np.asarray(x)
