"""Fixed band-limited bottom topography (``phy/topography.py:175-290``).

``Topography.mountain`` evaluates ``h exp(-(d/sigma)^2)`` at the state points
on the device, analyses it once with the backend transform, applies the
optional ``l_cut`` and the 0.2 relative projection-residual gate.
"""
from __future__ import annotations

import math

import torch

__all__ = ["Topography", "TopographyError", "MAX_PROJECTION_RESIDUAL", "MAX_MOUNTAIN_HEIGHT_M"]

MAX_PROJECTION_RESIDUAL = 0.2
MAX_MOUNTAIN_HEIGHT_M = 1.0e5


class TopographyError(ValueError):
    """Invalid or unrepresentable topography."""


def _require(name, value, *, positive=False, lo=None, hi=None):
    try:
        v = float(value)
    except (TypeError, ValueError) as exc:
        raise TopographyError(f"{name} must be a number, got {value!r}") from exc
    if not math.isfinite(v):
        raise TopographyError(f"{name} must be finite, got {value!r}")
    if positive and not v > 0:
        raise TopographyError(f"{name} must be > 0, got {value!r}")
    if lo is not None and v < lo:
        raise TopographyError(f"{name} must be >= {lo}, got {value!r}")
    if hi is not None and v > hi:
        raise TopographyError(f"{name} must be <= {hi}, got {value!r}")
    return v


class Topography:
    def __init__(self, elevation_lm: torch.Tensor, *, preset: str = "custom",
                 parameters: dict | None = None):
        e = torch.as_tensor(elevation_lm).to(torch.complex128)
        if e.dim() != 2 or e.shape[0] != e.shape[1]:
            raise TopographyError("elevation coefficients must be (l_max+1, l_max+1)")
        self._elevation_lm = e
        self._preset = preset
        self._parameters = dict(parameters or {})

    @property
    def l_max(self) -> int:
        return self._elevation_lm.shape[0] - 1

    @property
    def elevation_lm(self) -> torch.Tensor:
        return self._elevation_lm.clone()

    @property
    def is_flat(self) -> bool:
        return not bool(torch.any(self._elevation_lm != 0))

    @property
    def preset(self) -> str:
        return self._preset

    @property
    def parameters(self) -> dict:
        return dict(self._parameters)

    def surface_geopotential_lm(self, gravity: float) -> torch.Tensor:
        g = _require("gravity", gravity, positive=True)
        return g * self._elevation_lm

    def elevation_on(self, sh) -> torch.Tensor:
        return sh.inv_transform(self._elevation_lm.to(sh.device))

    @classmethod
    def flat(cls, l_max: int, device=None) -> "Topography":
        if int(l_max) < 1:
            raise TopographyError(f"l_max must be >= 1, got {l_max}")
        n = int(l_max) + 1
        return cls(torch.zeros((n, n), dtype=torch.complex128, device=device), preset="flat")

    @classmethod
    def mountain(cls, planet, *, height_m: float, lat_deg: float, lon_deg: float,
                 width_deg: float, l_cut: int | None = None) -> "Topography":
        height = _require("mountain height_m", height_m, positive=True, hi=MAX_MOUNTAIN_HEIGHT_M)
        lat0 = _require("mountain lat_deg", lat_deg, lo=-90.0, hi=90.0)
        lon0 = _require("mountain lon_deg", lon_deg, lo=-360.0, hi=360.0)
        width = _require("mountain width_deg", width_deg, positive=True, hi=90.0)
        sh = planet.sh
        if l_cut is not None and not 1 <= int(l_cut) <= sh.l_max:
            raise TopographyError(f"l_cut must be in [1, l_max={sh.l_max}], got {l_cut}")
        lat = torch.as_tensor(planet.grid.point_latitudes, device=sh.device)
        lon = torch.as_tensor(planet.grid.point_longitudes, device=sh.device)
        la0, lo0, sig = math.radians(lat0), math.radians(lon0), math.radians(width)
        cosd = math.sin(la0) * torch.sin(lat) + math.cos(la0) * torch.cos(lat) * torch.cos(lon - lo0)
        d = torch.arccos(torch.clamp(cosd, -1.0, 1.0))
        analytic = height * torch.exp(-(d / sig) ** 2)
        coeffs = sh.transform(analytic)
        if not bool(torch.isfinite(coeffs).all()):
            raise TopographyError("mountain projection produced non-finite coefficients")
        if l_cut is not None:
            coeffs[int(l_cut) + 1:, :] = 0.0
            coeffs[:, int(l_cut) + 1:] = 0.0
        synth = sh.inv_transform(coeffs)
        w = torch.as_tensor(sh.weights, device=sh.device)
        norm = float(torch.sqrt(torch.sum(w * analytic ** 2)))
        resid = float(torch.sqrt(torch.sum(w * (synth - analytic) ** 2)))
        rel = resid / norm if norm > 0 else math.inf
        if not rel <= MAX_PROJECTION_RESIDUAL:
            raise TopographyError(
                f"mountain (height {height:g} m, width {width:g} deg) is not representable: "
                f"relative projection residual {rel:.3f} exceeds {MAX_PROJECTION_RESIDUAL}.")
        params = {"height_m": height, "lat_deg": lat0, "lon_deg": lon0, "width_deg": width,
                  "projection_residual": rel}
        if l_cut is not None:
            params["l_cut"] = int(l_cut)
        return cls(coeffs, preset="mountain", parameters=params)
