"""Shallow-water run diagnostics (``run/swe/diagnostics.py:1-173``).

``SWEDiagnosticsRecorder`` writes the reference's ``diagnostics/timeseries.csv``
(same columns and definitions) after every accepted step.  The grid numbers
of a row come from one fused device reduction over the state sampling
(``aeo_grid_stats``: max wind, total-geopotential extrema, the compensated
energy quadrature, free-surface extrema) plus the product-sampling extrema;
mass and L2 norms are spectral.  One device-to-host read per row.
"""
from __future__ import annotations

import csv
import math
import pathlib

import torch

from .diagnostics import mode_power

__all__ = ["SWEDiagnosticsRecorder", "SWE_CSV_COLUMNS"]

SWE_CSV_COLUMNS = [
    "time_s",
    "dt_s",
    "step",
    "max_wind_ms",
    "max_char_speed_ms",
    "cfl",
    "phi_total_min",
    "phi_total_max",
    "total_mass",
    "total_energy",
    "h_min_m",
    "eta_min_m",
    "eta_max_m",
    "terrain_max_m",
    "zeta_l2",
    "delta_l2",
    "phi_l2",
]


class SWEDiagnosticsRecorder:
    """Append-only diagnostics writer (``run/swe/diagnostics.py:81-173``)."""

    def __init__(self, model, out_dir):
        self.model = model
        self.R = float(model.R)
        self.length_scale = getattr(model.grid, "cfl_length_scale", None)
        w = torch.as_tensor(model.sh.weights, dtype=torch.float64).reshape(-1)
        self._area_weights = (w * self.R ** 2).contiguous()
        self._terrain_max_m = float(model.surface_elevation_extrema[1])
        self.dir = pathlib.Path(out_dir) / "diagnostics"
        self.dir.mkdir(parents=True, exist_ok=True)
        self._csv_file = open(self.dir / "timeseries.csv", "w", newline="", encoding="utf-8")
        self._writer = csv.DictWriter(self._csv_file, fieldnames=SWE_CSV_COLUMNS)
        self._writer.writeheader()

    def record(self, t: float, state, dt: float, step: int) -> dict:
        model = self.model
        R = self.R
        _, _, _, gs, mm = model._stats(state, w=self._area_weights)
        c = state.coeffs
        l2 = torch.stack([(R ** 2 * mode_power(c[f]).sum()) for f in range(3)])
        host = torch.cat([gs[:6], mm, l2, c[2][0, 0].real.reshape(1).to(torch.float64)]) \
            .cpu().tolist()                             # the row's one device-to-host read
        max_wind, s_min, s_max, energy, eta_min, eta_max = host[0:6]
        phi_min = min(s_min, model.phi0 + host[6])
        phi_max = max(s_max, model.phi0 + host[7])
        max_char = max_wind + math.sqrt(max(phi_max, 0.0))
        cfl = (max_char * dt / self.length_scale
               if (self.length_scale and dt > 0) else float("nan"))
        total_mass = R ** 2 * (4.0 * math.pi * model.phi0 + math.sqrt(4.0 * math.pi) * host[11])
        row = {
            "time_s": t,
            "dt_s": dt,
            "step": step,
            "max_wind_ms": max_wind,
            "max_char_speed_ms": max_char,
            "cfl": cfl,
            "phi_total_min": phi_min,
            "phi_total_max": phi_max,
            "total_mass": total_mass,
            "total_energy": energy,
            "h_min_m": phi_min / model.gravity,
            "eta_min_m": eta_min,
            "eta_max_m": eta_max,
            "terrain_max_m": self._terrain_max_m,
            "zeta_l2": math.sqrt(host[8]),
            "delta_l2": math.sqrt(host[9]),
            "phi_l2": math.sqrt(host[10]),
        }
        self._writer.writerow(row)
        self._csv_file.flush()
        return row

    def close(self) -> None:
        if not self._csv_file.closed:
            self._csv_file.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
