"""BVE solver diagnostics (``run/bve/diagnostics.py:1-211``).

The four scalar invariants use the device kernel with compensated sums; the
per-degree spectra are device tensor reductions.  ``DiagnosticsRecorder``
writes the reference's ``diagnostics/timeseries.csv`` (same columns, one row
per accepted step) and ``diagnostics/spectra.npz``; every number of a row is
computed on the device (synthesis, velocity, the fused grid statistics
kernel) and read back in one transfer.
"""
from __future__ import annotations

import csv
import math
import pathlib

import numpy as np
import torch

from .. import _native as nat

__all__ = ["spectral_diagnostics", "grid_diagnostics", "mode_power", "DiagnosticsRecorder",
           "CSV_COLUMNS", "grid_stats"]

CSV_COLUMNS = [
    "time_s",
    "dt_s",
    "step",
    "max_speed_ms",
    "cfl",
    "circulation",
    "energy",
    "enstrophy_rel",
    "enstrophy_abs",
    "zeta_max",
    "zeta_rms",
    "energy_l1",
    "high_l_enstrophy_frac",
    "roundtrip_residual",
]


def grid_stats(mode: int, w, a, b=None, c=None, d=None, k=None, out=None) -> torch.Tensor:
    """Device grid statistics of one sampling (``aeo_grid_stats``); returns the
    device result tensor (6 doubles; the mode fixes how many are meaningful)."""
    ts = [None if t is None else t.reshape(-1).to(torch.float64).contiguous()
          for t in (w, a, b, c, d)]
    n = ts[1].numel()
    dev = ts[1].device
    kk = None if k is None else torch.as_tensor(k, dtype=torch.float64, device=dev).contiguous()
    out = torch.empty(6, dtype=torch.float64, device=dev) if out is None else out
    ptr = lambda t: None if t is None else t.data_ptr()
    nat.check(nat.lib().aeo_grid_stats(int(mode), n, *[ptr(t) for t in ts], ptr(kk),
                                       out.data_ptr(), nat.stream()), "grid_stats")
    return out


def mode_power(coeffs: torch.Tensor) -> torch.Tensor:
    n_l, n_m = coeffs.shape[-2:]
    li = torch.arange(n_l, device=coeffs.device)[:, None]
    mi = torch.arange(n_m, device=coeffs.device)[None, :]
    p = torch.where(mi == 0, coeffs.real ** 2, 2.0 * torch.abs(coeffs) ** 2)
    return torch.where(mi <= li, p, torch.zeros_like(p))


def spectral_diagnostics(zeta_lm: torch.Tensor, radius: float, omega: float) -> dict:
    z = zeta_lm.to(torch.complex128).contiguous()
    R = float(radius)
    L = z.shape[-1] - 1
    out = torch.empty(4, dtype=torch.float64, device=z.device)
    nat.check(nat.lib().aeo_spectral_diag(L, R, float(omega), z.data_ptr(), 1, out.data_ptr(),
                                          nat.stream()), "spectral_diagnostics")
    P = mode_power(z).sum(dim=1)
    l = torch.arange(L + 1, dtype=torch.float64, device=z.device)
    inv = torch.zeros_like(l)
    inv[1:] = 1.0 / (l[1:] * (l[1:] + 1.0))
    Z_l = 0.5 * R ** 2 * P
    E_l = 0.5 * R ** 4 * inv * P
    vals = out.cpu().tolist()
    cut = (2 * L) // 3
    Zt = vals[2]
    return {
        "circulation": vals[0],
        "energy": vals[1],
        "enstrophy_rel": vals[2],
        "enstrophy_abs": vals[3],
        "energy_l1": float(E_l[1]) if L >= 1 else 0.0,
        "high_l_enstrophy_frac": float(Z_l[cut + 1:].sum()) / Zt if Zt > 0 else 0.0,
        "energy_l": E_l.cpu().numpy(),
        "enstrophy_l": Z_l.cpu().numpy(),
    }


def grid_diagnostics(zeta_lm, sh, radius: float, omega: float, latitudes) -> dict:
    """Same invariants from the synthesised grid field (``run/bve/diagnostics.py:100-124``)."""
    R = float(radius)
    w_area = torch.as_tensor(sh.weights) * R ** 2
    zeta = sh.inv_transform(zeta_lm)
    f = 2.0 * omega * torch.sin(torch.as_tensor(latitudes, device=zeta.device))
    L = zeta_lm.shape[0] - 1
    l = torch.arange(L + 1, dtype=torch.float64, device=zeta.device)
    lam = -l * (l + 1.0) / R ** 2
    psi_lm = torch.zeros_like(zeta_lm)
    psi_lm[1:, :] = zeta_lm[1:, :] / lam[1:, None]
    psi = sh.inv_transform(psi_lm)
    return {
        "circulation": float(torch.sum(zeta * w_area)),
        "energy": float(-0.5 * torch.sum(psi * zeta * w_area)),
        "enstrophy_rel": float(0.5 * torch.sum(zeta ** 2 * w_area)),
        "enstrophy_abs": float(0.5 * torch.sum((zeta + f) ** 2 * w_area)),
    }


class DiagnosticsRecorder:
    """Append-only diagnostics writer, called after every accepted step
    (``run/bve/diagnostics.py:127-211``).

    Writes ``diagnostics/timeseries.csv`` row by row (flushed) and the degree
    spectra to ``diagnostics/spectra.npz`` on :meth:`close`.
    """

    def __init__(self, sh, so, grid, radius: float, omega: float, out_dir,
                 spectra_every: int = 10, roundtrip_every: int = 10):
        self.sh = sh
        self.so = so
        self.grid = grid
        self.R = float(radius)
        self.omega = float(omega)
        self.spectra_every = int(spectra_every)
        self.roundtrip_every = int(roundtrip_every)
        self.min_edge = getattr(grid, "cfl_length_scale", None)
        self.dir = pathlib.Path(out_dir) / "diagnostics"
        self.dir.mkdir(parents=True, exist_ok=True)
        self._csv_file = open(self.dir / "timeseries.csv", "w", newline="", encoding="utf-8")
        self._writer = csv.DictWriter(self._csv_file, fieldnames=CSV_COLUMNS)
        self._writer.writeheader()
        self._n_records = 0
        self._spectra_times: list[float] = []
        self._spectra_E: list[np.ndarray] = []
        self._spectra_Z: list[np.ndarray] = []
        self._w = torch.as_tensor(sh.weights, dtype=torch.float64).reshape(-1).contiguous()

    def record(self, t: float, zeta_lm, dt: float, step: int) -> dict:
        z = torch.as_tensor(zeta_lm).to(torch.complex128).contiguous()
        dev = z.device
        L = z.shape[-1] - 1
        R = self.R
        head = torch.empty(4, dtype=torch.float64, device=dev)
        nat.check(nat.lib().aeo_spectral_diag(L, R, self.omega, z.data_ptr(), 1, head.data_ptr(),
                                              nat.stream()), "DiagnosticsRecorder")
        P = mode_power(z).sum(dim=1)
        l = torch.arange(L + 1, dtype=torch.float64, device=dev)
        inv = torch.zeros_like(l)
        inv[1:] = 1.0 / (l[1:] * (l[1:] + 1.0))
        Z_l = 0.5 * R ** 2 * P
        E_l = 0.5 * R ** 4 * inv * P
        cut = (2 * L) // 3
        zeta = self.sh.inv_transform(z)
        u, v = self.so.velocity_from_streamfunction(self.so.inv_laplacian(z))
        gs = grid_stats(nat.GRID_STATS_BVE, self._w, zeta, u, v)
        do_rt = self.roundtrip_every > 0 and self._n_records % self.roundtrip_every == 0
        if do_rt:
            back = self.sh.transform(zeta)
            rt = torch.stack([torch.linalg.norm(back - z), torch.linalg.norm(z)])
        else:
            rt = torch.zeros(2, dtype=torch.float64, device=dev)
        do_spec = self.spectra_every > 0 and self._n_records % self.spectra_every == 0
        parts = [head, gs[:4], E_l[1:2] if L >= 1 else torch.zeros(1, dtype=torch.float64,
                                                                    device=dev),
                 Z_l[cut + 1:].sum().reshape(1), rt.to(torch.float64)]
        if do_spec:
            parts += [E_l, Z_l]
        host = torch.cat(parts).cpu().numpy()          # the step's one device-to-host read
        circ, energy, z_rel, z_abs = host[0:4]
        zeta_max, wz2, wsum, max_speed = host[4:8]
        e_l1, z_high = host[8], host[9]
        roundtrip = float(host[10]) / (float(host[11]) + 1e-300) if do_rt else float("nan")
        cfl = (max_speed * dt / self.min_edge) if (self.min_edge and dt > 0) else float("nan")
        row = {
            "time_s": t,
            "dt_s": dt,
            "step": step,
            "max_speed_ms": float(max_speed),
            "cfl": float(cfl),
            "circulation": float(circ),
            "energy": float(energy),
            "enstrophy_rel": float(z_rel),
            "enstrophy_abs": float(z_abs),
            "zeta_max": float(zeta_max),
            "zeta_rms": math.sqrt(wz2 / wsum),
            "energy_l1": float(e_l1),
            "high_l_enstrophy_frac": float(z_high) / float(z_rel) if z_rel > 0 else 0.0,
            "roundtrip_residual": roundtrip,
        }
        self._writer.writerow(row)
        self._csv_file.flush()
        if do_spec:
            self._spectra_times.append(t)
            self._spectra_E.append(host[12:13 + L].copy())
            self._spectra_Z.append(host[13 + L:14 + 2 * L].copy())
        self._n_records += 1
        return row

    def close(self) -> None:
        if self._spectra_times:
            np.savez(self.dir / "spectra.npz", times=np.asarray(self._spectra_times),
                     energy_l=np.asarray(self._spectra_E), enstrophy_l=np.asarray(self._spectra_Z))
        if not self._csv_file.closed:
            self._csv_file.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
