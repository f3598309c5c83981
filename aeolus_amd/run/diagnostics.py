"""Spectral invariants (``run/bve/diagnostics.py:52-124``).

The four scalar invariants use the device kernel with compensated sums; the
per-degree spectra are device tensor reductions.
"""
from __future__ import annotations

import math

import torch

from .. import _native as nat

__all__ = ["spectral_diagnostics", "grid_diagnostics", "mode_power"]


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
