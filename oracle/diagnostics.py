"""Spectral invariants (oracle, test infrastructure only).

Follows ``run/bve/diagnostics.py:52-97``: per-mode power
``P = Re(a_l0)^2`` for m = 0 and ``2 |a_lm|^2`` for m > 0, padding (m > l)
ignored; ``E = 1/2 R^4 sum_l P_l/(l(l+1))``; ``Z = 1/2 R^2 sum_l P_l``;
absolute enstrophy with ``f`` added to the (1, 0) coefficient;
circulation ``R^2 sqrt(4 pi) Re a_00``.  Sums use ``math.fsum`` (exactly
rounded), the yardstick for the device's compensated sums.

Pinned by (tests/test_oracle_models.py): the closed forms of
``tests/test_diagnostics.py:63-108`` (single mode, circulation, absolute
enstrophy of f, padding ignored).
"""
from __future__ import annotations

import math

import numpy as np

__all__ = ["mode_power", "spectral_diagnostics"]


def mode_power(a: np.ndarray) -> np.ndarray:
    n_l, n_m = a.shape
    li, mi = np.indices((n_l, n_m))
    p = np.where(mi == 0, a.real ** 2, 2.0 * np.abs(a) ** 2)
    return np.where(mi <= li, p, 0.0)


def spectral_diagnostics(zeta: np.ndarray, radius: float, omega: float) -> dict:
    R = float(radius)
    L = zeta.shape[0] - 1
    P = mode_power(zeta)
    Pl = np.array([math.fsum(row) for row in P])
    l = np.arange(L + 1, dtype=np.float64)
    inv = np.zeros(L + 1)
    inv[1:] = 1.0 / (l[1:] * (l[1:] + 1.0))
    q = zeta.copy()
    q[1, 0] = q[1, 0] + 2.0 * omega * math.sqrt(4.0 * math.pi / 3.0)
    return {
        "circulation": R ** 2 * math.sqrt(4.0 * math.pi) * float(zeta[0, 0].real),
        "energy": 0.5 * R ** 4 * math.fsum(inv * Pl),
        "enstrophy_rel": 0.5 * R ** 2 * math.fsum(Pl),
        "enstrophy_abs": 0.5 * R ** 2 * math.fsum(mode_power(q).ravel()),
    }
