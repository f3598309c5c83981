"""Spherical-harmonic basis restatements (oracle, test infrastructure only).

Two independent constructions of the orthonormal complex basis
``Y_lm(lat, lon) = N_lm P_l^m(sin lat) e^{i m lon}`` (Condon-Shortley phase
included, ``N_lm = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!)``):

* :func:`reference_basis` restates the reference kernel
  ``cu/sh_matrix.cu:7-76`` literally: unnormalised ``P_m^m`` seed
  (``:33-40``), upward three-term recurrence (``:47-56``), normalisation by
  ``exp(lgamma)`` (``:59-60``), flat index ``l(l+1)/2 + m`` (``:68``).  Like
  the reference it is accurate to ~1e-13 up to L = 85 and breaks down above
  L ~ 120 (``exp(-lgamma(l+m+1))`` leaves the FP64 range).
* :func:`normalized_legendre` is the stable normalised recurrence used by
  the oracle at every L (valid through L = 170 and beyond).

Pinned by (tests/test_oracle_basis.py): closed forms ``Y00 = 1/sqrt(4 pi)``,
``Y10 = sqrt(3/4pi) sin lat``, ``Y11``, ``Y20``; 50-digit ``mpmath.spherharm``
spot values at L = 170 on the polar-most Gauss node; Gram matrix = I under
exact Gauss quadrature; agreement of the two constructions to 1e-13 for
L <= 85.
"""
from __future__ import annotations

import math

import numpy as np

__all__ = [
    "n_basis",
    "flat_index",
    "lm_indices",
    "reference_basis",
    "normalized_legendre",
    "stable_basis",
]


def n_basis(l_max: int) -> int:
    """Number of stored (l, m >= 0) functions, ``(L+1)(L+2)/2``."""
    return (l_max + 1) * (l_max + 2) // 2


def flat_index(l: int, m: int) -> int:
    """Flat basis index ``l(l+1)/2 + m`` (``cu/sh_matrix.cu:68``)."""
    return l * (l + 1) // 2 + m


def lm_indices(l_max: int) -> tuple[np.ndarray, np.ndarray]:
    """``(l_indices, m_indices)`` in flat order (``num/fast_geodesic_sh.py:58-66``)."""
    ls, ms = [], []
    for l in range(l_max + 1):
        for m in range(l + 1):
            ls.append(l)
            ms.append(m)
    return np.asarray(ls, dtype=np.int32), np.asarray(ms, dtype=np.int32)


def reference_basis(lat: np.ndarray, lon: np.ndarray, l_max: int) -> np.ndarray:
    """Dense ``Y (N, B)`` complex128 exactly as ``cu/sh_matrix.cu:7-76`` builds it.

    Vectorised over points; the per-point arithmetic follows the kernel:
    ``x = cos(pi/2 - lat)``, ``somx2 = sqrt((1-x)(1+x))``; the ``P_m^m`` seed
    multiplies ``-fact * somx2`` for ``fact = 1, 3, 5, ...``; the recurrence is
    ``P_l = ((2l-1) x P_{l-1} - (l+m-1) P_{l-2}) / (l-m)``; the normalisation is
    ``sqrt((2l+1)/(4 pi) * exp(lgamma(l-m+1) - lgamma(l+m+1)))``.
    """
    lat = np.asarray(lat, dtype=np.float64)
    lon = np.asarray(lon, dtype=np.float64)
    npts = lat.shape[0]
    B = n_basis(l_max)
    Y = np.zeros((npts, B), dtype=np.complex128)
    phi = math.pi / 2.0 - lat                       # colatitude (:25)
    x = np.cos(phi)                                 # (:26)
    somx2 = np.sqrt((1.0 - x) * (1.0 + x))          # (:27)
    for m in range(l_max + 1):
        pmm = np.ones(npts)
        fact = 1.0
        for _ in range(1, m + 1):                   # (:36-39)
            pmm = pmm * (-fact * somx2)
            fact += 2.0
        p_prev = pmm.copy()
        p_prev2 = np.zeros(npts)
        cosm = np.cos(m * lon)
        sinm = np.sin(m * lon)
        for l in range(m, l_max + 1):
            if l == m:
                p_curr = pmm
            elif l == m + 1:
                p_curr = x * (2.0 * m + 1.0) * pmm
            else:
                p_curr = ((2.0 * l - 1.0) * x * p_prev
                          - (l + m - 1.0) * p_prev2) / (l - m)
            log_fact_diff = math.lgamma(l - m + 1.0) - math.lgamma(l + m + 1.0)
            norm = math.sqrt((2.0 * l + 1.0) / (4.0 * math.pi)
                             * math.exp(log_fact_diff))
            k = flat_index(l, m)
            Y[:, k] = p_curr * norm * cosm + 1j * (p_curr * norm * sinm)
            p_prev2 = p_prev
            p_prev = p_curr
    return Y


def normalized_legendre(x: np.ndarray, l_max: int) -> np.ndarray:
    """Stable orthonormal associated Legendre values ``Pbar[k, i]``.

    ``Pbar_lm(x) = N_lm P_l^m(x)`` with the Condon-Shortley phase, so that
    ``Y_lm = Pbar_lm(sin lat) e^{i m lon}`` is orthonormal on the unit sphere.
    Flat index ``k = l(l+1)/2 + m``.  Recurrence (vectorised over m and x):

    * ``Pbar_00 = 1/sqrt(4 pi)``
    * ``Pbar_mm = -sqrt((2m+1)/(2m)) s Pbar_{m-1,m-1}``, ``s = sqrt(1-x^2)``
    * ``Pbar_{m+1,m} = sqrt(2m+3) x Pbar_mm``
    * ``Pbar_lm = a_lm (x Pbar_{l-1,m} - b_lm Pbar_{l-2,m})`` with
      ``a_lm = sqrt((4l^2-1)/(l^2-m^2))`` and
      ``b_lm = sqrt(((l-1)^2-m^2)/(4(l-1)^2-1))``.
    """
    x = np.asarray(x, dtype=np.float64)
    s = np.sqrt((1.0 - x) * (1.0 + x))
    npts = x.shape[0]
    out = np.zeros((n_basis(l_max), npts), dtype=np.float64)
    rows = [np.full((1, npts), 1.0 / math.sqrt(4.0 * math.pi))]
    out[0] = rows[0][0]
    for l in range(1, l_max + 1):
        cur = np.zeros((l + 1, npts))
        prev = rows[l - 1]
        cur[l] = -math.sqrt((2.0 * l + 1.0) / (2.0 * l)) * s * prev[l - 1]
        cur[l - 1] = math.sqrt(2.0 * l + 1.0) * x * prev[l - 1]
        if l >= 2:
            m = np.arange(l - 1, dtype=np.float64)[:, None]
            a = np.sqrt((4.0 * l * l - 1.0) / (l * l - m * m))
            b = np.sqrt(((l - 1.0) ** 2 - m * m) / (4.0 * (l - 1.0) ** 2 - 1.0))
            prev2 = rows[l - 2]
            cur[: l - 1] = a * (x[None, :] * prev[: l - 1] - b * prev2[: l - 1])
        rows.append(cur)
        base = l * (l + 1) // 2
        out[base: base + l + 1] = cur
        if l >= 2:
            rows[l - 2] = None  # release
    return out


def stable_basis(lat: np.ndarray, lon: np.ndarray, l_max: int) -> np.ndarray:
    """Dense ``Y (N, B)`` complex128 from :func:`normalized_legendre`."""
    lat = np.asarray(lat, dtype=np.float64)
    lon = np.asarray(lon, dtype=np.float64)
    P = normalized_legendre(np.sin(lat), l_max)        # (B, N)
    _, m_idx = lm_indices(l_max)
    phase = np.exp(1j * np.outer(lon, m_idx.astype(np.float64)))  # (N, B)
    return P.T * phase
