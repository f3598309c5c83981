"""Spectral operator restatements (oracle, test infrastructure only).

Follows ``num/spectral_operators.py``:

* ``lap_eigs = -l(l+1)/R^2`` (``:114``); ``i m / R`` with the m = 0 column
  zero (``:117-118, 220-226``);
* ``C+_{lm} = l sqrt(((l+1)^2 - m^2)/((2l+1)(2l+3)))``,
  ``C-_{lm} = -(l+1) sqrt((l^2 - m^2)/((2l-1)(2l+1)))`` with the maximum(., 0)
  and maximum(., 1) clamps, m > l zero, ``C+[L,:] = 0``, ``C-[0,:] = 0``
  (``:135-176``);
* ``S a`` (``:228-241``), ``S^T b`` (``:243-270``), inverse Laplacian with
  ``1/R^2`` at l = 0 (``:275-292``), 2/3 truncation (``:329-334``);
* pseudospectral Jacobian on the product space (``:590-658``);
* velocity from streamfunction with the 1e-6 NaN mask (``:517-538``) and the
  gradient with the unclamped ``coslat`` (``:542-558``).

Pinned by (tests/test_oracle_operators.py): ``S`` applied to ``Y_lm``
equals the analytic ``sin(theta) dY/dtheta`` truncated at L on a Gauss grid
(1e-12); ``S^T`` satisfies ``<S a, b> = <a, S^T b>`` exactly; solid-body
``J(psi_sb, Y_lm) = i m omega Y_lm`` (1e-8), ``J(a, a) = 0`` (1e-15),
``integral J = 0`` (1e-11), solid-body velocity (1e-8)
(``tests/test_backend_parity.py:106-140``).
"""
from __future__ import annotations

import numpy as np

from .grids import COSLAT_FLOOR

__all__ = ["SpectralOps"]


class SpectralOps:
    def __init__(self, l_max: int, radius: float):
        self.l_max = L = int(l_max)
        self.R = float(radius)
        l = np.arange(L + 1, dtype=np.float64)
        m = np.arange(L + 1, dtype=np.float64)
        self.lap_eigs = -l * (l + 1.0) / self.R ** 2
        im = 1j * m[None, :] / self.R
        im[0, 0] = 0.0
        self.im_m_over_R = im
        Lb, Mb = np.meshgrid(l, m, indexing="ij")
        valid = Mb <= Lb
        cp_ = Lb * np.sqrt(np.maximum((Lb + 1.0) ** 2 - Mb ** 2, 0.0)
                           / np.maximum((2.0 * Lb + 1.0) * (2.0 * Lb + 3.0), 1.0))
        cm_ = -(Lb + 1.0) * np.sqrt(np.maximum(Lb ** 2 - Mb ** 2, 0.0)
                                    / np.maximum((2.0 * Lb - 1.0) * (2.0 * Lb + 1.0), 1.0))
        cp_ = np.where(valid, cp_, 0.0)
        cm_ = np.where(valid, cm_, 0.0)
        cp_[L, :] = 0.0
        cm_[0, :] = 0.0
        self.C_plus = cp_
        self.C_minus = cm_
        self.cut = (2 * L) // 3

    # -- diagonal / recurrence operators ---------------------------------
    def laplacian(self, a):
        return self.lap_eigs[:, None] * a

    def inv_laplacian(self, a):
        inv = self.lap_eigs.copy()
        inv[0] = 1.0 / self.R ** 2
        return a / inv[:, None]

    def d_lambda(self, a):
        return self.im_m_over_R * a

    def sin_theta_d_theta(self, a):
        g = np.zeros(a.shape, dtype=np.complex128)
        g[1:, :] += self.C_plus[:-1, :] * a[:-1, :]
        g[:-1, :] += self.C_minus[1:, :] * a[1:, :]
        return g

    def sin_theta_d_theta_T(self, b):
        g = np.zeros(b.shape, dtype=np.complex128)
        g[:-1, :] += self.C_plus[:-1, :] * b[1:, :]
        g[1:, :] += self.C_minus[1:, :] * b[:-1, :]
        return g

    def truncate(self, a):
        a = np.array(a, dtype=np.complex128, copy=True)
        a[self.cut + 1:, :] = 0.0
        a[:, self.cut + 1:] = 0.0
        return a

    # -- grid-space operators --------------------------------------------
    def jacobian(self, a, b, prod_sh, prod_coslat, dealias=True):
        """Spectral J(a, b) on the product space (``:590-658``, return_spectral)."""
        a_lam = prod_sh.inv_transform(self.d_lambda(a))
        b_lam = prod_sh.inv_transform(self.d_lambda(b))
        a_s = prod_sh.inv_transform(self.sin_theta_d_theta(a)) / self.R
        b_s = prod_sh.inv_transform(self.sin_theta_d_theta(b)) / self.R
        J = (a_s * b_lam - a_lam * b_s) / prod_coslat ** 2
        J_lm = prod_sh.transform(J)
        return self.truncate(J_lm) if dealias else J_lm

    def velocity_from_streamfunction(self, psi, sh, lat):
        coslat = np.cos(lat)
        safe = np.where(np.abs(coslat) < 1e-6, np.nan, coslat)
        lam = sh.inv_transform(self.d_lambda(psi))
        g = sh.inv_transform(self.sin_theta_d_theta(psi))
        u = np.nan_to_num(g / (self.R * safe), nan=0.0)
        v = np.nan_to_num(lam / safe, nan=0.0)
        return u, v

    def grad_from_scalar(self, q, sh, coslat):
        lam = sh.inv_transform(self.d_lambda(q))
        g = sh.inv_transform(self.sin_theta_d_theta(q))
        return lam / coslat, -g / (self.R * coslat)


def clamp_coslat(lat):
    return np.maximum(np.cos(lat), COSLAT_FLOOR)
