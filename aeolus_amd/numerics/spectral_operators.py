"""Spectral differential operators on the dense ``(L+1, L+1)`` layout.

Mirrors ``SpectralOperators`` (``num/spectral_operators.py:53-658``); every
operator runs on the HIP device through ``aeo_spec_op`` / ``aeo_jacobian`` /
``aeo_velocity`` / ``aeo_gradient``.
"""
from __future__ import annotations

import torch

from .. import _native as nat
from .spherical_backend import make_backend

__all__ = ["SpectralOperators"]


class SpectralOperators:
    def __init__(self, sh, radius: float, grid, product_quadrature: str = "coarse",
                 backend=None):
        self.sh = sh
        self.R = float(radius)
        self.l_max = sh.l_max
        self.grid = grid
        self.device = sh.device
        self.backend = backend if backend is not None else make_backend(grid, sh)
        self.product_quadrature = product_quadrature
        self._product_space = self.backend.product_space(product_quadrature)
        self.product_grid = self._product_space.geometry
        self.product_sh = (self._product_space.sh
                           if self._product_space.geometry is not None else None)
        l = torch.arange(self.l_max + 1, dtype=torch.float64, device=self.device)
        self.lap_eigs = -l * (l + 1.0) / self.R ** 2

    # -- helpers -----------------------------------------------------------
    def _coeffs(self, c) -> torch.Tensor:
        t = torch.as_tensor(c, device=self.device).to(torch.complex128).contiguous()
        n = self.l_max + 1
        if t.shape[-2:] != (n, n):
            raise ValueError(f"coefficients must have trailing shape ({n}, {n})")
        return t

    def _op(self, op: int, coeffs) -> torch.Tensor:
        a = self._coeffs(coeffs)
        out = torch.empty_like(a)
        nf = a.numel() // ((self.l_max + 1) ** 2)
        nat.check(nat.lib().aeo_spec_op(op, self.l_max, self.R, a.data_ptr(), out.data_ptr(),
                                        nf, nat.stream()), "spec_op")
        return out

    # -- spectral operators (num/spectral_operators.py:213-334) ------------
    def laplacian_coeffs(self, coeffs):
        return self._op(nat.OP_LAP, coeffs)

    def inv_laplacian(self, coeffs):
        return self._op(nat.OP_INVLAP, coeffs)

    def d_lambda_coeffs(self, coeffs):
        return self._op(nat.OP_DLAMBDA, coeffs)

    def sin_theta_d_theta_coeffs(self, coeffs):
        return self._op(nat.OP_SINTH_DTH, coeffs)

    def adjoint_sin_theta_d_theta_coeffs(self, coeffs):
        return self._op(nat.OP_SINTH_DTH_T, coeffs)

    def _truncate_product(self, coeffs):
        return self._op(nat.OP_TRUNC23, coeffs)

    def laplacian_field(self, coeffs):
        return self.sh.inv_transform(self.laplacian_coeffs(coeffs))

    def d_lambda_field(self, coeffs):
        return self.sh.inv_transform(self.d_lambda_coeffs(coeffs))

    def sin_theta_d_theta_field(self, coeffs):
        return self.sh.inv_transform(self.sin_theta_d_theta_coeffs(coeffs))

    # -- Jacobian (num/spectral_operators.py:590-658) ----------------------
    def jacobian_pseudospectral(self, a_lm, b_lm, dealias: bool = True,
                                return_spectral: bool = False):
        ps = self._product_space
        sh_p = ps.sh
        a = self._coeffs(a_lm)
        b = self._coeffs(b_lm)
        if not (dealias or return_spectral):
            # Grid-valued legacy path: synthesise the four derivative fields on
            # the product sampling and form J there.
            a_lam = sh_p.inv_transform(self.d_lambda_coeffs(a))
            b_lam = sh_p.inv_transform(self.d_lambda_coeffs(b))
            a_s = sh_p.inv_transform(self.sin_theta_d_theta_coeffs(a)) / self.R
            b_s = sh_p.inv_transform(self.sin_theta_d_theta_coeffs(b)) / self.R
            J = (a_s * b_lam - a_lam * b_s) / ps.coslat ** 2
            if sh_p is self.sh:
                return J
            return self.sh.inv_transform(sh_p.transform(J))
        out = torch.empty_like(a)
        nat.check(nat.lib().aeo_jacobian(sh_p._plan, self.R, a.data_ptr(), b.data_ptr(),
                                         out.data_ptr(), 1 if dealias else 0, nat.stream()),
                  "jacobian_pseudospectral")
        if return_spectral:
            return out
        return self.sh.inv_transform(out)

    # -- state-grid velocity / gradient (num/spectral_operators.py:517-558)
    def velocity_from_streamfunction(self, psi_lm):
        psi = self._coeffs(psi_lm)
        u = torch.empty(self.sh.n_points, dtype=torch.float64, device=self.device)
        v = torch.empty_like(u)
        nat.check(nat.lib().aeo_velocity(self.sh._plan, self.R, psi.data_ptr(), u.data_ptr(),
                                         v.data_ptr(), nat.stream()), "velocity_from_streamfunction")
        return u, v

    def grad_from_scalar(self, q_lm):
        q = self._coeffs(q_lm)
        dx = torch.empty(self.sh.n_points, dtype=torch.float64, device=self.device)
        dy = torch.empty_like(dx)
        nat.check(nat.lib().aeo_gradient(self.sh._plan, self.R, q.data_ptr(), dx.data_ptr(),
                                         dy.data_ptr(), nat.stream()), "grad_from_scalar")
        return dx, dy
