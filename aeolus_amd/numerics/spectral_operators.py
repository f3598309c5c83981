"""Spectral differential operators on the dense ``(L+1, L+1)`` layout.

Mirrors ``SpectralOperators`` (``num/spectral_operators.py:53-658``); every
operator runs on the HIP device through ``aeo_spec_op`` / ``aeo_jacobian`` /
``aeo_velocity`` / ``aeo_gradient`` / ``aeo_vector_curl_div``.
"""
from __future__ import annotations

import ctypes
import warnings

import torch

from .. import _native as nat
from .latlon_grid import GaussLatLonSphericalHarmonics
from .point_set_sh import PointSetSphericalHarmonics
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
        self._ext = None

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

    # -- weak-form vector analysis (num/spectral_operators.py:399-502) ------
    def product_transform(self, l_max: int, max_fields: int = 16):
        """A new transform of degree ``l_max`` on the product sampling (same
        points, same weights) that takes ``max_fields`` fields per launch."""
        sh_p = self._product_space.sh
        with warnings.catch_warnings():
            # a Gauss grid sized for degree L is under-resolved at L + 1; the
            # quadrature is the product sampling's either way
            warnings.simplefilter("ignore")
            if isinstance(sh_p, GaussLatLonSphericalHarmonics):
                return GaussLatLonSphericalHarmonics(sh_p.grid, l_max, max_fields=max_fields)
            return PointSetSphericalHarmonics(sh_p.latitudes, sh_p.longitudes, l_max,
                                              weights=sh_p.weights, max_fields=max_fields,
                                              device=self.device)

    def extended_product_transform(self):
        """The product-sampling transform one degree higher, built once
        (``num/spectral_operators.py:480-502``)."""
        if self._ext is None:
            self._ext = self.product_transform(self.l_max + 1)
        return self._ext

    def vector_curl_div_spectral(self, f_east, f_north, *, truncate: bool = True):
        """``(curl_lm, div_lm)`` of a tangent field sampled on the product
        points (``(N,)`` or ``(F, N)`` per component): degree-(L+1) analyses of
        the half-metric components, ``div = (i m/R) a + S^T_ext b / R``,
        ``curl = (i m/R) b - S^T_ext a / R``, 2/3-truncated when ``truncate``."""
        ext = self.extended_product_transform()
        fe = torch.as_tensor(f_east, dtype=torch.float64, device=self.device).contiguous()
        fn = torch.as_tensor(f_north, dtype=torch.float64, device=self.device).contiguous()
        N = ext.n_points
        if fe.shape != fn.shape or fe.numel() == 0 or fe.numel() % N:
            raise ValueError(f"f_east and f_north must both be (N,) or (F, N) with N = {N}, "
                             f"got {tuple(fe.shape)} and {tuple(fn.shape)}")
        nf = fe.numel() // N
        n = self.l_max + 1
        curl = torch.empty((nf, n, n), dtype=torch.complex128, device=self.device)
        div = torch.empty_like(curl)
        nw = ctypes.c_longlong()
        lib = nat.lib()
        nat.check(lib.aeo_vector_curl_div_work(ext._plan, nf, ctypes.byref(nw)),
                  "vector_curl_div_spectral")
        work = torch.empty(nw.value, dtype=torch.float64, device=self.device)
        nat.check(lib.aeo_vector_curl_div(ext._plan, self.R, fe.data_ptr(), fn.data_ptr(), nf,
                                          1 if truncate else 0, curl.data_ptr(), div.data_ptr(),
                                          work.data_ptr(), nat.stream()),
                  "vector_curl_div_spectral")
        if fe.dim() == 1:
            return curl[0], div[0]
        return curl, div

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
