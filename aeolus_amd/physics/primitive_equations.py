"""Dry hydrostatic primitive equations on the sigma grid
(``phy/primitive_equations.py:1-753``), tendency on the HIP device.

State: one ``(3K+1, L+1, L+1)`` complex128 stack, rows
``[zeta_1..zeta_K, delta_1..delta_K, T_1..T_K, ln p_s]`` top to bottom.
:meth:`PrimitiveEquationsModel.tendency` is one ``aeo_pe_tendency`` call:
10K + 2 syntheses on the product sampling, one column kernel for the
vertical physics (continuity, sigma_dot, omega/p, centred vertical advection,
the thermodynamic and momentum products), 4K + 1 analyses, 2K degree-(L+1)
analyses of the half-metric residual vector (the weak-form curl/divergence)
and one spectral assembly kernel (2/3 truncations, spectral hydrostatic
geopotential, ``-lap(Phi + E)``, pinned l = 0 momentum rows).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch

from .. import _native as nat
from .sigma_coordinate import SigmaGrid

__all__ = ["R_DRY", "CP_DRY", "KAPPA_DRY", "PROGNOSTICS", "product_truncation_cut",
           "PrimitiveEquationsStateError", "PrimitiveEquationsState", "isothermal_rest_state",
           "PrimitiveEquationsModel"]

R_DRY = 287.04
CP_DRY = 1004.64
KAPPA_DRY = R_DRY / CP_DRY
PROGNOSTICS = ("zeta", "delta", "temperature", "ln_ps")


def product_truncation_cut(l_max: int) -> int:
    """2/3-rule truncation degree of analysed products (``:61-68``)."""
    return (2 * int(l_max)) // 3


class PrimitiveEquationsStateError(ValueError):
    """A primitive-equation state violated a hard physical/numerical rule."""


@dataclass
class PrimitiveEquationsState:
    """``(3 nlev + 1, L+1, L+1)`` complex stack (``:76-137``)."""

    coeffs: torch.Tensor

    def __post_init__(self) -> None:
        c = self.coeffs
        if c.dim() != 3 or (c.shape[0] - 1) % 3 != 0 or c.shape[0] < 4 or c.shape[1] != c.shape[2]:
            raise PrimitiveEquationsStateError(
                "primitive-equation coefficients must have shape (3*nlev+1, l_max+1, l_max+1) "
                f"with nlev >= 1, got {tuple(c.shape)}")

    @property
    def nlev(self) -> int:
        return (self.coeffs.shape[0] - 1) // 3

    @property
    def zeta(self) -> torch.Tensor:
        return self.coeffs[0:self.nlev]

    @property
    def delta(self) -> torch.Tensor:
        return self.coeffs[self.nlev:2 * self.nlev]

    @property
    def temperature(self) -> torch.Tensor:
        return self.coeffs[2 * self.nlev:3 * self.nlev]

    @property
    def ln_ps(self) -> torch.Tensor:
        return self.coeffs[3 * self.nlev]

    @classmethod
    def from_fields(cls, zeta_lm, delta_lm, temperature_lm, ln_ps_lm) -> "PrimitiveEquationsState":
        z, d, t = (torch.as_tensor(x).to(torch.complex128) for x in (zeta_lm, delta_lm,
                                                                     temperature_lm))
        q = torch.as_tensor(ln_ps_lm).to(torch.complex128).to(z.device)
        return cls(torch.cat([z, d.to(z.device), t.to(z.device), q[None]], dim=0))

    @classmethod
    def zeros(cls, l_max: int, nlev: int, device=None) -> "PrimitiveEquationsState":
        n = l_max + 1
        return cls(torch.zeros((3 * nlev + 1, n, n), dtype=torch.complex128, device=device))


def isothermal_rest_state(l_max: int, nlev: int, *, temperature: float, surface_pressure: float,
                          device=None) -> PrimitiveEquationsState:
    """zeta = delta = 0, T_k = temperature, p_s uniform: pure (0, 0) modes
    scaled by sqrt(4 pi) (``:139-161``)."""
    if not (math.isfinite(temperature) and temperature > 0):
        raise ValueError(f"temperature must be finite and > 0, got {temperature}")
    if not (math.isfinite(surface_pressure) and surface_pressure > 0):
        raise ValueError(f"surface_pressure must be finite and > 0, got {surface_pressure}")
    st = PrimitiveEquationsState.zeros(l_max, nlev, device=device)
    mono = math.sqrt(4.0 * math.pi)
    st.temperature[:, 0, 0] = temperature * mono
    st.ln_ps[0, 0] = math.log(surface_pressure) * mono
    return st


class PrimitiveEquationsModel:
    """Explicit, unsplit dry hydrostatic PE tendency (``:164-753``)."""

    def __init__(self, planet, sigma: SigmaGrid, *, r_dry: float = R_DRY, cp_dry: float = CP_DRY,
                 surface_geopotential_lm=None):
        if not (math.isfinite(r_dry) and r_dry > 0):
            raise ValueError(f"r_dry must be finite and > 0, got {r_dry}")
        if not (math.isfinite(cp_dry) and cp_dry > r_dry):
            raise ValueError(f"cp_dry must be finite and > r_dry, got {cp_dry}")
        if not isinstance(sigma, SigmaGrid):
            raise TypeError(f"sigma must be a SigmaGrid, got {type(sigma).__name__}")
        self.planet = planet
        self.sh = planet.sh
        self.so = planet.so
        self.grid = planet.grid
        self.device = self.sh.device
        self.R = float(planet.params.radius)
        self.Omega = float(planet.params.angular_velocity)
        self.sigma = sigma
        self.nlev = sigma.nlev
        self.r_dry = float(r_dry)
        self.cp_dry = float(cp_dry)
        self.kappa = self.r_dry / self.cp_dry
        self.gamma = self.cp_dry / (self.cp_dry - self.r_dry)
        self.l_max = self.sh.l_max
        n = self.l_max + 1
        self._phi_cut = product_truncation_cut(self.l_max)
        if surface_geopotential_lm is None:
            self.phi_surface_lm = torch.zeros((n, n), dtype=torch.complex128, device=self.device)
        else:
            phi_s = torch.as_tensor(surface_geopotential_lm, device=self.device).to(torch.complex128)
            if tuple(phi_s.shape) != (n, n):
                raise ValueError(f"surface_geopotential_lm must have shape {(n, n)}, got "
                                 f"{tuple(phi_s.shape)}")
            if not bool(torch.isfinite(phi_s).all()):
                raise ValueError("surface_geopotential_lm contains NaN/Inf")
            cut = self._phi_cut
            if bool((phi_s[cut + 1:, :] != 0).any()) or bool((phi_s[:, cut + 1:] != 0).any()):
                raise ValueError(
                    "surface_geopotential_lm has spectral content above the dealiased product "
                    f"truncation l = {cut} (2*l_max/3 for l_max = {self.l_max}); it could never "
                    "be balanced by the truncated pressure-gradient pathway. Band-limit the "
                    "terrain at the cut or raise l_max.")
            self.phi_surface_lm = phi_s.contiguous()
        l = torch.arange(n, dtype=torch.float64, device=self.device)
        self.lap_eigs = -l * (l + 1.0) / self.R ** 2
        self.f_lm = torch.zeros((n, n), dtype=torch.complex128, device=self.device)
        self.f_lm[1, 0] = 2.0 * self.Omega * math.sqrt(4.0 * math.pi / 3.0)
        self._ps = self.so.backend.product_space(self.so.product_quadrature)
        coslat = getattr(self.grid, "coslat", None)
        if coslat is None:
            coslat = torch.cos(torch.as_tensor(self.grid.point_latitudes))
        self._state_coslat = torch.clamp(torch.as_tensor(coslat, device=self.device), min=1e-8)
        # own product-sampling plans sized so that every batch of one tendency
        # is a single synthesis / analysis launch sequence
        self._prod = self.so.product_transform(self.l_max, max_fields=10 * self.nlev + 2)
        self._ext = self.so.product_transform(self.l_max + 1, max_fields=2 * self.nlev)
        interfaces = np.ascontiguousarray(sigma.interfaces_array())
        handle = ctypes.c_void_p()
        nat.check(nat.lib().aeo_pe_create(
            self._prod._plan, self._ext._plan, self.nlev, nat.host_ptr(interfaces), self.R,
            self.Omega, self.r_dry, self.cp_dry,
            None if surface_geopotential_lm is None else self.phi_surface_lm.data_ptr(),
            ctypes.byref(handle)), "PrimitiveEquationsModel")
        self._model = handle

    def __del__(self):
        h = getattr(self, "_model", None)
        if h is not None and h.value:
            try:
                nat.lib().aeo_pe_destroy(h)
            except Exception:
                pass
            self._model = ctypes.c_void_p()

    def _stack(self, coeffs) -> torch.Tensor:
        y = torch.as_tensor(coeffs, device=self.device).to(torch.complex128).contiguous()
        n = self.l_max + 1
        if tuple(y.shape) != (3 * self.nlev + 1, n, n):
            raise PrimitiveEquationsStateError(
                f"coefficients must have shape {(3 * self.nlev + 1, n, n)}, got {tuple(y.shape)}")
        return y

    def tendency(self, coeffs) -> torch.Tensor:
        """d/dt of the ``(3K+1, L+1, L+1)`` stack (``:727-753``)."""
        y = self._stack(coeffs)
        out = torch.empty_like(y)
        nat.check(nat.lib().aeo_pe_tendency(self._model, y.data_ptr(), out.data_ptr(),
                                            nat.stream()), "PrimitiveEquationsModel.tendency")
        return out

    def tendency_state(self, state: PrimitiveEquationsState) -> PrimitiveEquationsState:
        return PrimitiveEquationsState(self.tendency(state.coeffs))

    # -- state-grid fields (phy/primitive_equations.py:266-313, 472-480) ----
    def _inv_laplacian(self, coeffs) -> torch.Tensor:
        """lap^-1 per (L+1, L+1) slice with l = 0 zeroed (``:266-270``)."""
        c = coeffs.to(torch.complex128).contiguous()
        out = torch.empty_like(c)
        nat.check(nat.lib().aeo_spec_op(nat.OP_INVLAP_ZERO, self.l_max, self.R, c.data_ptr(),
                                        out.data_ptr(), c.numel() // (self.l_max + 1) ** 2,
                                        nat.stream()), "inv_laplacian")
        return out

    def wind_on_state_grid(self, state: PrimitiveEquationsState):
        """Per-level (u, v), each ``(nlev, N)``, on the state sampling: the
        Helmholtz wind ``k x grad psi + grad chi`` (``:280-307``)."""
        K = self.nlev
        y = self._stack(state.coeffs)
        pc = self._inv_laplacian(y[:2 * K])                     # psi_1..K, chi_1..K
        lam = self.sh.inv_transform(self.so.d_lambda_coeffs(pc))
        snt = self.sh.inv_transform(self.so.sin_theta_d_theta_coeffs(pc)) / self.R
        c = self._state_coslat
        return (snt[:K] + lam[K:]) / c, (lam[:K] - snt[K:]) / c

    def temperature_on_state_grid(self, state: PrimitiveEquationsState) -> torch.Tensor:
        return self.sh.inv_transform(self._stack(state.coeffs)[2 * self.nlev:3 * self.nlev])

    def surface_pressure_on_state_grid(self, state: PrimitiveEquationsState) -> torch.Tensor:
        """``exp(ln p_s)`` on the state grid (``:472-480``)."""
        return torch.exp(self.sh.inv_transform(self._stack(state.coeffs)[3 * self.nlev]))

    def _t_extrema_dev(self, y) -> torch.Tensor:
        """Device ``[min, max]`` of T over every level and every sampling."""
        from ..run.diagnostics import grid_stats
        t = y[2 * self.nlev:3 * self.nlev]
        samplings = [self.sh] + ([self._ps.sh] if self._ps.sh is not self.sh else [])
        mm = [grid_stats(nat.GRID_STATS_MINMAX, None, sh.inv_transform(t))[:2] for sh in samplings]
        lo = torch.stack([x[0] for x in mm]).min()
        hi = torch.stack([x[1] for x in mm]).max()
        return torch.stack([lo, hi])

    def temperature_extrema(self, state: PrimitiveEquationsState) -> tuple[float, float]:
        """(min, max) of T over every level on the state and product samplings
        (``:764-782``); one host read."""
        lo, hi = self._t_extrema_dev(self._stack(state.coeffs)).tolist()
        return lo, hi

    def max_characteristic_speed(self, state: PrimitiveEquationsState) -> float:
        """``max_k max|V_k| + sqrt(gamma R_d max(T_max, 0))`` (``:784-806``);
        one host read."""
        y = self._stack(state.coeffs)
        u, v = self.wind_on_state_grid(state)
        vals = torch.stack([torch.sqrt(u * u + v * v).max(), self._t_extrema_dev(y)[1]]).tolist()
        return vals[0] + math.sqrt(self.gamma * self.r_dry * max(vals[1], 0.0))

    #: monopole tolerance relative to the per-level spectral norm (``:808-811``)
    _MONOPOLE_RTOL = 1e-10

    def validate_state(self, state: PrimitiveEquationsState, *, context: str = "") -> None:
        """Raise :class:`PrimitiveEquationsStateError` on any hard violation
        (``:813-854``): shape, finiteness, per-level zeta / delta monopoles,
        strict temperature positivity over every sampling, finite p_s."""
        where = f" {context}" if context else ""
        n = self.l_max + 1
        expected = (3 * self.nlev + 1, n, n)
        c = state.coeffs
        if tuple(c.shape) != expected:
            raise PrimitiveEquationsStateError(
                f"state shape {tuple(c.shape)} does not match the model's expected "
                f"{expected}{where}")
        y = torch.as_tensor(c, device=self.device).to(torch.complex128).contiguous()
        K = self.nlev
        if not bool(torch.isfinite(y).all()):
            raise PrimitiveEquationsStateError(
                f"primitive-equation state contains NaN/Inf coefficients{where}")
        mono = y[:2 * K, 0, 0].abs()
        scale = torch.linalg.norm(y[:2 * K].reshape(2 * K, -1), dim=1)
        t_min = self._t_extrema_dev(y)[0]
        ps_ok = torch.isfinite(torch.exp(self.sh.inv_transform(y[3 * K]))).all()
        host = torch.cat([mono, scale, t_min[None], ps_ok[None].to(torch.float64)]).tolist()
        for idx in range(2 * K):
            name, k = ("zeta", idx) if idx < K else ("delta", idx - K)
            m, sc = host[idx], host[2 * K + idx]
            if m > self._MONOPOLE_RTOL * max(sc, 1e-30):
                raise PrimitiveEquationsStateError(
                    f"{name} monopole is nonzero at level {k + 1}{where}: |a00| = {m:g} "
                    f"(field norm {sc:g}); the global mean of {name} must stay exactly zero "
                    "on every level")
        if not host[4 * K] > 0.0:
            raise PrimitiveEquationsStateError(
                f"temperature is not strictly positive{where}: min(T) = {host[4 * K]:g} K over "
                "the state/product samplings")
        if host[4 * K + 1] != 1.0:
            raise PrimitiveEquationsStateError(
                f"surface pressure exp(ln p_s) overflowed to a non-finite value on the state "
                f"grid{where}")
